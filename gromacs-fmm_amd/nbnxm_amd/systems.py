"""Synthetic, seeded molecular systems for the nonbonded cluster-pair kernel.

A System holds what nonbonded_verlet_t needs from the topology: coordinates,
charges, per-atom LJ type, the per-type-pair (C6, C12) matrix and the
exclusion lists (ListOfLists<int> as CSR ranges/elements).  The rigid
3-site water follows the model of the benchmark configurations: O-H 0.1 nm,
HOH 109.47 deg, q_O = -0.8476, q_H = +0.4238, LJ on O only
(sigma 0.3166 nm, epsilon 0.650 kJ/mol); intramolecular pairs excluded
(3x3 per molecule, self included, as src/gromacs/nbnxm/tests/kernel_test.cpp:265-270).
"""
from dataclasses import dataclass, field
import math

import numpy as np

O_H_BOND = 0.1
HOH_ANGLE_DEG = 109.47
Q_O = -0.8476
Q_H = 0.4238
SIGMA_O = 0.3166
EPSILON_O = 0.650
WATER_MOLECULES_PER_NM3 = 33.4
WATER_MASS_AMU = 18.01528
AMU_PER_NM3_AT_1G_CM3 = 602.214076


@dataclass
class System:
    box: np.ndarray                 # (3,) rectangular box edges, nm
    x: np.ndarray                   # (n, 3) float64
    q: np.ndarray                   # (n,) float64, e
    types: np.ndarray               # (n,) int32
    c6c12: np.ndarray               # (ntypes, ntypes, 2) float64: C6, C12
    excl_ranges: np.ndarray         # (n+1,) int32
    excl_elements: np.ndarray       # (m,) int32
    comb_rule: str = "none"         # "geom", "lb" or "none" (how c6c12 was built)
    per_type_sigma_eps: np.ndarray = field(default=None)  # (ntypes, 2), when known
    name: str = ""

    @property
    def n(self):
        return int(self.x.shape[0])

    @property
    def num_types(self):
        return int(self.c6c12.shape[0])


def lj_c6c12(sigma, eps):
    s6 = sigma ** 6
    return 4.0 * eps * s6, 4.0 * eps * s6 * s6


def combine_params(sig_eps, rule):
    """(ntypes, 2) sigma/epsilon -> (ntypes, ntypes, 2) C6/C12 by combination rule."""
    nt = sig_eps.shape[0]
    out = np.zeros((nt, nt, 2))
    for i in range(nt):
        for j in range(nt):
            si, ei = sig_eps[i]
            sj, ej = sig_eps[j]
            if ei == 0.0 or ej == 0.0:
                continue
            eps = math.sqrt(ei * ej)
            if rule == "geom":
                s6 = (si * sj) ** 3
            else:
                s6 = (0.5 * (si + sj)) ** 6
            out[i, j, 0] = 4.0 * eps * s6
            out[i, j, 1] = 4.0 * eps * s6 * s6
    return out


def random_rotations(rng, n):
    """Uniform random rotation matrices (Shoemake's quaternion method)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a = np.sqrt(1.0 - u1)
    b = np.sqrt(u1)
    qw = a * np.sin(2 * np.pi * u2)
    qx = a * np.cos(2 * np.pi * u2)
    qy = b * np.sin(2 * np.pi * u3)
    qz = b * np.cos(2 * np.pi * u3)
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (qy * qy + qz * qz)
    R[:, 0, 1] = 2 * (qx * qy - qz * qw)
    R[:, 0, 2] = 2 * (qx * qz + qy * qw)
    R[:, 1, 0] = 2 * (qx * qy + qz * qw)
    R[:, 1, 1] = 1 - 2 * (qx * qx + qz * qz)
    R[:, 1, 2] = 2 * (qy * qz - qx * qw)
    R[:, 2, 0] = 2 * (qx * qz - qy * qw)
    R[:, 2, 1] = 2 * (qy * qz + qx * qw)
    R[:, 2, 2] = 1 - 2 * (qx * qx + qy * qy)
    return R


def _water_geometry():
    half = math.radians(HOH_ANGLE_DEG) / 2.0
    return np.array([[0.0, 0.0, 0.0],
                     [O_H_BOND * math.sin(half), O_H_BOND * math.cos(half), 0.0],
                     [-O_H_BOND * math.sin(half), O_H_BOND * math.cos(half), 0.0]])


def water_exclusions(n_mol, offset=0):
    """3x3 intramolecular exclusions (self included) in CSR form."""
    n = 3 * n_mol
    ranges = np.arange(0, 3 * n + 1, 3, dtype=np.int32)
    first = (np.arange(n) // 3) * 3 + offset
    elements = (first[:, None] + np.arange(3)[None, :]).reshape(-1).astype(np.int32)
    return ranges, elements


def water_positions(n_mol, box_edge, rng, jitter=0.02):
    """O on a cubic lattice with uniform +-jitter, random orientations."""
    n_side = int(math.ceil(n_mol ** (1.0 / 3.0) - 1e-9))
    spacing = box_edge / n_side
    sites = np.stack(np.meshgrid(np.arange(n_side), np.arange(n_side), np.arange(n_side), indexing="ij"), -1).reshape(-1, 3)
    if sites.shape[0] > n_mol:
        pick = np.sort(rng.choice(sites.shape[0], size=n_mol, replace=False))
        sites = sites[pick]
    o = (sites + 0.5) * spacing + rng.uniform(-jitter, jitter, size=(n_mol, 3))
    rot = random_rotations(rng, n_mol)
    geom = _water_geometry()
    x = o[:, None, :] + np.einsum("mij,aj->mai", rot, geom)
    return x.reshape(-1, 3)


def water_box(n_mol, box_edge=None, seed=1, lj=True, jitter=0.02, name=""):
    """Rigid 3-site water box; box_edge defaults to 33.4 molecules/nm^3."""
    if box_edge is None:
        box_edge = (n_mol / WATER_MOLECULES_PER_NM3) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    x = water_positions(n_mol, box_edge, rng, jitter)
    q = np.tile([Q_O, Q_H, Q_H], n_mol).astype(np.float64)
    types = np.tile([0, 1, 1], n_mol).astype(np.int32)
    sig_eps = np.array([[SIGMA_O, EPSILON_O if lj else 0.0], [0.0, 0.0]])
    c6c12 = combine_params(sig_eps, "geom")
    ranges, elements = water_exclusions(n_mol)
    return System(box=np.array([box_edge] * 3), x=x, q=q, types=types, c6c12=c6c12,
                  excl_ranges=ranges, excl_elements=elements, comb_rule="geom",
                  per_type_sigma_eps=sig_eps, name=name or f"water{n_mol}")


def config_water_small(seed=1):
    """3,000 waters in a 4.48 nm cube (9,000 atoms)."""
    return water_box(3000, 4.48, seed=seed, name="water9k")


def config_water_96k(seed=2, lj=True):
    """32,000 waters in a 9.79 nm cube (96,000 atoms), with or without LJ."""
    return water_box(32000, 9.79, seed=seed, lj=lj, name="water96k" + ("" if lj else "_coulomb_only"))


def config_water_3m(seed=5):
    """1,000,000 waters in a 31.1 nm cube (3.0 M atoms)."""
    return water_box(1000000, 31.1, seed=seed, name="water3M")


def ensemble_sizes(num_systems=256, n_min=648, n_max=24000, seed=4):
    """Atom counts log-uniform in [n_min, n_max], rounded to whole molecules."""
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(math.log(n_min), math.log(n_max), size=num_systems))
    mols = np.clip(np.round(n / 3.0).astype(int), (n_min + 2) // 3, n_max // 3)
    return mols


def config_ensemble(num_systems=256, seed=4):
    """Water boxes at 33.4 molecules/nm^3 with log-uniform sizes."""
    mols = ensemble_sizes(num_systems, seed=seed)
    return [water_box(int(m), None, seed=seed * 100003 + k, name=f"ens{k}_{3 * int(m)}") for k, m in enumerate(mols)]


def spc27_system(golden, comb_rule):
    """The 243-atom box of the reference's kernel test (kernel_test.cpp:150-271)."""
    opt = golden["options"]
    x = np.array(golden["coordinates"], dtype=np.float64)
    n = x.shape[0]
    box = np.array([golden["box"]] * 3)
    sig_eps = np.array([[opt["sigma_O"], opt["epsilon_O"]],
                        [opt["sigma_H"], opt["epsilon_H"]],
                        [0.0, 0.0]])
    rule = "geom" if comb_rule == "geom" else "lb"
    # The test hands 4*eps*sigma^6 and 4*eps*sigma^12 to nbnxn_atomdata_t, which
    # takes its input as 6*C6 and 12*C12 (kernel_test.cpp:179-190 with
    # atomdata.cpp:457-466), so the effective C6/C12 carry 1/6 and 1/12.
    c6c12 = combine_params(sig_eps, rule)
    c6c12[:, :, 0] /= 6.0
    c6c12[:, :, 1] /= 12.0
    types = np.zeros(n, dtype=np.int32)
    q = np.zeros(n)
    for a in range(n):
        has_charge = a >= n // 3
        if a % 3 == 0:
            types[a] = 0
            q[a] = opt["charge_O"] if has_charge else 0.0
        else:
            types[a] = 1 if a >= n * 2 // 3 else 2
            q[a] = opt["charge_H"] if has_charge else 0.0
    ranges, elements = water_exclusions(n // 3)
    return System(box=box, x=x, q=q, types=types, c6c12=c6c12, excl_ranges=ranges,
                  excl_elements=elements, comb_rule=comb_rule, per_type_sigma_eps=sig_eps,
                  name=f"spc27_{comb_rule}")
