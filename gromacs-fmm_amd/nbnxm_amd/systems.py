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


SOLUTE_BOND = 0.15
SOLUTE_ATOM_MASS_AMU = 14.027  # united-atom CH2: the mass that sets the solvated box volume
SOLUTE_MIN_DISTANCE = 0.2      # self-avoidance of the random walk (non-bonded solute atoms)
WATER_SOLUTE_MIN_DISTANCE = 0.3


def _random_unit(rng):
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def _random_walk_chains(rng, n_chains, chain_len, box):
    """Self-avoiding-by-rejection random walks with fixed bond length in a
    periodic box: every new atom keeps SOLUTE_MIN_DISTANCE from all earlier
    solute atoms except its bonded predecessor."""
    cell = SOLUTE_MIN_DISTANCE
    ncell = np.maximum((box / cell).astype(int), 1)
    grid = {}
    placed = np.zeros((n_chains * chain_len, 3))

    def key(p):
        return tuple((np.floor(p / box * ncell).astype(int)) % ncell)

    def min_dist2(p, skip):
        k = key(p)
        best = np.inf
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    for a in grid.get(((k[0] + dx) % ncell[0], (k[1] + dy) % ncell[1], (k[2] + dz) % ncell[2]), ()):
                        if a == skip:
                            continue
                        d = p - placed[a]
                        d -= np.round(d / box) * box
                        best = min(best, float(d @ d))
        return best

    for c in range(n_chains):
        for m in range(chain_len):
            a = c * chain_len + m
            best_p, best_d = None, -1.0
            for _ in range(50):
                if m == 0:
                    p = rng.uniform(0.0, 1.0, 3) * box
                else:
                    p = placed[a - 1] + SOLUTE_BOND * _random_unit(rng)
                p = p - np.floor(p / box) * box
                d2 = min_dist2(p, a - 1 if m > 0 else -1)
                if d2 > best_d:
                    best_p, best_d = p, d2
                if d2 >= SOLUTE_MIN_DISTANCE ** 2:
                    break
            placed[a] = best_p
            grid.setdefault(key(best_p), []).append(a)
    return placed


def hetero_system(n_chains=480, chain_len=50, total_atoms=1_000_000, n_types=20, seed=6, name=""):
    """Solute of random-walk chains (bond 0.15 nm) with n_types LJ types
    (sigma U[0.25, 0.40] nm, epsilon U[0.1, 1.0] kJ/mol, Lorentz-Berthelot),
    charges N(0, 0.3 e) neutralised per chain and 1-2/1-3 exclusions,
    solvated with the rigid water to 1.0 g/cm^3 (solute atoms weigh
    SOLUTE_ATOM_MASS_AMU) up to total_atoms.  Types 0..n_types-1 are solute,
    n_types the water O and n_types+1 the water H."""
    rng = np.random.default_rng(seed)
    n_sol = n_chains * chain_len
    n_wat = (total_atoms - n_sol) // 3
    volume = n_wat * WATER_MASS_AMU / AMU_PER_NM3_AT_1G_CM3 + n_sol * SOLUTE_ATOM_MASS_AMU / AMU_PER_NM3_AT_1G_CM3
    edge = volume ** (1.0 / 3.0)
    box = np.array([edge] * 3)

    sig = rng.uniform(0.25, 0.40, n_types)
    eps = rng.uniform(0.1, 1.0, n_types)
    sig_eps = np.zeros((n_types + 2, 2))
    sig_eps[:n_types, 0] = sig
    sig_eps[:n_types, 1] = eps
    sig_eps[n_types] = (SIGMA_O, EPSILON_O)
    c6c12 = combine_params(sig_eps, "lb")

    xs = _random_walk_chains(rng, n_chains, chain_len, box)
    ts = rng.integers(0, n_types, n_sol).astype(np.int32)
    qs = rng.normal(0.0, 0.3, (n_chains, chain_len))
    qs -= qs.mean(axis=1, keepdims=True)
    qs = qs.reshape(-1)

    # Water: lattice candidates away from the solute, then n_wat of them at random.
    from scipy.spatial import cKDTree
    tree = cKDTree(xs, boxsize=edge)
    n_cand = int(n_wat * volume / (n_wat / WATER_MOLECULES_PER_NM3) * 1.05) + 8
    while True:
        xw = water_positions(n_cand, edge, rng).reshape(-1, 3, 3)
        xw -= np.floor(xw / box) * box
        d, _ = tree.query(xw[:, 0, :], k=1)
        free = np.flatnonzero(d >= WATER_SOLUTE_MIN_DISTANCE)
        if free.size >= n_wat:
            break
        n_cand = int(n_cand * 1.1) + 8
    keep = np.sort(rng.choice(free, size=n_wat, replace=False))
    xw = xw[keep].reshape(-1, 3)

    x = np.concatenate([xs, xw])
    q = np.concatenate([qs, np.tile([Q_O, Q_H, Q_H], n_wat)])
    types = np.concatenate([ts, np.tile(np.array([n_types, n_types + 1], dtype=np.int32)[[0, 1, 1]], n_wat)]).astype(np.int32)

    # Exclusions: self, 1-2 and 1-3 along each chain; 3x3 per water.
    elements = []
    ranges = [0]
    for a in range(n_sol):
        m = a % chain_len
        lo = a - min(m, 2)
        hi = a + min(chain_len - 1 - m, 2)
        elements.extend(range(lo, hi + 1))
        ranges.append(len(elements))
    wr, we = water_exclusions(n_wat, offset=n_sol)
    ranges = np.concatenate([np.array(ranges, dtype=np.int64), wr[1:].astype(np.int64) + len(elements)]).astype(np.int32)
    elements = np.concatenate([np.array(elements, dtype=np.int32), we]).astype(np.int32)
    return System(box=box, x=x, q=q, types=types, c6c12=c6c12, excl_ranges=ranges, excl_elements=elements,
                  comb_rule="lb", per_type_sigma_eps=sig_eps, name=name or f"hetero{x.shape[0]}")


def config_hetero_1m(seed=6):
    """480 chains x 50 atoms, 20 LJ types (LB), solvated to 1.0 M atoms."""
    return hetero_system(480, 50, 1_000_000, 20, seed=seed, name="hetero1M")
