"""Edge cases of the nonbonded kernel through the C ABI: an empty pair list, a
dilute system with no pair inside the cut-off, a lone water molecule whose
only pairs are excluded, and an LJ type table too large to stage in LDS.
The first three are checked against closed forms as well as the oracle."""
import dataclasses
import math

import numpy as np
import pytest

import nbnxm_amd as nb
import oracle
from conftest import announce
from helpers import rel_rms, wrapped_float_system

pytestmark = pytest.mark.gpu

RC, RLIST, RTOL = 1.0, 1.1, 1e-5


def run(system, vdw=nb.VDW_CUT_TYPE_TABLE, accumulation=nb.ACCUM_FIXED64, plist=None):
    ic = nb.interaction_const(RC, RTOL)
    search, built, atoms = nb.build(system, RLIST, vdw)
    plist = built if plist is None else plist
    task = nb.GpuTask(ic, atoms, RLIST, None, accumulation)
    task.init_pairlist(plist)
    task.launch(compute_energy=True, compute_virial=True)
    f, fd, elj, eel, fshift = task.fetch()
    forces = (fd if fd is not None else f).reshape(-1, 3)
    return ic, plist, atoms, forces, elj, eel, fshift.reshape(-1, 3)


def self_energy(ic, q):
    """Ewald self term of the list's diagonal entries: -epsfac beta/sqrt(pi) sum q^2."""
    return -ic.epsfac * ic.ewaldcoeff_q / math.sqrt(math.pi) * float(np.sum(np.asarray(q, dtype=np.float64) ** 2))


@pytest.fixture(scope="module")
def dilute_gas():
    """64 charged LJ atoms on a 1.5 nm lattice: every pair is beyond the cut-off."""
    rng = np.random.default_rng(17)
    g = np.stack(np.meshgrid(*[np.arange(4)] * 3, indexing="ij"), -1).reshape(-1, 3)
    x = (g + 0.5) * 1.5 + rng.uniform(-0.05, 0.05, size=g.shape)
    n = len(x)
    q = rng.choice([-0.5, 0.5], size=n)
    return nb.systems.System(box=np.array([6.0, 6.0, 6.0]), x=x, q=q, types=np.zeros(n, dtype=np.int32),
                             c6c12=np.array([[nb.systems.lj_c6c12(0.32, 0.8)]]),
                             excl_ranges=np.arange(n + 1, dtype=np.int32), excl_elements=np.arange(n, dtype=np.int32))


@pytest.mark.parametrize("accumulation", [nb.ACCUM_FIXED64, nb.ACCUM_FLOAT, nb.ACCUM_FP64])
def test_no_pair_inside_cutoff(dilute_gas, accumulation):
    """Only the self term remains: zero forces, zero LJ energy, Coulomb energy
    -epsfac beta/sqrt(pi) sum q^2 (nbnxm_cuda_kernel.cuh:381-418)."""
    announce(f"edge: dilute gas accumulation={accumulation}")
    ic, plist, atoms, f, elj, eel, fshift = run(dilute_gas, accumulation=accumulation)
    assert len(plist.sci) > 0
    assert np.all(f == 0.0)
    assert np.all(fshift == 0.0)
    assert elj == 0.0
    expect = self_energy(ic, dilute_gas.q)
    assert eel == pytest.approx(expect, rel=1e-6)
    _, _, elj64, eel64 = oracle.kernel(plist, atoms, ic, True, "f64")
    assert eel == pytest.approx(eel64, rel=1e-6)


def test_empty_pair_list(dilute_gas):
    """A list with no entries: no self term either, everything zero."""
    announce("edge: empty pair list")
    _, built, _ = nb.build(dilute_gas, RLIST)
    empty = nb.core.Pairlist(built.sci[:0], built.cj_packed[:0], built.excl[:1], RLIST)
    _, _, _, f, elj, eel, fshift = run(dilute_gas, plist=empty)
    assert np.all(f == 0.0) and np.all(fshift == 0.0)
    assert elj == 0.0 and eel == 0.0


def lone_water():
    w = nb.systems.water_box(1, box_edge=3.5, seed=23)
    x = w.x - w.x.mean(axis=0) + 1.75
    return dataclasses.replace(w, x=x)


def excluded_pairs_closed_form(system, ic):
    """Energy and forces of the excluded intramolecular pairs (int_bit = 0):
    V = -qq erf(beta r)/r, plus the self term."""
    beta, epsfac = ic.ewaldcoeff_q, ic.epsfac
    n = system.n
    f = np.zeros((n, 3))
    e = self_energy(ic, system.q)
    for i in range(n):
        for j in range(i + 1, n):
            d = system.x[i] - system.x[j]
            r = float(np.linalg.norm(d))
            qq = epsfac * system.q[i] * system.q[j]
            e -= qq * math.erf(beta * r) / r
            # F_i = -dV/dr r_hat, dV/dr = qq (erf(beta r)/r^2 - 2 beta/sqrt(pi) exp(-beta^2 r^2)/r)
            dvdr = qq * (math.erf(beta * r) / r ** 2 - 2.0 * beta / math.sqrt(math.pi) * math.exp(-(beta * r) ** 2) / r)
            f[i] -= dvdr * d / r
            f[j] += dvdr * d / r
    return e, f


@pytest.mark.parametrize("accumulation", [nb.ACCUM_FIXED64, nb.ACCUM_FP64])
def test_lone_molecule_only_excluded_pairs(accumulation):
    """One water in a 3.5 nm box: its three pairs are excluded, so the forces
    are the Ewald exclusion correction alone; closed form with math.erf."""
    announce(f"edge: lone water accumulation={accumulation}")
    w = lone_water()
    ic, plist, atoms, f, elj, eel, _ = run(w, accumulation=accumulation)
    # The closed form on the FP32 coordinates and charges the kernel reads.
    e_ref, f_ref = excluded_pairs_closed_form(wrapped_float_system(w), ic)
    fo = atoms.to_original_order(f, w.n)
    tol = 1e-9 if accumulation == nb.ACCUM_FP64 else 2e-5
    assert elj == 0.0
    assert eel == pytest.approx(e_ref, rel=tol)
    assert np.abs(fo - f_ref).max() <= tol * np.abs(f_ref).max()


def many_type_water(n_types=70, seed=29):
    """3000-atom water whose O atoms draw from n_types - 1 LB types (H keeps
    an LJ-free type): n_types + 1 table entries with the dummy type, beyond
    the 65 the kernel stages in LDS."""
    w = nb.systems.water_box(1000, None, seed=seed)
    rng = np.random.default_rng(seed)
    sig_eps = np.column_stack([rng.uniform(0.28, 0.36, n_types), rng.uniform(0.3, 1.0, n_types)])
    sig_eps[-1] = (0.0, 0.0)
    types = np.full(w.n, n_types - 1, dtype=np.int32)
    types[0::3] = rng.integers(0, n_types - 1, size=w.n // 3)
    return dataclasses.replace(w, types=types, c6c12=nb.systems.combine_params(sig_eps, "lb"), comb_rule="lb",
                               per_type_sigma_eps=sig_eps)


def test_type_table_beyond_lds():
    """71 type-table entries: the launch takes the kernel that reads the table
    from global memory; FP64 oracle on the same list."""
    announce("edge: 70-type table")
    s = many_type_water()
    assert s.num_types + 1 > 65
    ic, plist, atoms, f, elj, eel, fshift = run(s)
    f64, fs64, elj64, eel64 = oracle.kernel(plist, atoms, ic, True, "f64")
    real = atoms.atom_indices >= 0
    assert rel_rms(f[real], f64[real]) < 5e-6
    assert elj == pytest.approx(elj64, rel=2e-5)
    assert eel == pytest.approx(eel64, rel=2e-5)
    assert np.abs(fshift - fs64).max() <= 1e-4 * max(1.0, np.abs(fs64).max())


def test_dilute_gas_md_device_search(dilute_gas):
    """The nonbonded part of MD steps on the dilute gas with the device search
    (searches every 2 steps; every search after the first stays on the device,
    whose entry lists the task derives on the device): zero forces and LJ
    energy, the self Coulomb energy, and the same bits as the host search."""
    announce("edge: dilute gas md steps, device vs host search")
    ic = nb.interaction_const_ex(RC, ewald_rtol=RTOL)
    params = np.asarray(dilute_gas.c6c12, dtype=np.float64) * np.array([6.0, 12.0])
    runs = {}
    for gpu_search in (False, True):
        ver = nb.NonbondedVerlet(dilute_gas.types, params, dilute_gas.q, ic, nstlist=2, rlist_outer=RLIST,
                                 excl_ranges=dilute_gas.excl_ranges, excl_elements=dilute_gas.excl_elements,
                                 gpu_search=gpu_search)
        d_x = nb.DeviceArray(dilute_gas.x.astype(np.float32))
        d_f = nb.DeviceArray.zeros(dilute_gas.x.shape)
        out = []
        for step in range(6):
            ver.step(step, dilute_gas.box, d_x.ptr, d_f.ptr, 0, compute_energy=True, compute_virial=True)
            out.append((d_f.download(), ver.energies(), ver.virial()))
        assert ver.stats()["searches"] == 3
        runs[gpu_search] = out
    expect = self_energy(ic, dilute_gas.q)
    for step, (host, dev) in enumerate(zip(runs[False], runs[True])):
        assert np.all(dev[0] == 0.0), step
        assert dev[1][0] == 0.0, step
        assert dev[1][1] == pytest.approx(expect, rel=1e-6), step
        assert np.array_equal(host[0], dev[0]) and host[1] == dev[1], step
        assert np.array_equal(np.asarray(host[2]), np.asarray(dev[2])), step
