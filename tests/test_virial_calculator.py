"""Virial oracle, combination-rule detection, the heterogeneous chain system and
the force calculator's argument checks (CPU only: no call here reaches a GPU)."""
import ctypes
import math

import numpy as np
import pytest

import nbnxm_amd as nb
import oracle
from helpers import rel_rms
from nbnxm_amd import _lib


def test_pair_virial_two_atoms_closed_form():
    """-1/2 r (x) F for one charged LJ pair along x, from the closed-form force."""
    ic = nb.interaction_const(1.0, 1e-5)
    beta, epsfac = ic.ewaldcoeff_q, ic.epsfac
    c6, c12 = 1e-3, 1e-6
    r = 0.5
    s = nb.systems.System(box=np.array([3.0, 3.0, 3.0]), x=np.array([[0.5, 0.5, 0.5], [0.5 + r, 0.5, 0.5]]),
                          q=np.array([1.0, -1.0]), types=np.array([0, 0], dtype=np.int32),
                          c6c12=np.array([[[c6, c12]]]), excl_ranges=np.array([0, 1, 2], dtype=np.int32),
                          excl_elements=np.array([0, 1], dtype=np.int32))
    _, vir = oracle.bruteforce_virial(s, ic)
    fr = (12 * c12 * r ** -12 - 6 * c6 * r ** -6) / r ** 2 \
        - epsfac * (math.erfc(beta * r) / r ** 3 + 2 * beta / math.sqrt(math.pi) * math.exp(-(beta * r) ** 2) / r ** 2)
    expect = np.zeros((3, 3))
    expect[0, 0] = -0.5 * r * fr * r
    assert vir[0, 0] == pytest.approx(expect[0, 0], rel=1e-12)
    assert np.abs(vir - expect).max() <= 1e-12 * abs(expect[0, 0])


@pytest.mark.parametrize("n_mol", [120, 400])
def test_single_sum_virial_equals_pair_virial(n_mol):
    """The single-sum virial over list forces and shift forces (calc_vir) equals
    the pair virial of the brute-force sum: pins the shift-force convention."""
    s = nb.systems.water_box(n_mol, None, seed=5)
    ic = nb.interaction_const(1.0, 1e-5)
    search, plist, atoms = nb.build(s, 1.1, nb.VDW_CUT_COMB_GEOM)
    f, fs, _, _ = oracle.kernel(plist, atoms, ic, False, "f64")
    xq = atoms.xq.reshape(-1, 4)[:, :3].astype(np.float64)
    real = atoms.atom_indices >= 0
    vir_ss = oracle.single_sum_virial(xq[real], f[real], atoms.shift_vec, fs)
    # Brute force on the same (FP32-rounded, wrapped) coordinates and parameters.
    from helpers import wrapped_float_system
    ws = wrapped_float_system(s)
    _, vir_pair = oracle.bruteforce_virial(ws, ic)
    assert np.abs(vir_ss - vir_pair).max() < 1e-6 * np.abs(vir_pair).max()


def test_detect_comb_rule():
    rng = np.random.default_rng(1)
    sig_eps = np.column_stack([rng.uniform(0.25, 0.4, 5), rng.uniform(0.1, 1.0, 5)])

    def table(c6c12):
        return np.stack([6.0 * c6c12[..., 0], 12.0 * c6c12[..., 1]], axis=-1)

    geom = nb.systems.combine_params(sig_eps, "geom")
    lb = nb.systems.combine_params(sig_eps, "lb")
    assert nb.detect_comb_rule(table(geom)) == nb.VDW_CUT_COMB_GEOM
    assert nb.detect_comb_rule(table(lb)) == nb.VDW_CUT_COMB_LB
    other = lb.copy()
    other[1, 2] *= 1.5
    other[2, 1] *= 1.5
    assert nb.detect_comb_rule(table(other)) == nb.VDW_CUT_TYPE_TABLE
    # Water: LJ on O only; the zero H row satisfies both rules, geometric wins
    # (atomdata.cpp:518-531).
    w = nb.systems.water_box(10)
    assert nb.detect_comb_rule(table(w.c6c12)) == nb.VDW_CUT_COMB_GEOM
    # One type with only dispersion rules out LB.
    d = lb.copy()
    d[0, 0, 1] = 0.0
    assert nb.detect_comb_rule(table(d)) == nb.VDW_CUT_TYPE_TABLE
    assert nb.load().nbnxm_detect_comb_rule(0, None, 1e-5) == -1


def test_hetero_system_generator():
    s = nb.systems.hetero_system(6, 50, 2400, 20, seed=11)
    n_sol = 300
    assert s.n == n_sol + 3 * ((2400 - n_sol) // 3)
    # Neutral chains, waters with the model charges.
    assert np.abs(s.q[:n_sol].reshape(6, 50).sum(axis=1)).max() < 1e-12
    assert s.q[n_sol:n_sol + 3].tolist() == [nb.systems.Q_O, nb.systems.Q_H, nb.systems.Q_H]
    # Bonds of 0.15 nm along each chain (minimum image).
    d = np.diff(s.x[:n_sol].reshape(6, 50, 3), axis=1)
    d -= np.round(d / s.box) * s.box
    assert np.allclose(np.linalg.norm(d, axis=-1), 0.15)
    # Exclusions: self, 1-2 and 1-3 inside the chain; 3x3 per water.
    r, e = s.excl_ranges, s.excl_elements
    assert sorted(e[r[0]:r[1]].tolist()) == [0, 1, 2]
    assert sorted(e[r[25]:r[26]].tolist()) == [23, 24, 25, 26, 27]
    assert sorted(e[r[49]:r[50]].tolist()) == [47, 48, 49]
    assert sorted(e[r[n_sol + 1]:r[n_sol + 2]].tolist()) == [n_sol, n_sol + 1, n_sol + 2]
    c = np.stack([6.0 * s.c6c12[..., 0], 12.0 * s.c6c12[..., 1]], axis=-1)
    assert nb.detect_comb_rule(c) == nb.VDW_CUT_COMB_LB
    assert s.num_types == 22


@pytest.mark.parametrize("vdw", [nb.VDW_CUT_TYPE_TABLE, nb.VDW_CUT_COMB_LB])
def test_hetero_list_kernel_equals_bruteforce(vdw):
    """The FP64 list oracle on the chain system (20 LB types, 1-2/1-3
    exclusions) against the all-image brute force."""
    from helpers import wrapped_float_system
    s = nb.systems.hetero_system(6, 50, 2400, 20, seed=11)
    ic = nb.interaction_const(1.0, 1e-5)
    search, plist, atoms = nb.build(s, 1.1, vdw)
    f, fs, elj, eel = oracle.kernel(plist, atoms, ic, True, "f64")
    fo = atoms.to_original_order(f, s.n)
    ws = wrapped_float_system(s)
    if vdw == nb.VDW_CUT_COMB_LB:
        # The LB kernel combines per-type parameters in FP32; the brute force uses the table.
        tol = 2e-6
    else:
        tol = 1e-9
    fb, eljb, eelb, _ = oracle.bruteforce(ws, ic)
    assert rel_rms(fo, fb) < tol
    assert eel == pytest.approx(eelb, rel=1e-9)
    assert elj == pytest.approx(eljb, rel=max(tol, 1e-9) * 10)


def _options(**kw):
    o = nb.NBKernelOptions(**kw)
    return o


def test_force_calculator_argument_checks():
    w = nb.systems.water_box(30, None, seed=1)
    params = np.stack([6.0 * w.c6c12[..., 0], 12.0 * w.c6c12[..., 1]], axis=-1)
    with pytest.raises(nb.NbnxmError, match="GPU only"):
        nb.ForceCalculator(w.types, params, w.q, w.excl_ranges, w.excl_elements, _options(useGpu=False))
    with pytest.raises(nb.NbnxmError, match="Coulomb type"):
        nb.ForceCalculator(w.types, params, w.q, options=_options(coulombType=3))
    bad = w.types.copy()
    bad[0] = 7
    with pytest.raises(nb.NbnxmError, match="type id"):
        nb.ForceCalculator(bad, params, w.q)
    with pytest.raises(nb.NbnxmError, match="rlist"):
        nb.ForceCalculator(w.types, params, w.q, options=_options(rlist=0.5))
    calc = nb.ForceCalculator(w.types, params, w.q, w.excl_ranges, w.excl_elements)
    with pytest.raises(nb.NbnxmError, match="without updating pairlist"):
        calc.compute(w.x, w.box)
    with pytest.raises(nb.NbnxmError, match="different number"):
        calc.update_pairlist(w.x[:-1].copy(), w.box)
    # A 30-molecule box is ~0.96 nm: below the 1.0 nm list radius.
    with pytest.raises(nb.NbnxmError, match="box edge"):
        calc.update_pairlist(w.x.copy(), w.box)
    lib = nb.load()
    assert lib.nbnxm_force_calculator_create(1, None, 1, None, None, None, None, None, None) == -1
    assert lib.nbnxm_force_calculator_update_pairlist(None, 0, None, None) == -1
    assert lib.nbnxm_force_calculator_compute(None, 0, None, None, None, None, None) == -1
    assert lib.nbnxm_gpu_get_virial(None, None) == -1


def test_verlet_argument_checks():
    """nbnxm_verlet_* rejects bad arguments before touching the GPU."""
    lib = nb.load()
    ic = nb.interaction_const(1.0, 1e-5)
    p = nb._lib.VerletParams()
    p.nstlist, p.rlistOuter, p.rlistInner, p.nstlistPrune, p.accumulation = 20, 1.1, 1.05, 4, nb.ACCUM_FIXED64
    types = np.zeros(4, dtype=np.int32)
    nbfp = np.array([1e-3, 1e-6])
    q = np.zeros(4)
    out = ctypes.c_void_p()
    ptr = nb._lib.ptr
    assert lib.nbnxm_verlet_create(4, ptr(types), 1, ptr(nbfp), ptr(q), None, None, ctypes.byref(ic), ctypes.byref(p), None) == -1
    bad = nb._lib.VerletParams()
    bad.nstlist, bad.rlistOuter, bad.rlistInner, bad.accumulation = 20, 1.1, 0.9, nb.ACCUM_FIXED64
    assert lib.nbnxm_verlet_create(4, ptr(types), 1, ptr(nbfp), ptr(q), None, None, ctypes.byref(ic), ctypes.byref(bad), ctypes.byref(out)) == -1
    assert "rlistInner" in lib.nbnxm_last_error().decode()
    bad_types = np.array([0, 0, 3, 0], dtype=np.int32)
    assert lib.nbnxm_verlet_create(4, ptr(bad_types), 1, ptr(nbfp), ptr(q), None, None, ctypes.byref(ic), ctypes.byref(p), ctypes.byref(out)) == -1
    assert lib.nbnxm_verlet_step(None, 0, None, None, None, 0, 0, None, 0) == -1
    assert lib.nbnxm_verlet_get_stats(None, None, None) == -1
    assert lib.nbnxm_verlet_get_list(None, None, None, None, None, None) == -1
