"""Pins the CPU oracle: reference golden vectors, closed forms, brute force."""
import math

import numpy as np
import pytest
from scipy.special import erf

import nbnxm_amd as nb
import oracle
from helpers import rel_rms, wrapped_float_system

GOLDEN_CASES = [
    ("CoulombEwald_VdwCutCombNone", "none", nb.VDW_CUT_TYPE_TABLE),
    ("CoulombEwald_VdwCutCombGeom", "geom", nb.VDW_CUT_COMB_GEOM),
    ("CoulombEwald_VdwCutCombLB", "lb", nb.VDW_CUT_COMB_LB),
]
REF_FORCE_TOL = 0.0119  # kernel_test.cpp:663-698 in FP32


def test_ewald_coeff_closed_form():
    for rc, rtol in [(1.0, 1e-5), (0.9, 1e-6), (1.2, 1e-4)]:
        beta = oracle.ewald_coeff(rc, rtol)
        assert math.erfc(beta * rc) == pytest.approx(rtol, rel=1e-12)
        assert nb.ewald_coeff(rc, rtol) == beta
    # The value quoted for the benchmark configurations.
    assert oracle.ewald_coeff(1.0, 1e-5) == pytest.approx(3.1234132743, abs=1e-9)


def _pmecorr_exact(z2):
    z = math.sqrt(z2)
    return 2.0 * math.exp(-z2) / (math.sqrt(math.pi) * z2) - erf(z) / (z2 * z)


def test_pmecorrF_closed_form():
    # simd_math.h:1604-1609: absolute error below 1e-6 for beta*r <= 4 (z2 <= 16).
    z2 = np.linspace(1e-3, 16.0, 4001)
    err = [abs(oracle.pmecorrF(v) - _pmecorr_exact(v)) for v in z2]
    assert max(err) < 1e-6
    # Small-argument limit -4/(3 sqrt(pi)).
    assert oracle.pmecorrF(1e-8) == pytest.approx(-4.0 / (3.0 * math.sqrt(math.pi)), rel=1e-6)


def test_bruteforce_two_and_three_atoms():
    """Hand-checkable pair sums: two ions and an excluded pair."""
    ic = nb.interaction_const(1.0, 1e-5)
    beta, epsfac = ic.ewaldcoeff_q, ic.epsfac
    c6, c12 = 1e-3, 1e-6
    s = nb.systems.System(box=np.array([3.0, 3.0, 3.0]), x=np.array([[0.5, 0.5, 0.5], [1.0, 0.5, 0.5]]),
                          q=np.array([1.0, -1.0]), types=np.array([0, 0], dtype=np.int32),
                          c6c12=np.array([[[c6, c12]]]), excl_ranges=np.array([0, 1, 2], dtype=np.int32),
                          excl_elements=np.array([0, 1], dtype=np.int32))
    f, elj, eel, npairs = oracle.bruteforce(s, ic)
    r = 0.5
    e_coul = -epsfac * (math.erfc(beta * r) / r - ic.sh_ewald) - 2 * epsfac * beta / math.sqrt(math.pi)
    e_lj = c12 * (r ** -12 - 1.0) - c6 * (r ** -6 - 1.0)
    fr = (12 * c12 * r ** -12 - 6 * c6 * r ** -6) / r ** 2 \
        - epsfac * (math.erfc(beta * r) / r ** 3 + 2 * beta / math.sqrt(math.pi) * math.exp(-(beta * r) ** 2) / r ** 2)
    assert npairs == 1
    assert eel == pytest.approx(e_coul, rel=1e-12)
    assert elj == pytest.approx(e_lj, rel=1e-12)
    assert f[0, 0] == pytest.approx(-fr * r, rel=1e-12)
    assert f[1, 0] == pytest.approx(fr * r, rel=1e-12)
    # Excluding the pair leaves only the Ewald correction -qq erf(beta r)/r, no LJ.
    s.excl_ranges = np.array([0, 2, 4], dtype=np.int32)
    s.excl_elements = np.array([0, 1, 0, 1], dtype=np.int32)
    f, elj, eel, npairs = oracle.bruteforce(s, ic)
    assert npairs == 0 and elj == 0.0
    assert eel == pytest.approx(epsfac * math.erf(beta * r) / r - 2 * epsfac * beta / math.sqrt(math.pi), rel=1e-12)


def test_bruteforce_translation_invariance():
    s = nb.systems.water_box(60, None, seed=3)
    ic = nb.interaction_const(1.0, 1e-5)
    f1, elj1, eel1, n1 = oracle.bruteforce(s, ic)
    s.x = s.x + np.array([0.37, -1.1, 2.9])
    f2, elj2, eel2, n2 = oracle.bruteforce(s, ic)
    assert n1 == n2
    assert rel_rms(f2, f1) < 1e-10
    assert eel2 == pytest.approx(eel1, rel=1e-10)


@pytest.mark.parametrize("case,rule,vdw", GOLDEN_CASES)
def test_golden_bruteforce(golden, case, rule, vdw):
    """The FP64 all-image sum reproduces the reference's double-precision refdata."""
    system = nb.systems.spc27_system(golden, rule)
    ic = nb.interaction_const(0.9, 1e-6)
    f, elj, eel, _ = oracle.bruteforce(system, ic)
    ref = np.array(golden["cases"][case]["forces"])
    assert np.abs(f - ref).max() < 1e-5
    assert elj == pytest.approx(golden["cases"][case]["vdw_energy"], abs=1e-8)
    assert eel == pytest.approx(golden["cases"][case]["coulomb_energy"], abs=1e-5)


@pytest.mark.parametrize("case,rule,vdw", GOLDEN_CASES)
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_golden_list_kernel(golden, case, rule, vdw, precision):
    """The list-based restatements match the refdata within the reference test tolerance."""
    system = nb.systems.spc27_system(golden, rule)
    ic = nb.interaction_const(0.9, 1e-6)
    search, plist, atoms = nb.build(system, 0.9, vdw)
    f, fs, elj, eel = oracle.kernel(plist, atoms, ic, True, precision)
    ref = np.array(golden["cases"][case]["forces"])
    fo = atoms.to_original_order(f, system.n)
    tol = 3e-3 if precision == "f64" else REF_FORCE_TOL
    assert np.abs(fo - ref).max() < tol
    assert abs(elj - golden["cases"][case]["vdw_energy"]) < tol
    assert abs(eel - golden["cases"][case]["coulomb_energy"]) < 10 * tol


@pytest.mark.parametrize("n_mol,max_pairs", [(216, 0), (216, 16), (500, 0), (700, 24)])
def test_list_kernel_equals_bruteforce(n_mol, max_pairs):
    """List semantics (images, exclusions, Newton masks, splitting) vs the list-free sum."""
    s = wrapped_float_system(nb.systems.water_box(n_mol, None, seed=n_mol))
    ic = nb.interaction_const(1.0, 1e-5)
    search, plist, atoms = nb.build(s, 1.1, nb.VDW_CUT_TYPE_TABLE, max_pairs)
    f, fs, elj, eel = oracle.kernel(plist, atoms, ic, True, "f64")
    fb, eljb, eelb, npairs = oracle.bruteforce(s, ic)
    assert oracle.count_pairs(plist, atoms, 1.0) == npairs
    assert rel_rms(atoms.to_original_order(f, s.n), fb) < 1e-6
    assert eel == pytest.approx(eelb, rel=1e-7)
    assert elj == pytest.approx(eljb, rel=1e-6)


def test_f32_restatement_close_to_f64():
    s = nb.systems.water_box(700, None, seed=9)
    ic = nb.interaction_const(1.0, 1e-5)
    search, plist, atoms = nb.build(s, 1.1, nb.VDW_CUT_COMB_GEOM)
    f64, _, elj64, eel64 = oracle.kernel(plist, atoms, ic, True, "f64")
    f32, _, elj32, eel32 = oracle.kernel(plist, atoms, ic, True, "f32")
    real = atoms.atom_indices >= 0
    assert rel_rms(f32[real], f64[real]) < 5e-6


def test_newton_shift_forces():
    """Pair forces cancel (sum_i f_i = 0) and the central shift collects no
    shift force (nbnxm_cuda_kernel.cuh:694-698)."""
    s = nb.systems.water_box(300, None, seed=2)
    ic = nb.interaction_const(1.0, 1e-5)
    search, plist, atoms = nb.build(s, 1.1, nb.VDW_CUT_COMB_GEOM)
    f, fs, _, _ = oracle.kernel(plist, atoms, ic, False, "f64")
    assert np.abs(f.sum(axis=0)).max() < 1e-9 * np.abs(f).max()
    assert fs[nb.CENTRAL_SHIFT].tolist() == [0.0, 0.0, 0.0]


def _prune_predicate(plist, xq, shift_vec, r):
    """Independent FP64 any-pair-within-r test per (group, bit), with the margin
    to the threshold, so that FP32 rounding cannot flip a decided bit."""
    xq = xq.reshape(-1, 4).astype(np.float64)
    sv = shift_vec.reshape(-1, 3).astype(np.float64)
    decided = {}
    for e in plist.sci:
        for g in range(e["cjPackedBegin"], e["cjPackedEnd"]):
            for jm in range(4):
                cj = plist.cj_packed[g]["cj"][jm]
                xj = xq[cj * 8:(cj + 1) * 8, :3]
                for im in range(8):
                    bit = jm * 8 + im
                    if not (int(plist.cj_packed[g]["imask"]) >> bit) & 1:
                        continue
                    ci = e["sci"] * 8 + im
                    xi = xq[ci * 8:(ci + 1) * 8, :3] + sv[e["shift"]]
                    d2 = ((xi[:, None, :] - xj[None, :, :]) ** 2).sum(-1)
                    dmin = float(d2.min())
                    decided[(g, bit)] = (dmin < r * r, abs(dmin - r * r) > 1e-4)
    return decided


def test_prune_oracle_against_predicate():
    s = nb.systems.water_box(300, None, seed=21)
    search, plist, atoms = nb.build(s, 1.1, nb.VDW_CUT_COMB_GEOM)
    inner, outer, count = oracle.prune(plist, atoms.xq, atoms.shift_vec, 1.1, 1.02, True)
    pred_outer = _prune_predicate(plist, atoms.xq, atoms.shift_vec, 1.1)
    pred_inner = _prune_predicate(plist, atoms.xq, atoms.shift_vec, 1.02)
    for (g, bit), (inside, clear) in pred_outer.items():
        if clear:
            assert bool((int(outer[g]) >> bit) & 1) == inside
    for (g, bit), (inside, clear) in pred_inner.items():
        if clear:
            assert bool((int(inner[g]) >> bit) & 1) == inside
    assert np.all((inner & ~outer) == 0)
    # Histogram bins: 8192 - count - 1, clamped at 0.
    pairs = [sum(bin(int(inner[g])).count("1") for g in range(e["cjPackedBegin"], e["cjPackedEnd"])) for e in plist.sci]
    assert count.tolist() == [max(8192 - p - 1, 0) for p in pairs]


def test_rolling_prune_only_adds_back():
    s = nb.systems.water_box(300, None, seed=22)
    search, plist, atoms = nb.build(s, 1.1, nb.VDW_CUT_COMB_GEOM)
    inner, outer, _ = oracle.prune(plist, atoms.xq, atoms.shift_vec, 1.1, 1.02, True)
    xq = atoms.xq.reshape(-1, 4).copy()
    real = atoms.atom_indices >= 0
    xq[real, :3] += np.random.default_rng(0).normal(0, 0.02, size=(real.sum(), 3)).astype(np.float32)
    inner2 = inner.copy()
    for part in range(2):
        inner2, _, _ = oracle.prune(plist, xq.reshape(-1), atoms.shift_vec, 1.1, 1.02, False, part, 2,
                                    imask_outer=outer, imask_inner=inner2)
    assert np.all((inner & ~inner2) == 0)       # nothing removed
    assert np.all((inner2 & ~outer) == 0)       # bounded by the outer list
    assert (inner2 != inner).any()


def test_sci_sort_stable_descending():
    counts = np.array([8190, 8000, 8190, 0, 8000, 5], dtype=np.int32)
    order = oracle.sci_sort_order(counts)
    assert order.tolist() == [3, 5, 1, 4, 0, 2]


def test_cpu4xm_baseline_matches_fp64():
    """The CPU 4xM SIMD baseline computes the same forces and energies."""
    s = nb.systems.water_box(900, None, seed=31)
    ic = nb.interaction_const(1.0, 1e-5)
    search, plist, atoms = nb.build(s, 1.1, nb.VDW_CUT_COMB_GEOM)
    f64, _, elj64, eel64 = oracle.kernel(plist, atoms, ic, True, "f64")
    k = oracle.Cpu4xm(plist, atoms, ic, 1.04, threads=2)
    f, vlj, vc = k.compute(True)
    real = atoms.atom_indices >= 0
    assert rel_rms(f[real], f64[real]) < 1e-5
    assert vlj == pytest.approx(elj64, rel=1e-5)
    assert vc == pytest.approx(eel64, rel=1e-4)
    f2, _, _ = k.compute(False)
    assert rel_rms(f2[real], f[real]) < 1e-6
