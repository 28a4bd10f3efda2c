"""Pins the CPU oracle: reference golden vectors, closed forms, brute force."""
import math

import numpy as np
import pytest
from scipy.special import erf

import nbnxm_amd as nb
import oracle
from helpers import GOLDEN_ALL, golden_case, rel_rms, wrapped_float_system


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


@pytest.mark.parametrize("case", GOLDEN_ALL)
def test_golden_bruteforce(golden, case):
    """The FP64 all-image sum reproduces the reference's double-precision refdata
    for every Coulomb / LJ flavour."""
    system, ic, _, _ = golden_case(golden, case)
    f, elj, eel, _ = oracle.bruteforce(system, ic)
    ref = np.array(golden["cases"][case]["forces"])
    assert np.abs(f - ref).max() < 1e-5
    assert elj == pytest.approx(golden["cases"][case]["vdw_energy"], abs=1e-8)
    assert eel == pytest.approx(golden["cases"][case]["coulomb_energy"], abs=1e-5)


@pytest.mark.parametrize("case", GOLDEN_ALL)
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_golden_list_kernel(golden, case, precision):
    """The list-based restatements match the refdata within the reference test
    tolerance (kernel_test.cpp:663-698), energies included."""
    system, ic, vdw, tol = golden_case(golden, case)
    search, plist, atoms = nb.build(system, 0.9, vdw)
    f, fs, elj, eel = oracle.kernel(plist, atoms, ic, True, precision)
    ref = np.array(golden["cases"][case]["forces"])
    fo = atoms.to_original_order(f, system.n)
    if precision == "f64":
        tol = 3e-3
    assert np.abs(fo - ref).max() < tol
    assert abs(elj - golden["cases"][case]["vdw_energy"]) < tol
    assert abs(eel - golden["cases"][case]["coulomb_energy"]) < 10 * tol
    # The force-only pass gives the same forces within 10 ulp at magnitude 1000
    # (kernel_test.cpp:720-729; the LB force-only form rounds differently).
    f2, _, _, _ = oracle.kernel(plist, atoms, ic, False, precision)
    assert np.abs(atoms.to_original_order(f2, system.n) - fo).max() < (1e-9 if precision == "f64" else 1000 * 10 * 2.0 ** -23)


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


FLAVOURS = {
    "rf": dict(coulomb="rf"),
    "cutoff": dict(coulomb="rf", plain_cutoff=True),
    "twin": dict(rvdw=0.8),
    "forceswitch": dict(vdw_modifier="forceswitch", rvdw_switch=0.8),
    "potswitch": dict(vdw_modifier="potswitch", rvdw_switch=0.75),
    "ljpme": dict(lj_pme=True, ewald_rtol_lj=1e-3),
    "rf_twin_potswitch": dict(coulomb="rf", rvdw=0.9, vdw_modifier="potswitch", rvdw_switch=0.7),
    "ljpme_twin": dict(lj_pme=True, rvdw=0.85),
}


@pytest.mark.parametrize("name", sorted(FLAVOURS))
def test_flavour_list_kernel_equals_bruteforce(name):
    """List semantics of every flavour (excluded pairs, Newton masks, twin cut-off
    check) against the list-free FP64 sum on identical FP32-rounded inputs."""
    s = wrapped_float_system(nb.systems.water_box(216, None, seed=5))
    ic = nb.interaction_const_ex(1.0, ewald_rtol=1e-5, **FLAVOURS[name])
    search, plist, atoms = nb.build(s, 1.1, nb.VDW_CUT_TYPE_TABLE, 16)
    f, fs, elj, eel = oracle.kernel(plist, atoms, ic, True, "f64")
    fb, eljb, eelb, npairs = oracle.bruteforce(s, ic)
    assert rel_rms(atoms.to_original_order(f, s.n), fb) < 1e-6
    assert eel == pytest.approx(eelb, rel=1e-7)
    assert elj == pytest.approx(eljb, rel=1e-6)
    f32, _, elj32, eel32 = oracle.kernel(plist, atoms, ic, True, "f32")
    real = atoms.atom_indices >= 0
    assert rel_rms(f32[real], f[real]) < 5e-6
    # Energies accumulate in FP32 per entry, as the CUDA kernel does.
    assert eel32 == pytest.approx(eel, rel=1e-4)
    assert elj32 == pytest.approx(elj, rel=1e-4)


def _two_atoms(r, q=(1.0, -1.0), c6=1e-3, c12=1e-6, excluded=False):
    ranges = np.array([0, 2, 4], dtype=np.int32) if excluded else np.array([0, 1, 2], dtype=np.int32)
    elements = np.array([0, 1, 0, 1], dtype=np.int32) if excluded else np.array([0, 1], dtype=np.int32)
    return nb.systems.System(box=np.array([3.0, 3.0, 3.0]), x=np.array([[0.5, 0.5, 0.5], [0.5 + r, 0.5, 0.5]]),
                             q=np.array(q), types=np.array([0, 0], dtype=np.int32),
                             c6c12=np.array([[[c6, c12]]]), excl_ranges=ranges, excl_elements=elements)


def test_reaction_field_closed_form():
    """Two ions: V = qq (1/r + k_rf r^2 - c_rf), F/r = qq (1/r^3 - 2 k_rf), self
    energy -c_rf/2 epsfac sum q^2 (rf_util.cpp:50-62 constants, eps_rf = inf)."""
    rc, r = 1.0, 0.4
    ic = nb.interaction_const_ex(rc, coulomb="rf", epsilon_rf=0.0)
    assert ic.k_rf == pytest.approx(0.5 / rc ** 3) and ic.c_rf == pytest.approx(1.5 / rc)
    s = _two_atoms(r, c6=0.0, c12=0.0)
    f, elj, eel, _ = oracle.bruteforce(s, ic)
    qq = -ic.epsfac
    assert eel == pytest.approx(qq * (1 / r + ic.k_rf * r * r - ic.c_rf) - 0.5 * ic.c_rf * ic.epsfac * 2, rel=1e-12)
    assert f[0, 0] == pytest.approx(-qq * (1 / r ** 3 - 2 * ic.k_rf) * r, rel=1e-12)
    # Excluded: only the reaction-field correction qq (k_rf r^2 - c_rf), force -2 k_rf qq r.
    f, _, eel, _ = oracle.bruteforce(_two_atoms(r, c6=0.0, c12=0.0, excluded=True), ic)
    assert eel == pytest.approx(qq * (ic.k_rf * r * r - ic.c_rf) - ic.c_rf * ic.epsfac, rel=1e-12)
    assert f[0, 0] == pytest.approx(2 * ic.k_rf * qq * r, rel=1e-12)
    # Finite epsilon_rf: k_rf = (eps_rf - eps_r) / (2 eps_rf + eps_r) / rc^3.
    ic2 = nb.interaction_const_ex(1.2, coulomb="rf", epsilon_rf=78.0)
    assert ic2.k_rf == pytest.approx(77.0 / 157.0 / 1.2 ** 3) and ic2.c_rf == pytest.approx(1 / 1.2 + ic2.k_rf * 1.44)


def test_lj_switch_continuity():
    """Force switch: force and energy vanish at rvdw, force equals plain LJ
    below rvdw_switch; potential switch: energy and force vanish at rvdw and equal
    the unshifted LJ below rvdw_switch (interaction_const.cpp:199-246)."""
    c6, c12 = 2e-3, 3e-6
    rc, rsw = 1.0, 0.7
    for modifier in ("forceswitch", "potswitch"):
        ic = nb.interaction_const_ex(rc, vdw_modifier=modifier, rvdw_switch=rsw)
        ic.epsfac = 0.0  # LJ only
        for r in (0.5, 0.69):
            f, elj, _, _ = oracle.bruteforce(_two_atoms(r, q=(0.0, 0.0), c6=c6, c12=c12), ic)
            assert -f[0, 0] == pytest.approx((12 * c12 / r ** 13 - 6 * c6 / r ** 7), rel=1e-10)
            if modifier == "potswitch":
                assert elj == pytest.approx(c12 / r ** 12 - c6 / r ** 6, rel=1e-10)
        r = rc - 1e-7
        f, elj, _, _ = oracle.bruteforce(_two_atoms(r, q=(0.0, 0.0), c6=c6, c12=c12), ic)
        assert abs(elj) < 1e-12 and abs(f[0, 0]) < 1e-9
        # Energy is the integral of the force: central difference in the switch region.
        h, r = 1e-5, 0.85
        e1 = oracle.bruteforce(_two_atoms(r - h, q=(0.0, 0.0), c6=c6, c12=c12), ic)[1]
        e2 = oracle.bruteforce(_two_atoms(r + h, q=(0.0, 0.0), c6=c6, c12=c12), ic)[1]
        f, _, _, _ = oracle.bruteforce(_two_atoms(r, q=(0.0, 0.0), c6=c6, c12=c12), ic)
        assert -f[0, 0] == pytest.approx(-(e2 - e1) / (2 * h), rel=1e-6)


def test_lj_pme_closed_form():
    """With equal table and grid C6 the dispersion becomes the real-space Ewald
    kernel -C6 g(beta r)/r^6, g(x) = exp(-x^2)(1 + x^2 + x^4/2) (ewald_utils.cpp:73-82),
    plus constants; calc_ewaldcoeff_lj solves g(beta rc) = rtol."""
    rc, rtol = 1.0, 1e-3
    beta = nb.ewald_coeff_lj(rc, rtol)
    g = lambda x: math.exp(-x * x) * (1 + x * x + x ** 4 / 2)
    assert g(beta * rc) == pytest.approx(rtol, rel=1e-12)
    ic = nb.interaction_const_ex(rc, lj_pme=True, ewald_rtol_lj=rtol)
    assert ic.ewaldcoeff_lj == beta
    ic.epsfac = 0.0
    c6 = 2e-3
    self_e = 2 * c6 * beta ** 6 / 12.0   # + sum C6_ii beta^6 / 12 (nbnxm_cuda_kernel.cuh:395-409)
    for r in (0.3, 0.6, 0.95):
        _, elj, _, _ = oracle.bruteforce(_two_atoms(r, q=(0.0, 0.0), c6=c6, c12=0.0), ic)
        expect = -c6 * g(beta * r) / r ** 6 - c6 * (-1 / rc ** 6) + c6 * ic.sh_lj_ewald
        assert elj - self_e == pytest.approx(expect, rel=1e-10)
    # An excluded pair keeps only the grid correction C6 (1 - g(beta r))/r^6.
    r = 0.3
    _, elj, _, _ = oracle.bruteforce(_two_atoms(r, q=(0.0, 0.0), c6=c6, c12=0.0, excluded=True), ic)
    assert elj - self_e == pytest.approx(c6 * (1 - g(beta * r)) / r ** 6, rel=1e-10)


def test_lj_pme_lb_closed_form():
    """Lorentz-Berthelot grid rule: for two types the grid C6 the oracle forms
    from the nbfp diagonals (nbfp_comb, atomdata.cpp:311-324, combined as
    nbnxm_cuda_kernel_utils.cuh:307-324) equals the textbook
    C6_ij = 4 sqrt(eps_i eps_j) ((sigma_i + sigma_j)/2)^6, so with the same
    real-space C6 the dispersion is -C6_ij g(beta r)/r^6 + constants; the
    geometric grid rule gives a different value for unequal sigma."""
    rc, rtol = 1.0, 1e-3
    beta = nb.ewald_coeff_lj(rc, rtol)
    g = lambda x: math.exp(-x * x) * (1 + x * x + x ** 4 / 2)
    sig_eps = np.array([[0.28, 0.3], [0.39, 0.8]])
    c6c12 = nb.systems.combine_params(sig_eps, "lb")
    c6ij = 4.0 * math.sqrt(0.3 * 0.8) * ((0.28 + 0.39) / 2) ** 6
    assert c6c12[0, 1, 0] == pytest.approx(c6ij, rel=1e-14)
    c6ii = [4.0 * e * sg ** 6 for sg, e in sig_eps]
    ic = nb.interaction_const_ex(rc, lj_pme="lb", ewald_rtol_lj=rtol)
    assert ic.ljPme == 2
    ic.epsfac = 0.0
    self_e = sum(c6ii) * beta ** 6 / 12.0
    for r in (0.3, 0.6, 0.95):
        s = nb.systems.System(box=np.array([3.0, 3.0, 3.0]), x=np.array([[0.5, 0.5, 0.5], [0.5 + r, 0.5, 0.5]]),
                              q=np.zeros(2), types=np.array([0, 1], dtype=np.int32), c6c12=c6c12,
                              excl_ranges=np.array([0, 1, 2], dtype=np.int32), excl_elements=np.array([0, 1], dtype=np.int32))
        f, elj, _, _ = oracle.bruteforce(s, ic)
        c12 = c6c12[0, 1, 1]
        expect = (-c6ij * g(beta * r) / r ** 6 + c6ij / rc ** 6 + c6ij * ic.sh_lj_ewald
                  + c12 * (1 / r ** 12 - 1 / rc ** 12))
        assert elj - self_e == pytest.approx(expect, rel=1e-10)
        # dV/dr of the same closed form (the C12 term is the plain LJ repulsion).
        dg = lambda x: -x ** 5 * math.exp(-x * x)  # d/dx g(x)
        dvdr = 6 * c6ij * g(beta * r) / r ** 7 - c6ij * beta * dg(beta * r) / r ** 6 - 12 * c12 / r ** 13
        assert f[1, 0] == pytest.approx(-dvdr, rel=1e-9)
        icg = nb.interaction_const_ex(rc, lj_pme="geom", ewald_rtol_lj=rtol)
        icg.epsfac = 0.0
        _, eljg, _, _ = oracle.bruteforce(s, icg)
        assert abs(eljg - elj) > 1e-6 * abs(elj)


def _mixed_lb_system(seed=7):
    """About 1,500 atoms: 5 chains of 50 atoms with 6 Lorentz-Berthelot LJ
    types and charges, 1-2/1-3 exclusions, in water."""
    return nb.systems.hetero_system(n_chains=5, chain_len=50, total_atoms=1500, n_types=6, seed=seed)


@pytest.mark.parametrize("rule", ["lb", "geom"])
def test_lj_pme_mixed_types_list_equals_bruteforce(rule):
    """Both grid rules on several LJ types (type table kernel): the FP64 list
    restatement equals the brute force, the FP32 one follows it."""
    s = wrapped_float_system(_mixed_lb_system())
    ic = nb.interaction_const_ex(1.0, lj_pme=rule, ewald_rtol_lj=1e-3)
    search, plist, atoms = nb.build(s, 1.1, nb.VDW_CUT_TYPE_TABLE, 16)
    f, fs, elj, eel = oracle.kernel(plist, atoms, ic, True, "f64")
    fb, eljb, eelb, npairs = oracle.bruteforce(s, ic)
    assert rel_rms(atoms.to_original_order(f, s.n), fb) < 1e-6
    assert eel == pytest.approx(eelb, rel=1e-7)
    assert elj == pytest.approx(eljb, rel=1e-6)
    f32, _, elj32, eel32 = oracle.kernel(plist, atoms, ic, True, "f32")
    real = atoms.atom_indices >= 0
    assert rel_rms(f32[real], f[real]) < 5e-6
    assert elj32 == pytest.approx(elj, rel=1e-4)


def test_ewald_table_against_exact():
    """The tabulated Ewald correction interpolates -d/dr erf(beta r)/r with the
    FP32 spline error eps * beta^3 of its spacing (forcetable.cpp:178-181,
    :280-294), and the library builds the same table."""
    for rc, rtol in [(1.0, 1e-5), (0.9, 1e-6), (1.2, 1e-4)]:
        beta = oracle.ewald_coeff(rc, rtol)
        scale, t32, t64 = oracle.ewald_table(beta, rc)
        assert len(t32) == int(np.float32(rc) * np.float32(scale)) + 2
        b = float(np.float32(beta))
        r = np.linspace(0.02, rc, 5001)
        x = scale * r
        i = x.astype(int)
        fr = x - i
        interp = t64[i] + fr * (t64[i + 1] - t64[i])
        exact = erf(b * r) / r ** 2 - 2 * b / math.sqrt(math.pi) * np.exp(-(b * r) ** 2) / r
        eps = 2.0 ** -23
        assert np.abs(interp - exact).max() < 2 * eps * b ** 3
        s2, p32, p64 = nb.ewald_correction_table(beta, rc)
        assert s2 == scale
        assert np.array_equal(p32, t32)
        assert np.allclose(p64, t64, rtol=1e-14, atol=1e-14)


TABULATED_CASES = [c for c in GOLDEN_ALL if not c.startswith("CoulombReactionField")]


@pytest.mark.parametrize("case", TABULATED_CASES)
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_golden_tabulated_ewald(golden, case, precision):
    """The tabulated Ewald kernels share the analytical cases' refdata
    (kernel_test.cpp:478-489); the FP32 tolerance is the same (:663-698)."""
    system, ic, vdw, tol = golden_case(golden, case)
    ic.tabulatedEwald = 1
    search, plist, atoms = nb.build(system, 0.9, vdw)
    f, fs, elj, eel = oracle.kernel(plist, atoms, ic, True, precision)
    ref = np.array(golden["cases"][case]["forces"])
    fo = atoms.to_original_order(f, system.n)
    assert np.abs(fo - ref).max() < tol
    assert abs(elj - golden["cases"][case]["vdw_energy"]) < tol
    assert abs(eel - golden["cases"][case]["coulomb_energy"]) < 10 * tol
    fb, eljb, eelb, _ = oracle.bruteforce(system, ic)
    assert np.abs(fb - ref).max() < tol
