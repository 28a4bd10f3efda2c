"""GPU parity of every interaction flavour (reaction field, plain cut-off, twin
cut-off, LJ force / potential switch, LJ-PME with both grid rules) against the reference's 18
refdata cases and the CPU oracle.  Calls go through the C ABI."""
import numpy as np
import pytest

import nbnxm_amd as nb
import oracle
from conftest import announce
from helpers import GOLDEN_ALL, golden_case, rel_rms

pytestmark = pytest.mark.gpu

ACCUMULATIONS = [nb.ACCUM_FIXED64, nb.ACCUM_FLOAT, nb.ACCUM_FP64]
# kernel_test.cpp:720-729: force-only vs energy kernel within 10 ulp at magnitude 1000.
F_VS_VF_TOL = 1000 * 10 * 2.0 ** -23


def _run(ic, atoms, plist, rlist, accumulation, energy=True):
    task = nb.GpuTask(ic, atoms, rlist, None, accumulation)
    task.init_pairlist(plist)
    task.launch(compute_energy=energy, compute_virial=True)
    f, fd, elj, eel, fshift = task.fetch()
    forces = (fd if fd is not None else f).reshape(-1, 3)
    return task, forces, elj, eel, fshift.reshape(-1, 3)


@pytest.mark.parametrize("case", GOLDEN_ALL)
@pytest.mark.parametrize("accumulation", ACCUMULATIONS)
def test_golden_flavours(golden, case, accumulation):
    announce(f"flavour golden {case} accumulation={accumulation}")
    system, ic, vdw, tol = golden_case(golden, case)
    search, plist, atoms = nb.build(system, 0.9, vdw)
    task, f, elj, eel, fshift = _run(ic, atoms, plist, 0.9, accumulation)
    ref = np.array(golden["cases"][case]["forces"])
    fo = atoms.to_original_order(f, system.n)
    assert np.abs(fo - ref).max() < tol
    assert abs(elj - golden["cases"][case]["vdw_energy"]) < tol
    assert abs(eel - golden["cases"][case]["coulomb_energy"]) < 10 * tol
    precision = "f64" if accumulation == nb.ACCUM_FP64 else "f32"
    fo_, fso, eljo, eelo = oracle.kernel(plist, atoms, ic, True, precision)
    real = atoms.atom_indices >= 0
    if precision == "f64":
        # Same FP64 arithmetic with exact erf/exp: agreement to rounding.
        assert rel_rms(f[real], fo_[real]) < 1e-10
        assert elj == pytest.approx(eljo, rel=1e-10, abs=1e-10)
        assert eel == pytest.approx(eelo, rel=1e-10)
        assert np.abs(fshift - fso).max() < 1e-8 * max(1.0, np.abs(fso).max())
    else:
        assert rel_rms(f[real], fo_[real]) < 2e-6
        assert np.abs(fshift - fso).max() < 1e-4 * max(1.0, np.abs(fso).max())
    # Force-only pass, after gpu_clear_outputs as at the start of an MD step
    # (float accumulation keeps adding into the same buffer otherwise).
    task.clear_outputs()
    task.launch(compute_energy=False, compute_virial=False)
    f2, fd2, _, _, _ = task.fetch()
    f2 = (fd2 if fd2 is not None else f2).reshape(-1, 3)
    assert np.abs(f2 - f).max() < F_VS_VF_TOL


TABULATED_CASES = [c for c in GOLDEN_ALL if not c.startswith("CoulombReactionField")]


@pytest.mark.parametrize("case", TABULATED_CASES)
@pytest.mark.parametrize("accumulation", ACCUMULATIONS)
def test_golden_tabulated_ewald(golden, case, accumulation):
    """Tabulated Ewald kernels against the shared refdata (kernel_test.cpp:478-489)
    and the oracle's own table interpolation."""
    announce(f"tabulated golden {case} accumulation={accumulation}")
    system, ic, vdw, tol = golden_case(golden, case)
    ic.tabulatedEwald = 1
    search, plist, atoms = nb.build(system, 0.9, vdw)
    task, f, elj, eel, fshift = _run(ic, atoms, plist, 0.9, accumulation)
    ref = np.array(golden["cases"][case]["forces"])
    assert np.abs(atoms.to_original_order(f, system.n) - ref).max() < tol
    assert abs(elj - golden["cases"][case]["vdw_energy"]) < tol
    assert abs(eel - golden["cases"][case]["coulomb_energy"]) < 10 * tol
    precision = "f64" if accumulation == nb.ACCUM_FP64 else "f32"
    fo, _, eljo, eelo = oracle.kernel(plist, atoms, ic, True, precision)
    real = atoms.atom_indices >= 0
    assert rel_rms(f[real], fo[real]) < (1e-10 if precision == "f64" else 2e-6)
    assert task._lib.nbnxm_gpu_is_kernel_ewald_analytical(task._h) == 0


FLAVOURS = {
    "tabulated": dict(tabulated=True),
    "tabulated_twin_forceswitch": dict(tabulated=True, rvdw=0.9, vdw_modifier="forceswitch", rvdw_switch=0.7),
    "rf": dict(coulomb="rf"),
    "rf_epsrf78": dict(coulomb="rf", epsilon_rf=78.0),
    "cutoff": dict(coulomb="rf", plain_cutoff=True),
    "twin": dict(rvdw=0.8),
    "forceswitch": dict(vdw_modifier="forceswitch", rvdw_switch=0.8),
    "potswitch": dict(vdw_modifier="potswitch", rvdw_switch=0.75),
    "ljpme": dict(lj_pme=True, ewald_rtol_lj=1e-3),
    "rf_twin_potswitch": dict(coulomb="rf", rvdw=0.9, vdw_modifier="potswitch", rvdw_switch=0.7),
    "ljpme_twin": dict(lj_pme=True, rvdw=0.85),
}


@pytest.fixture(scope="module")
def water3k():
    s = nb.systems.water_box(1000, None, seed=41)
    search, plist, atoms = nb.build(s, 1.1, nb.VDW_CUT_TYPE_TABLE)
    return s, plist, atoms


@pytest.mark.parametrize("name", ["tabulated", "rf", "twin", "forceswitch", "potswitch", "ljpme"])
def test_i_cluster_placement_flavours(water3k, name):
    """Each compile-time flavour class with the i-cluster data in LDS and in
    registers, F and F+E: fixed-point forces, energies and shift forces agree
    bit for bit (combination rule; LJ-PME with the type table), and match the
    FP64 oracle."""
    announce(f"flavour i-cluster placement {name}")
    s, plist, atoms = water3k
    if name != "ljpme":
        _, plist, atoms = nb.build(s, 1.1, nb.VDW_CUT_COMB_GEOM)
    ic = nb.interaction_const_ex(1.0, ewald_rtol=1e-5, **FLAVOURS[name])
    real = atoms.atom_indices >= 0
    for energy in (False, True):
        out = []
        for mode in (0, 1):
            task = nb.GpuTask(ic, atoms, 1.1, None, nb.ACCUM_FIXED64)
            task.set_i_cluster_in_lds(mode)
            task.init_pairlist(plist)
            task.launch(compute_energy=energy, compute_virial=True)
            out.append(task.fetch())
            del task
        (_, fd0, elj0, eel0, fs0), (_, fd1, elj1, eel1, fs1) = out
        assert np.array_equal(fd0, fd1), energy
        assert np.array_equal(fs0, fs1) and elj0 == elj1 and eel0 == eel1
    f64, _, elj64, eel64 = oracle.kernel(plist, atoms, ic, True, "f64")
    assert rel_rms(fd1.reshape(-1, 3)[real], f64[real]) < 5e-6
    assert elj1 == pytest.approx(elj64, rel=2e-5)
    assert eel1 == pytest.approx(eel64, rel=2e-5)


@pytest.mark.parametrize("name", sorted(FLAVOURS))
def test_flavour_water3k(water3k, name):
    """3000-atom water with a buffered list: FP32 kernel vs the FP64 oracle,
    FP64 mode vs the FP64 oracle, fixed point bitwise reproducible."""
    announce(f"flavour water3k {name}")
    s, plist, atoms = water3k
    ic = nb.interaction_const_ex(1.0, ewald_rtol=1e-5, **FLAVOURS[name])
    f64, fs64, elj64, eel64 = oracle.kernel(plist, atoms, ic, True, "f64")
    real = atoms.atom_indices >= 0
    task, f, elj, eel, fshift = _run(ic, atoms, plist, 1.1, nb.ACCUM_FIXED64)
    assert rel_rms(f[real], f64[real]) < 5e-6
    assert elj == pytest.approx(elj64, rel=2e-5)
    assert eel == pytest.approx(eel64, rel=2e-5)
    assert np.abs(fshift - fs64).max() <= 1e-4 * max(1.0, np.abs(fs64).max())
    task.launch(compute_energy=True, compute_virial=True)
    _, fd2, elj2, eel2, _ = task.fetch()
    assert np.array_equal(fd2.reshape(-1, 3).view(np.uint64), f.view(np.uint64))
    assert elj2 == elj and eel2 == eel
    _, fd, eljd, eeld, _ = _run(ic, atoms, plist, 1.1, nb.ACCUM_FP64)
    assert rel_rms(fd[real], f64[real]) < 1e-10
    assert eljd == pytest.approx(elj64, rel=1e-10)
    assert eeld == pytest.approx(eel64, rel=1e-10)


@pytest.mark.parametrize("vdw", [nb.VDW_CUT_COMB_GEOM, nb.VDW_CUT_COMB_LB])
def test_ljpme_with_combination_kernel(vdw):
    """LJ-PME with a combination-rule LJ kernel: the grid part still reads the
    types and the type table (the standalone flavour kernel path)."""
    announce(f"ljpme comb vdw={vdw}")
    s = nb.systems.water_box(1000, None, seed=43)
    search, plist, atoms = nb.build(s, 1.1, vdw)
    ic = nb.interaction_const_ex(1.0, ewald_rtol=1e-5, lj_pme=True, ewald_rtol_lj=1e-3)
    f64, fs64, elj64, eel64 = oracle.kernel(plist, atoms, ic, True, "f64")
    real = atoms.atom_indices >= 0
    task, f, elj, eel, fshift = _run(ic, atoms, plist, 1.1, nb.ACCUM_FIXED64)
    assert rel_rms(f[real], f64[real]) < 5e-6
    assert elj == pytest.approx(elj64, rel=2e-5)
    assert eel == pytest.approx(eel64, rel=2e-5)


@pytest.fixture(scope="module")
def mixed_lb():
    """About 1,500 atoms: 5 chains of 50 atoms with 6 Lorentz-Berthelot LJ
    types, charges and 1-2/1-3 exclusions, in water (8 LJ types in all)."""
    return nb.systems.hetero_system(n_chains=5, chain_len=50, total_atoms=1500, n_types=6, seed=7)


LJPME_LB_CASES = {
    "lb": dict(lj_pme="lb", ewald_rtol_lj=1e-3),            # compile-time LJ-PME class
    "lb_twin": dict(lj_pme="lb", ewald_rtol_lj=1e-3, rvdw=0.85),  # run-time class
    "geom": dict(lj_pme="geom", ewald_rtol_lj=1e-3),
}


@pytest.mark.parametrize("case", sorted(LJPME_LB_CASES))
@pytest.mark.parametrize("vdw", [nb.VDW_CUT_TYPE_TABLE, nb.VDW_CUT_COMB_LB])
@pytest.mark.parametrize("accumulation", ACCUMULATIONS)
def test_ljpme_grid_rules_mixed_types(mixed_lb, case, vdw, accumulation):
    """LJ-PME with the Lorentz-Berthelot grid rule (VdwType::EwaldLB,
    nbnxm_gpu_data_mgmt.cpp:335-346; calculate_lj_ewald_comb_LB_F_E,
    nbnxm_cuda_kernel_utils.cuh:290-337) and the geometric one on a system
    with 8 LJ types, F and F+E, against the FP64 oracle: FP32 modes within
    5e-6 force relative RMS and 2e-5 in the energies, the FP64 mode to 1e-10."""
    announce(f"ljpme grid rule {case} vdw={vdw} accumulation={accumulation}")
    s = mixed_lb
    search, plist, atoms = nb.build(s, 1.1, vdw)
    ic = nb.interaction_const_ex(1.0, ewald_rtol=1e-5, **LJPME_LB_CASES[case])
    f64, fs64, elj64, eel64 = oracle.kernel(plist, atoms, ic, True, "f64")
    real = atoms.atom_indices >= 0
    task, f, elj, eel, fshift = _run(ic, atoms, plist, 1.1, accumulation)
    if accumulation == nb.ACCUM_FP64:
        assert rel_rms(f[real], f64[real]) < 1e-10
        assert elj == pytest.approx(elj64, rel=1e-10)
        assert eel == pytest.approx(eel64, rel=1e-10)
    else:
        assert rel_rms(f[real], f64[real]) < 5e-6
        assert elj == pytest.approx(elj64, rel=2e-5)
        assert eel == pytest.approx(eel64, rel=2e-5)
    assert np.abs(fshift - fs64).max() <= 1e-4 * max(1.0, np.abs(fs64).max())
    # Force-only pass equals the F+E forces.
    task.clear_outputs()
    task.launch(compute_energy=False, compute_virial=False)
    f2, fd2, _, _, _ = task.fetch()
    f2 = (fd2 if fd2 is not None else f2).reshape(-1, 3)
    assert rel_rms(f2[real], f[real]) < 1e-6


def test_flavour_is_kernel_ewald(water3k):
    s, plist, atoms = water3k
    for coulomb, expect in (("ewald", 1), ("rf", 0)):
        task = nb.GpuTask(nb.interaction_const_ex(1.0, coulomb=coulomb), atoms, 1.1)
        assert task._lib.nbnxm_gpu_is_kernel_ewald_analytical(task._h) == expect
