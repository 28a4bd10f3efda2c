"""GPU parity of the nonbonded cluster-pair kernel against the CPU oracle and
the reference's golden vectors.  Calls go through the C ABI."""
import numpy as np
import pytest

import nbnxm_amd as nb
import oracle
from conftest import announce
from helpers import rel_rms

pytestmark = pytest.mark.gpu

GOLDEN_CASES = [
    ("CoulombEwald_VdwCutCombNone", "none", nb.VDW_CUT_TYPE_TABLE),
    ("CoulombEwald_VdwCutCombGeom", "geom", nb.VDW_CUT_COMB_GEOM),
    ("CoulombEwald_VdwCutCombLB", "lb", nb.VDW_CUT_COMB_LB),
]
# kernel_test.cpp:663-698: max(1000*2^-22*50, 20*epsfac*1e-6*(0.8476/rc)^2) in FP32
REF_FORCE_TOL = 0.0119


def run_gpu(system, rc, rlist, rtol, vdw, accumulation=nb.ACCUM_FIXED64, energy=True, rlist_inner=None):
    ic = nb.interaction_const(rc, rtol)
    search, plist, atoms = nb.build(system, rlist, vdw)
    task = nb.GpuTask(ic, atoms, rlist, rlist_inner, accumulation)
    task.init_pairlist(plist)
    task.launch(compute_energy=energy, compute_virial=True)
    f, fd, elj, eel, fshift = task.fetch()
    return dict(ic=ic, search=search, plist=plist, atoms=atoms, task=task, f=f, fd=fd, elj=elj, eel=eel, fshift=fshift)


@pytest.mark.parametrize("case,rule,vdw", GOLDEN_CASES)
@pytest.mark.parametrize("accumulation", [nb.ACCUM_FIXED64, nb.ACCUM_FLOAT, nb.ACCUM_FP64])
def test_golden_spc27(golden, case, rule, vdw, accumulation):
    announce(f"golden {case} accumulation={accumulation}")
    system = nb.systems.spc27_system(golden, rule)
    r = run_gpu(system, 0.9, 0.9, 1e-6, vdw, accumulation)
    ref = np.array(golden["cases"][case]["forces"])
    f = r["atoms"].to_original_order(r["f"], system.n)
    assert np.abs(f - ref).max() < REF_FORCE_TOL
    assert abs(r["elj"] - golden["cases"][case]["vdw_energy"]) < REF_FORCE_TOL
    assert abs(r["eel"] - golden["cases"][case]["coulomb_energy"]) < 10 * REF_FORCE_TOL


@pytest.fixture(scope="module")
def water9k():
    return nb.systems.config_water_small()


@pytest.mark.parametrize("vdw", [nb.VDW_CUT_COMB_GEOM, nb.VDW_CUT_TYPE_TABLE])
def test_water9k_fixed_vs_fp32_oracle(water9k, vdw):
    announce(f"water9k fixed64 vdw={vdw}")
    r = run_gpu(water9k, 1.0, 1.1, 1e-5, vdw)
    f32, fs32, elj32, eel32 = oracle.kernel(r["plist"], r["atoms"], r["ic"], True, "f32")
    f64, fs64, elj64, eel64 = oracle.kernel(r["plist"], r["atoms"], r["ic"], True, "f64")
    real = r["atoms"].atom_indices >= 0
    gpu = r["fd"].reshape(-1, 3)[real]
    assert rel_rms(gpu, f32[real]) <= 1e-5
    assert rel_rms(gpu, f64[real]) <= 5e-6
    assert abs(r["elj"] - elj64) <= 1e-5 * abs(elj64)
    assert abs(r["eel"] - eel64) <= 1e-5 * abs(eel64)
    fshift = r["fshift"].reshape(-1, 3)
    assert np.abs(fshift - fs64).max() <= 1e-4 * max(1.0, np.abs(fs64).max())


def test_water9k_fp64_mode_matches_oracle(water9k):
    announce("water9k fp64 mode")
    r = run_gpu(water9k, 1.0, 1.1, 1e-5, nb.VDW_CUT_COMB_GEOM, nb.ACCUM_FP64)
    f64, fs64, elj64, eel64 = oracle.kernel(r["plist"], r["atoms"], r["ic"], True, "f64")
    real = r["atoms"].atom_indices >= 0
    assert rel_rms(r["fd"].reshape(-1, 3)[real], f64[real]) <= 1e-10
    assert abs(r["elj"] - elj64) <= 1e-10 * abs(elj64)
    assert abs(r["eel"] - eel64) <= 1e-10 * abs(eel64)


def test_fixed_point_bitwise_reproducible(water9k):
    announce("water9k bitwise reproducibility")
    r = run_gpu(water9k, 1.0, 1.1, 1e-5, nb.VDW_CUT_COMB_GEOM)
    task = r["task"]
    for _ in range(2):
        task.launch(compute_energy=True, compute_virial=True)
        f, fd, elj, eel, fshift = task.fetch()
        assert np.array_equal(fd.view(np.uint64), r["fd"].view(np.uint64))
        assert elj == r["elj"] and eel == r["eel"]
        assert np.array_equal(fshift, r["fshift"])


def test_force_only_matches_energy_kernel(water9k):
    announce("water9k F vs VF")
    r = run_gpu(water9k, 1.0, 1.1, 1e-5, nb.VDW_CUT_COMB_GEOM)
    task = r["task"]
    task.launch(compute_energy=False, compute_virial=False)
    f, fd, elj, eel, fshift = task.fetch()
    real = r["atoms"].atom_indices >= 0
    assert rel_rms(fd.reshape(-1, 3)[real], r["fd"].reshape(-1, 3)[real]) <= 1e-6


def test_float_accumulation_mode(water9k):
    announce("water9k float accumulation")
    r = run_gpu(water9k, 1.0, 1.1, 1e-5, nb.VDW_CUT_COMB_GEOM, nb.ACCUM_FLOAT)
    f32, _, _, _ = oracle.kernel(r["plist"], r["atoms"], r["ic"], True, "f32")
    real = r["atoms"].atom_indices >= 0
    assert rel_rms(r["f"].reshape(-1, 3)[real], f32[real]) <= 1e-5


def test_newton_third_law(water9k):
    announce("water9k net force")
    r = run_gpu(water9k, 1.0, 1.1, 1e-5, nb.VDW_CUT_COMB_GEOM, nb.ACCUM_FP64)
    real = r["atoms"].atom_indices >= 0
    f = r["fd"].reshape(-1, 3)[real]
    # Pair forces cancel, so the total force is zero up to fixed-point rounding.
    assert np.abs(f.sum(axis=0)).max() <= 1e-6 * np.abs(f).max()


@pytest.mark.parametrize("vdw", [nb.VDW_CUT_COMB_GEOM, nb.VDW_CUT_TYPE_TABLE, nb.VDW_CUT_COMB_LB])
@pytest.mark.parametrize("energy", [True, False])
def test_i_cluster_placement_bit_identical(water9k, vdw, energy):
    """The tuned kernel with the i-cluster coordinates and charges in LDS (4
    waves per SIMD) and in registers (3 waves per SIMD): the same arithmetic
    in the same order, so the fixed-point forces, energies and shift forces
    agree bit for bit; both match the FP32 oracle."""
    announce(f"i-cluster placement vdw={vdw} energy={energy}")
    ic = nb.interaction_const(1.0, 1e-5)
    _, plist, atoms = nb.build(water9k, 1.1, vdw)
    out = []
    for mode in (0, 1):
        task = nb.GpuTask(ic, atoms, 1.1, None, nb.ACCUM_FIXED64)
        task.set_i_cluster_in_lds(mode)
        task.init_pairlist(plist)
        task.launch(compute_energy=energy, compute_virial=True)
        out.append(task.fetch())
    (_, fd0, elj0, eel0, fs0), (_, fd1, elj1, eel1, fs1) = out
    assert np.array_equal(fd0, fd1)
    assert np.array_equal(fs0, fs1)
    assert elj0 == elj1 and eel0 == eel1
    f32, _, _, _ = oracle.kernel(plist, atoms, ic, energy, "f32")
    real = atoms.atom_indices >= 0
    assert rel_rms(fd1.reshape(-1, 3)[real], f32[real]) <= 1e-5
