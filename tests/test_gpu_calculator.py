"""GPU: the nblib-style force calculator, the FP64 virial reduction and the
heterogeneous chain system, against the CPU oracle and the golden vectors.
Calls go through the C ABI."""
import numpy as np
import pytest

import nbnxm_amd as nb
import oracle
from conftest import announce
from helpers import rel_rms

pytestmark = pytest.mark.gpu

REF_FORCE_TOL = 0.0119  # kernel_test.cpp:663-698 in FP32


def nb_params(system):
    """(6*C6, 12*C12) per type pair, as nblib's createNonBondedParameters."""
    return np.stack([6.0 * system.c6c12[..., 0], 12.0 * system.c6c12[..., 1]], axis=-1)


@pytest.fixture(scope="module")
def water9k():
    return nb.systems.config_water_small()


@pytest.fixture(scope="module")
def water9k_oracle(water9k):
    ic = nb.interaction_const(1.0, 1e-5)
    search, plist, atoms = nb.build(water9k, 1.1, nb.VDW_CUT_COMB_GEOM)
    f, fs, elj, eel = oracle.kernel(plist, atoms, ic, True, "f64")
    real = atoms.atom_indices >= 0
    xq = atoms.xq.reshape(-1, 4)[:, :3].astype(np.float64)
    vir = oracle.single_sum_virial(xq[real], f[real], atoms.shift_vec, fs)
    # nblib's calculator has no Coulomb potential shift (nbnxmsetuphelpers.cpp:262-270).
    ic_nblib = nb.interaction_const(1.0, 1e-5)
    ic_nblib.sh_ewald = 0.0
    _, _, _, eel_nblib = oracle.kernel(plist, atoms, ic_nblib, True, "f64")
    return dict(f=atoms.to_original_order(f, water9k.n), elj=elj, eel=eel, eel_nblib=eel_nblib, vir=vir,
                atoms=atoms, plist=plist, ic=ic)



def test_second_pairlist_update_reuses_split(water9k):
    """A second update_pairlist builds its list once, with the entry split of
    the first list: forces on new coordinates equal those of a calculator
    whose first list is built on them, up to the FP32 summation order."""
    announce("calculator second pair-list update")
    opt = nb.NBKernelOptions(pairlistCutoff=1.0, rlist=1.1)
    args = (water9k.types, nb_params(water9k), water9k.q, water9k.excl_ranges, water9k.excl_elements, opt)
    calc = nb.ForceCalculator(*args)
    calc.update_pairlist(water9k.x.copy(), water9k.box)
    rng = np.random.default_rng(31)
    x_new = water9k.x + rng.normal(0.0, 0.01, size=water9k.x.shape)
    x2 = calc.update_pairlist(x_new.copy(), water9k.box)
    f2, _, en2 = calc.compute(x2, water9k.box, energies=True)
    fresh = nb.ForceCalculator(*args)
    x3 = fresh.update_pairlist(x_new.copy(), water9k.box)
    f3, _, en3 = fresh.compute(x3, water9k.box, energies=True)
    assert np.array_equal(x2, x3)
    assert rel_rms(f2, f3) <= 1e-6
    assert en2[:2] == pytest.approx(en3[:2], rel=1e-6)

def test_calculator_water9k(water9k, water9k_oracle):
    announce("calculator water9k fixed64")
    opt = nb.NBKernelOptions(pairlistCutoff=1.0, rlist=1.1)
    calc = nb.ForceCalculator(water9k.types, nb_params(water9k), water9k.q, water9k.excl_ranges,
                              water9k.excl_elements, opt)
    x = calc.update_pairlist(water9k.x.copy(), water9k.box)
    f, vir, en = calc.compute(x, water9k.box, virial=True, energies=True)
    ref = water9k_oracle
    assert rel_rms(f, ref["f"]) <= 5e-6
    assert en[1] == pytest.approx(ref["elj"], rel=1e-5)
    assert en[0] == pytest.approx(ref["eel_nblib"], rel=1e-5)
    assert en[2:].tolist() == [0.0, 0.0, 0.0]
    assert np.abs(vir - ref["vir"]).max() <= 1e-5 * np.abs(ref["vir"]).max()
    # Forces are added to the output buffer (atomdata_add_nbat_f_to_f); the
    # same kernel flavour (F+E) gives bitwise the same fixed-point forces.
    f2, vir2, en2 = calc.compute(x, water9k.box, forces=f.copy(), virial=True, energies=True)
    assert np.array_equal(f2, 2.0 * f)
    assert np.array_equal(vir2, vir) and np.array_equal(en2, en)
    # The force-only flavour agrees within FP32 rounding.
    f3, _, _ = calc.compute(x, water9k.box)
    assert rel_rms(f3, f) <= 1e-6


def test_calculator_fp64_mode(water9k, water9k_oracle):
    announce("calculator water9k fp64")
    opt = nb.NBKernelOptions(pairlistCutoff=1.0, rlist=1.1, accumulation=nb.ACCUM_FP64)
    calc = nb.ForceCalculator(water9k.types, nb_params(water9k), water9k.q, water9k.excl_ranges,
                              water9k.excl_elements, opt)
    x = calc.update_pairlist(water9k.x.copy(), water9k.box)
    f, vir, en = calc.compute(x, water9k.box, virial=True, energies=True)
    ref = water9k_oracle
    # FP64 on both sides; the kernel reads FP32 coordinates and parameters like the oracle.
    assert rel_rms(f, ref["f"]) <= 1e-9
    assert en[0] == pytest.approx(ref["eel_nblib"], rel=1e-9)
    assert np.abs(vir - ref["vir"]).max() <= 1e-8 * np.abs(ref["vir"]).max()


@pytest.mark.parametrize("rule", ["geom", "lb"])
def test_calculator_golden(golden, rule):
    """The reference kernel test through the calculator: the combination rule is
    detected from the type table, as enbnxninitcombruleDETECT does."""
    announce(f"calculator golden {rule}")
    system = nb.systems.spc27_system(golden, rule)
    case = {"geom": "CoulombEwald_VdwCutCombGeom", "lb": "CoulombEwald_VdwCutCombLB"}[rule]
    assert nb.detect_comb_rule(nb_params(system)) == (nb.VDW_CUT_COMB_GEOM if rule == "geom" else nb.VDW_CUT_COMB_LB)
    opt = nb.NBKernelOptions(pairlistCutoff=0.9, ewaldRtol=1e-6)
    calc = nb.ForceCalculator(system.types, nb_params(system), system.q, system.excl_ranges,
                              system.excl_elements, opt)
    x = calc.update_pairlist(system.x.copy(), system.box)
    f, _, en = calc.compute(x, system.box, energies=True)
    ref = np.array(golden["cases"][case]["forces"])
    assert np.abs(f - ref).max() < REF_FORCE_TOL
    assert abs(en[1] - golden["cases"][case]["vdw_energy"]) < REF_FORCE_TOL
    # The refdata carry the Ewald potential shift, which nblib leaves out: the
    # difference is sh_ewald * sum of qq over the pairs within the cut-off.
    assert abs(en[0] - golden["cases"][case]["coulomb_energy"]) < 10 * REF_FORCE_TOL


def test_calculator_tabulated_ewald(water9k, water9k_oracle):
    """useTabulatedEwaldCorr: the tabulated correction of the same Ewald
    interaction agrees with the analytical FP64 oracle to the table accuracy."""
    announce("calculator water9k tabulated Ewald")
    opt = nb.NBKernelOptions(pairlistCutoff=1.0, rlist=1.1, useTabulatedEwaldCorr=True)
    calc = nb.ForceCalculator(water9k.types, nb_params(water9k), water9k.q, water9k.excl_ranges,
                              water9k.excl_elements, opt)
    x = calc.update_pairlist(water9k.x.copy(), water9k.box)
    f, vir, en = calc.compute(x, water9k.box, virial=True, energies=True)
    ref = water9k_oracle
    assert rel_rms(f, ref["f"]) <= 5e-6
    assert en[0] == pytest.approx(ref["eel_nblib"], rel=1e-5)


@pytest.mark.parametrize("coulomb", ["cutoff", "rf"])
def test_calculator_cutoff_and_rf(water9k, coulomb):
    """nblib's Cutoff and ReactionField types: both run the reaction-field
    kernel with k_rf = 0, c_rf = 1/rc (calc_rffac with epsilon_rf = 1)."""
    announce(f"calculator water9k {coulomb}")
    ct = {"cutoff": nb._lib.COULOMB_CUTOFF, "rf": nb._lib.COULOMB_REACTION_FIELD}[coulomb]
    opt = nb.NBKernelOptions(pairlistCutoff=1.0, rlist=1.1, coulombType=ct)
    calc = nb.ForceCalculator(water9k.types, nb_params(water9k), water9k.q, water9k.excl_ranges,
                              water9k.excl_elements, opt)
    x = calc.update_pairlist(water9k.x.copy(), water9k.box)
    f, vir, en = calc.compute(x, water9k.box, virial=True, energies=True)
    ic = nb.interaction_const_ex(1.0, coulomb="rf", epsilon_rf=1.0)
    assert ic.k_rf == 0.0 and ic.c_rf == pytest.approx(1.0)
    search, plist, atoms = nb.build(water9k, 1.1, nb.VDW_CUT_COMB_GEOM)
    f64, fs64, elj64, eel64 = oracle.kernel(plist, atoms, ic, True, "f64")
    assert rel_rms(f, atoms.to_original_order(f64, water9k.n)) <= 5e-6
    assert en[0] == pytest.approx(eel64, rel=2e-5)
    assert en[1] == pytest.approx(elj64, rel=1e-5)


def test_gpu_task_virial(water9k, water9k_oracle):
    announce("gpu task virial water9k")
    ref = water9k_oracle
    for acc, tol in [(nb.ACCUM_FIXED64, 1e-5), (nb.ACCUM_FLOAT, 1e-5), (nb.ACCUM_FP64, 1e-8)]:
        task = nb.GpuTask(ref["ic"], ref["atoms"], 1.1, None, acc)
        task.init_pairlist(ref["plist"])
        task.launch(compute_energy=False, compute_virial=True)
        task.fetch()
        vir = task.virial()
        assert np.abs(vir - ref["vir"]).max() <= tol * np.abs(ref["vir"]).max(), acc
    # Without computeVirial in the last launch there is no virial to return.
    task.launch(compute_energy=True, compute_virial=False)
    task.fetch()
    with pytest.raises(nb.NbnxmError, match="virial"):
        task.virial()


@pytest.fixture(scope="module")
def hetero_small():
    return nb.systems.hetero_system(12, 50, 6000, 20, seed=12)


@pytest.mark.parametrize("vdw", [nb.VDW_CUT_TYPE_TABLE, nb.VDW_CUT_COMB_LB])
def test_hetero_chains(hetero_small, vdw):
    """20 LB types from the LDS-staged type table (or the LB rule), 1-2/1-3
    exclusions inside the chains."""
    announce(f"hetero chains vdw={vdw}")
    s = hetero_small
    ic = nb.interaction_const(1.0, 1e-5)
    search, plist, atoms = nb.build(s, 1.1, vdw)
    f64, fs64, elj64, eel64 = oracle.kernel(plist, atoms, ic, True, "f64")
    f32, _, _, _ = oracle.kernel(plist, atoms, ic, True, "f32")
    task = nb.GpuTask(ic, atoms, 1.1, None, nb.ACCUM_FIXED64)
    task.init_pairlist(plist)
    task.launch(compute_energy=True, compute_virial=True)
    f, fd, elj, eel, fshift = task.fetch()
    real = atoms.atom_indices >= 0
    gpu = fd.reshape(-1, 3)[real]
    assert rel_rms(gpu, f64[real]) <= 5e-6
    assert rel_rms(gpu, f32[real]) <= 1e-5
    assert elj == pytest.approx(elj64, rel=1e-5)
    assert eel == pytest.approx(eel64, rel=1e-5)
    assert np.abs(fshift.reshape(-1, 3) - fs64).max() <= 1e-5 * np.abs(fs64).max()
