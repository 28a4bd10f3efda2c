"""Shared helpers of the test suite (no GPU use at import time)."""
import numpy as np


def rel_rms(f, ref):
    f = np.asarray(f, dtype=np.float64).reshape(-1, 3)
    ref = np.asarray(ref, dtype=np.float64).reshape(-1, 3)
    return float(np.sqrt(np.sum((f - ref) ** 2) / np.sum(ref ** 2)))


def real_mask(atoms):
    return np.repeat(atoms.atom_indices >= 0, 1)


def wrapped_float_system(system):
    """Copy of a system whose coordinates are the FP32-rounded wrapped ones the
    kernels see, so brute-force FP64 sums and list kernels share inputs."""
    import copy
    s = copy.copy(system)
    x = np.asarray(system.x, dtype=np.float64)
    box = np.asarray(system.box, dtype=np.float32).astype(np.float64)
    s.box = box
    xw = x - np.floor(x / box) * box
    xw = xw.astype(np.float32)
    xw[xw >= box.astype(np.float32)] = 0.0
    s.x = xw.astype(np.float64)
    s.q = np.asarray(system.q, dtype=np.float32).astype(np.float64)
    # LJ parameters as the FP32 type table holds them (6*C6, 12*C12 in FP32).
    c = np.array(system.c6c12, dtype=np.float64)
    c[..., 0] = (6.0 * c[..., 0]).astype(np.float32).astype(np.float64) / 6.0
    c[..., 1] = (12.0 * c[..., 1]).astype(np.float32).astype(np.float64) / 12.0
    s.c6c12 = c
    return s


# The 18 interaction flavours of the reference kernel test
# (src/gromacs/nbnxm/tests/kernel_test.cpp:402-447, 577-611).
GOLDEN_COULOMB = ["Ewald", "EwaldTwin", "ReactionField"]
GOLDEN_VDW = ["CutCombGeom", "CutCombLB", "CutCombNone", "ForceSwitch", "PotSwitch", "EwaldCombGeom"]
GOLDEN_ALL = [f"Coulomb{c}_Vdw{v}" for c in GOLDEN_COULOMB for v in GOLDEN_VDW]
# kernel_test.cpp:663-698 in FP32: 1000 * 2^-22 * 50 ulp; LJ-PME 400 ulp.
REF_FORCE_TOL = 1000.0 * 0.5 ** 22 * 50
REF_FORCE_TOL_LJPME = 1000.0 * 0.5 ** 22 * 400


def golden_case(golden, case):
    """(system, ic, list vdw type, FP32 reference tolerance) of one refdata case."""
    import nbnxm_amd as nb
    coulomb, vdw = case[len("Coulomb"):].split("_Vdw")
    opt = golden["options"]
    rc = opt["rc"]
    rvdw = opt["rvdw_twin"] if coulomb == "EwaldTwin" else rc
    modifier = {"ForceSwitch": "forceswitch", "PotSwitch": "potswitch"}.get(vdw, "potshift")
    ic = nb.interaction_const_ex(rc, coulomb="rf" if coulomb == "ReactionField" else "ewald",
                                 vdw_modifier=modifier, lj_pme=(vdw == "EwaldCombGeom"), rvdw=rvdw,
                                 rvdw_switch=rvdw - opt["switch_width"], ewald_rtol=opt["ewald_rtol"],
                                 ewald_rtol_lj=opt["ewald_rtol_lj"], epsilon_r=opt["epsilon_r"],
                                 epsilon_rf=opt["epsilon_rf"])
    system = nb.systems.spc27_system(golden, "geom" if vdw == "CutCombGeom" else "lb")
    vdw_type = {"CutCombGeom": nb.VDW_CUT_COMB_GEOM, "CutCombLB": nb.VDW_CUT_COMB_LB}.get(vdw, nb.VDW_CUT_TYPE_TABLE)
    tol = REF_FORCE_TOL_LJPME if vdw == "EwaldCombGeom" else REF_FORCE_TOL
    return system, ic, vdw_type, tol
