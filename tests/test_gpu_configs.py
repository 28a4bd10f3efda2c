"""GPU parity at the full sizes of the benchmark configurations: the 96k-atom
water box (with and without LJ) over 100 perturbed evaluations, the 1.0 M-atom
solvated chain system and the 3.0 M-atom precision study.  The reference at
these sizes is the FP64 GPU mode, itself pinned to the CPU oracle on smaller
systems (test_gpu_kernel.py); the properties checked are size-independent:
error bounds, bitwise reproducibility, Newton's third law."""
import numpy as np
import pytest

import nbnxm_amd as nb
from conftest import announce
from helpers import rel_rms

pytestmark = pytest.mark.gpu

SIGMA = 0.002  # nm, per-coordinate perturbation


def nb_params(system):
    return np.stack([6.0 * system.c6c12[..., 0], 12.0 * system.c6c12[..., 1]], axis=-1)


def calculator(system, accumulation):
    opt = nb.NBKernelOptions(pairlistCutoff=1.0, rlist=1.1, accumulation=accumulation)
    return nb.ForceCalculator(system.types, nb_params(system), system.q, system.excl_ranges, system.excl_elements, opt)


@pytest.mark.parametrize("lj", [True, False], ids=["lj", "coulomb_only"])
def test_water96k_perturbed_evaluations(lj):
    """100 evaluations at x0 + N(0, 0.002 nm) on one list built at x0 with a
    0.1 nm buffer: fixed-point FP32 against FP64 at every evaluation."""
    announce(f"water96k lj={lj}: 100 perturbed evaluations")
    s = nb.systems.config_water_96k(lj=lj)
    fixed = calculator(s, nb.ACCUM_FIXED64)
    ref = calculator(s, nb.ACCUM_FP64)
    x0 = fixed.update_pairlist(s.x.copy(), s.box)
    ref.update_pairlist(x0.copy(), s.box)
    rng = np.random.default_rng(31)
    worst = 0.0
    for k in range(100):
        x = x0 + rng.normal(0.0, SIGMA, x0.shape)
        f, vir, en = fixed.compute(x, s.box, virial=True, energies=True)
        fr, virr, enr = ref.compute(x, s.box, virial=True, energies=True)
        err = rel_rms(f, fr)
        worst = max(worst, err)
        assert err <= 5e-6, (k, err)
        assert en[0] == pytest.approx(enr[0], rel=2e-5)
        if lj:
            assert en[1] == pytest.approx(enr[1], rel=2e-5)
        else:
            assert en[1] == 0.0 and enr[1] == 0.0
        # The single-sum virial -1/2 sum x (x) f inherits the force error:
        # |d vir| <= 1/2 |x| |d f| (Cauchy-Schwarz), with a margin for the
        # shift-force term.
        assert np.abs(vir - virr).max() <= 1.5 * np.linalg.norm(x) * np.linalg.norm(f - fr) + 1e-9 * np.abs(virr).max()
        # Net force: pair forces cancel up to the FP32 rounding of each atom's
        # summed force, a random walk over atoms; the FP64 mode cancels exactly.
        assert np.abs(f.sum(axis=0)).max() <= 1e-6 * np.sqrt((f ** 2).sum())
        assert np.abs(fr.sum(axis=0)).max() <= 1e-9 * np.sqrt((fr ** 2).sum())
    announce(f"water96k lj={lj}: worst force rel RMS vs FP64 {worst:.3e}")


def _task_forces(ic, atoms, plist, accumulation):
    task = nb.GpuTask(ic, atoms, 1.1, None, accumulation)
    task.init_pairlist(plist)
    task.launch(compute_energy=True, compute_virial=True)
    f, fd, elj, eel, fshift = task.fetch()
    return task, (fd if fd is not None else f.astype(np.float64)), elj, eel


def test_hetero_1m():
    """1.0 M atoms: 480 chains x 50 atoms with 20 LB types and 1-2/1-3
    exclusions in water, type-table kernel, fixed point against FP64."""
    announce("hetero1M: build")
    s = nb.systems.config_hetero_1m()
    ic = nb.interaction_const(1.0, 1e-5)
    search, plist, atoms = nb.build(s, 1.1, nb.VDW_CUT_TYPE_TABLE)
    real = atoms.atom_indices >= 0
    announce("hetero1M: fixed64 and fp64 launches")
    task, f, elj, eel = _task_forces(ic, atoms, plist, nb.ACCUM_FIXED64)
    task.launch(compute_energy=True, compute_virial=True)
    _, f_again, elj_again, eel_again, _ = task.fetch()
    _, fr, eljr, eelr = _task_forces(ic, atoms, plist, nb.ACCUM_FP64)
    err = rel_rms(f.reshape(-1, 3)[real], fr.reshape(-1, 3)[real])
    announce(f"hetero1M: force rel RMS vs FP64 {err:.3e}")
    assert err <= 5e-6
    assert np.array_equal(f, f_again) and elj == elj_again and eel == eel_again
    assert elj == pytest.approx(eljr, rel=1e-5)
    assert eel == pytest.approx(eelr, rel=1e-5)


def test_water3m_precision_study():
    """3.0 M atoms, three modes: FP32 with float atomics, FP32 with int64
    fixed point + FP64 reductions, and FP64.  Fixed point is bitwise
    reproducible; both FP32 modes stay within 5e-6 of FP64."""
    announce("water3M: build")
    s = nb.systems.config_water_3m()
    ic = nb.interaction_const(1.0, 1e-5)
    search, plist, atoms = nb.build(s, 1.1, nb.VDW_CUT_COMB_GEOM)
    real = atoms.atom_indices >= 0
    announce("water3M: fp64")
    _, fr, eljr, eelr = _task_forces(ic, atoms, plist, nb.ACCUM_FP64)
    fr = fr.reshape(-1, 3)[real]
    for acc in (nb.ACCUM_FIXED64, nb.ACCUM_FLOAT):
        announce(f"water3M: accumulation {acc}")
        task, f, elj, eel = _task_forces(ic, atoms, plist, acc)
        task.clear_outputs()
        task.launch(compute_energy=True, compute_virial=True)
        f2, fd2, elj2, eel2, _ = task.fetch()
        f2 = fd2 if fd2 is not None else f2.astype(np.float64)
        err = rel_rms(f.reshape(-1, 3)[real], fr)
        announce(f"water3M: accumulation {acc}: force rel RMS {err:.3e}, "
                 f"E_el rel {abs(eel - eelr) / abs(eelr):.3e}, E_lj rel {abs(elj - eljr) / abs(eljr):.3e}, "
                 f"bitwise repeat {np.array_equal(f, f2)}")
        assert err <= 5e-6
        assert eel == pytest.approx(eelr, rel=2e-5)
        assert elj == pytest.approx(eljr, rel=2e-5)
        if acc == nb.ACCUM_FIXED64:
            assert np.array_equal(f, f2) and eel == eel2 and elj == elj2
