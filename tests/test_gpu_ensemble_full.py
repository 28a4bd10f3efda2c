"""The benchmark workload at full size: 256 water boxes of 648-24,000 atoms
(1.9 M atoms, seed 4), device pair search, int64 fixed-point accumulation,
several F+E evaluations with N(0, 0.002 nm) perturbations captured in one
hipGraph.  Per box: forces within 5e-6 relative RMS of the FP64 mode at the
same positions, batched equal to the plain per-system launches bit for bit;
the FP64 mode equals the CPU brute force on the 16 smallest boxes."""
import numpy as np
import pytest

import nbnxm_amd as nb
import oracle
from conftest import announce
from helpers import rel_rms, wrapped_float_system

pytestmark = pytest.mark.gpu

RC, RTOL, RLIST_OUTER, RLIST_INNER, SIGMA, SEED = 1.0, 1e-5, 1.1, 1.04, 0.002, 11
STEPS = 4


@pytest.fixture(scope="module")
def systems():
    return nb.systems.config_ensemble(256)


@pytest.fixture(scope="module")
def batched(systems):
    """Forces, energies and positions of the last of STEPS graph-captured
    perturbed F+E evaluations (steps 0..STEPS-1), with a bucketed entry split."""
    announce("full ensemble: create (device search) and graph run")
    ens = nb.Ensemble(systems, RC, RTOL, RLIST_OUTER, RLIST_INNER, nb.VDW_CUT_COMB_GEOM, nb.ACCUM_FIXED64, gpu_search=True)
    ens.set_split([1500, 3500, 8000], [256, 0, 0, 512], 2)
    ens.set_step(0)
    ens.run(STEPS, nb.MODE_BATCHED, SIGMA, SEED, True, True, STEPS)
    ens.sync()
    forces = [ens.forces(k) for k in range(len(systems))]
    positions = [ens.positions(k) for k in range(len(systems))]
    energies, fshift = ens.energies()
    return ens, forces, positions, energies, fshift


def test_full_ensemble_vs_fp64_mode(systems, batched):
    ens, forces, positions, energies, _ = batched
    announce("full ensemble: FP64 mode at the same perturbed positions")
    ref = nb.Ensemble(systems, RC, RTOL, RLIST_OUTER, RLIST_INNER, nb.VDW_CUT_COMB_GEOM, nb.ACCUM_FP64)
    ref.set_step(STEPS - 1)
    ref.run(1, nb.MODE_BATCHED, SIGMA, SEED, True)
    ref.sync()
    e64, _ = ref.energies()
    worst = 0.0
    for k in range(len(systems)):
        assert np.array_equal(ref.positions(k), positions[k]), k
        err = rel_rms(forces[k], ref.forces(k))
        worst = max(worst, err)
        assert err <= 5e-6, (k, err)
        # A box's Coulomb total is a small difference of large terms (self
        # energy against the pair sum): FP32 pair sums carry ~1e-5 of it.
        n = systems[k].n
        assert abs(energies[k, 1] - e64[k, 1]) <= 5e-5 * abs(e64[k, 1]), k
        assert abs(energies[k, 0] - e64[k, 0]) <= 1e-5 * n, k
    print(f"[full ensemble] worst per-box force rel RMS vs FP64 mode: {worst:.3e}", flush=True)


def test_full_ensemble_batched_equals_plain(systems, batched):
    ens, forces, positions, energies, fshift = batched
    announce("full ensemble: plain per-system launches at the last step")
    ens.set_step(STEPS - 1)
    ens.run(1, nb.MODE_PLAIN, SIGMA, SEED, True)
    ens.sync()
    for k in range(len(systems)):
        assert np.array_equal(ens.forces(k), forces[k]), k
    ep, fsp = ens.energies()
    assert np.array_equal(ep, energies) and np.array_equal(fsp, fshift)


def test_full_ensemble_fp64_mode_vs_bruteforce(systems):
    """The FP64 mode, the reference the two tests above compare with, against
    the CPU FP64 brute force on the 16 smallest boxes (several periodic images
    of a pair lie within the cut-off in the smallest), on the FP32-rounded
    inputs both see (type-table LJ, so the FP32 table holds the same values)."""
    announce("full ensemble: FP64 mode vs CPU brute force, 16 smallest boxes")
    order = np.argsort([s.n for s in systems], kind="stable")[:16]
    small = [wrapped_float_system(systems[k]) for k in order]
    ens64 = nb.Ensemble(small, RC, RTOL, RLIST_OUTER, RLIST_INNER, nb.VDW_CUT_TYPE_TABLE, nb.ACCUM_FP64)
    ens64.run(1, nb.MODE_BATCHED, 0.0, SEED, True)
    ens64.sync()
    e, _ = ens64.energies()
    ic = nb.interaction_const(RC, RTOL)
    for k, s in enumerate(small):
        f_ref, elj, eel, _ = oracle.bruteforce(s, ic)
        assert rel_rms(ens64.forces(k), f_ref) <= 1e-9, k
        assert abs(e[k, 0] - elj) <= 1e-9 * abs(elj)
        assert abs(e[k, 1] - eel) <= 1e-9 * abs(eel)
