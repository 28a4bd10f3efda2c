"""Batched ensemble on the GPU: per-system forces against the FP64 brute-force
sum, batched == plain bitwise, graph replays == eager bitwise, perturbed
positions stay covered by the pruned list."""
import copy

import numpy as np
import pytest

import nbnxm_amd as nb
import oracle
from conftest import announce
from helpers import rel_rms, wrapped_float_system

pytestmark = pytest.mark.gpu

RC, RTOL = 1.0, 1e-5


@pytest.fixture(scope="module")
def systems():
    # Smallest box (648 atoms, L = 1.86 nm < 2 rc: several images of a pair in range) up to 6000 atoms.
    sizes = [216, 300, 420, 1000, 2000]
    return [wrapped_float_system(nb.systems.water_box(m, None, seed=40 + k)) for k, m in enumerate(sizes)]


@pytest.fixture(scope="module")
def ens(systems):
    announce("ensemble create fixed64")
    # Type-table LJ: the FP32 table holds exactly the (rounded) parameters the
    # brute-force sum uses, so only FP32 arithmetic separates the two.
    return nb.Ensemble(systems, RC, RTOL, 1.1, 1.04, nb.VDW_CUT_TYPE_TABLE, nb.ACCUM_FIXED64, max_pairs_per_entry=48)


def _bruteforce(s):
    ic = nb.interaction_const(RC, RTOL)
    f, elj, eel, npairs = oracle.bruteforce(s, ic)
    return f, elj, eel, npairs


def test_ensemble_vs_bruteforce(systems, ens):
    announce("ensemble vs brute force, sigma 0")
    ens.run(1, nb.MODE_BATCHED, 0.0, 1, True)
    e, fshift = ens.energies()
    total, per = ens.count_pairs()
    for k, s in enumerate(systems):
        f_ref, elj, eel, npairs = _bruteforce(s)
        assert per[k] == npairs
        assert rel_rms(ens.forces(k), f_ref) <= 5e-6, k
        # FP32 pair energies sum to a small difference of large terms
        # (repulsion vs dispersion, self energy vs pair sum): ~1e-5 relative.
        assert abs(e[k, 0] - elj) <= 5e-5 * abs(elj)
        assert abs(e[k, 1] - eel) <= 5e-5 * abs(eel)


def test_fp64_mode_vs_bruteforce(systems):
    announce("ensemble fp64 mode vs brute force")
    ens64 = nb.Ensemble(systems, RC, RTOL, 1.1, 1.04, nb.VDW_CUT_TYPE_TABLE, nb.ACCUM_FP64)
    ens64.run(1, nb.MODE_BATCHED, 0.0, 1, True)
    e, _ = ens64.energies()
    for k, s in enumerate(systems):
        f_ref, elj, eel, _ = _bruteforce(s)
        # Same inputs, FP64 arithmetic on both sides; 2^-32 fixed-point force resolution.
        assert rel_rms(ens64.forces(k), f_ref) <= 1e-9, k
        assert abs(e[k, 0] - elj) <= 1e-9 * abs(elj)
        assert abs(e[k, 1] - eel) <= 1e-9 * abs(eel)


def test_batched_equals_plain_bitwise(systems, ens):
    announce("batched vs plain bitwise")
    ens.set_step(3)
    ens.run(1, nb.MODE_BATCHED, 0.002, 9, True)
    fb = [ens.forces(k) for k in range(len(systems))]
    eb, fsb = ens.energies()
    ens.set_step(3)
    ens.run(1, nb.MODE_PLAIN, 0.002, 9, True)
    for k in range(len(systems)):
        assert np.array_equal(ens.forces(k), fb[k])
    ep, fsp = ens.energies()
    assert np.array_equal(ep, eb) and np.array_equal(fsp, fsb)


@pytest.mark.parametrize("mode", [1, 0, -1])
def test_i_cluster_placement_bitwise(systems, ens, mode):
    """The batched launch with the i-cluster data in LDS (4 waves per SIMD)
    or in registers gives the per-system launches' results bit for bit."""
    announce(f"batched i-cluster placement {mode} vs plain bitwise")
    ens.set_i_cluster_in_lds(mode)
    # Automatic choice: LDS from 8192 work items.
    expect = mode if mode >= 0 else int(ens.stats()["work_items"] >= 8192)
    assert ens.launch_config()[1] == expect
    # The type-table F+E kernel keeps registers; the F kernel takes the setting.
    for energy in (True, False):
        ens.set_step(5)
        ens.run(1, nb.MODE_BATCHED, 0.002, 9, energy)
        fb = [ens.forces(k) for k in range(len(systems))]
        eb, fsb = ens.energies()
        ens.set_step(5)
        ens.run(1, nb.MODE_PLAIN, 0.002, 9, energy)
        for k in range(len(systems)):
            assert np.array_equal(ens.forces(k), fb[k]), (energy, k)
        if energy:
            ep, fsp = ens.energies()
            assert np.array_equal(ep, eb) and np.array_equal(fsp, fsb)


def test_graph_replays_equal_eager(systems, ens):
    announce("graph replay vs eager")
    ens.set_step(7)
    ens.run(1, nb.MODE_BATCHED, 0.002, 5, True, use_graph=False)
    ref = [ens.forces(k) for k in range(len(systems))]
    eref, _ = ens.energies()
    for rep in range(3):
        ens.set_step(7)
        ens.run(1, nb.MODE_BATCHED, 0.002, 5, True, use_graph=True, steps_per_graph=1)
        ens.sync()
        for k in range(len(systems)):
            assert np.array_equal(ens.forces(k), ref[k]), (rep, k)
        e, _ = ens.energies()
        assert np.array_equal(e, eref)
        # eager work between replays
        ens.run(1, nb.MODE_BATCHED, 0.0, 1, False)


@pytest.mark.parametrize("steps,use_graph,per_graph", [(2, False, 0), (3, False, 0), (5, True, 5), (4, True, 2)])
def test_multi_step_runs_equal_single_steps(systems, ens, steps, use_graph, per_graph):
    """A batched run of several fixed-point steps, eager or in graphs of
    `per_graph` steps, leaves the state of the same number of one-step runs
    bit for bit: forces, energies, shift forces and positions of the last
    step, and the step counter (the next run continues from it)."""
    announce(f"{steps} steps in one run (graph {use_graph}, {per_graph} per graph) vs one-step runs")
    ens.set_step(11)
    for _ in range(steps + 1):
        ens.run(1, nb.MODE_BATCHED, 0.002, 3, True)
    next_f = [ens.forces(k) for k in range(len(systems))]
    ens.set_step(11)
    for _ in range(steps):
        ens.run(1, nb.MODE_BATCHED, 0.002, 3, True)
    ref_f = [ens.forces(k) for k in range(len(systems))]
    ref_x = [ens.positions(k) for k in range(len(systems))]
    ref_e, ref_fs = ens.energies()
    for rep in range(2):
        ens.set_step(11)
        ens.run(steps, nb.MODE_BATCHED, 0.002, 3, True, use_graph=use_graph, steps_per_graph=per_graph)
        ens.sync()
        for k in range(len(systems)):
            assert np.array_equal(ens.forces(k), ref_f[k]), (rep, k)
            assert np.array_equal(ens.positions(k), ref_x[k]), (rep, k)
        e, fs = ens.energies()
        assert np.array_equal(e, ref_e) and np.array_equal(fs, ref_fs), rep
        # A one-step run afterwards continues from the step counter.
        ens.run(1, nb.MODE_BATCHED, 0.002, 3, True)
        for k in range(len(systems)):
            assert np.array_equal(ens.forces(k), next_f[k]), (rep, k)


def test_multi_step_runs_equal_single_steps_fp64(systems):
    """FP64 accumulation mode: three steps in one run equal three one-step
    runs bit for bit."""
    announce("3 steps in one run vs one-step runs, fp64 mode")
    ens64 = nb.Ensemble(systems[:3], RC, RTOL, 1.1, 1.04, nb.VDW_CUT_TYPE_TABLE, nb.ACCUM_FP64)
    ens64.set_step(5)
    for _ in range(3):
        ens64.run(1, nb.MODE_BATCHED, 0.002, 9, True)
    ref_f = [ens64.forces(k) for k in range(3)]
    ref_e, ref_fs = ens64.energies()
    ens64.set_step(5)
    ens64.run(3, nb.MODE_BATCHED, 0.002, 9, True)
    ens64.sync()
    for k in range(3):
        assert np.array_equal(ens64.forces(k), ref_f[k]), k
    e, fs = ens64.energies()
    assert np.array_equal(e, ref_e) and np.array_equal(fs, ref_fs)


def test_perturbed_positions_covered_by_list(systems, ens):
    announce("perturbed positions vs brute force")
    ens.set_step(123)
    ens.run(1, nb.MODE_BATCHED, 0.002, 77, True)
    e, _ = ens.energies()
    for k, s in enumerate(systems):
        x = ens.positions(k)
        moved = np.abs(x - s.x).max()
        assert 0.0 < moved < 0.02
        sp = copy.copy(s)
        sp.x = x
        f_ref, elj, eel, _ = _bruteforce(sp)
        assert rel_rms(ens.forces(k), f_ref) <= 5e-6, k
        assert abs(e[k, 1] - eel) <= 5e-5 * abs(eel)


def test_split_and_autotune_keep_forces(systems, ens):
    announce("autotune keeps forces")
    ens.run(1, nb.MODE_BATCHED, 0.0, 1, True)
    ref = [ens.forces(k) for k in range(len(systems))]
    chosen, wpb = ens.autotune([1000, 3000], compute_energy=True)
    assert len(chosen) == 3 and 1 <= wpb <= 4
    ens.set_split([1000, 3000], [0, 32, 16], 2)
    ens.run(1, nb.MODE_BATCHED, 0.0, 1, True)
    for k in range(len(systems)):
        # A different split regroups the FP32 i-force partial sums: equal to rounding.
        assert rel_rms(ens.forces(k), ref[k]) <= 1e-6
    # For a fixed split the result is bitwise reproducible.
    again = [ens.forces(k) for k in range(len(systems))]
    ens.run(1, nb.MODE_BATCHED, 0.0, 1, True)
    for k in range(len(systems)):
        assert np.array_equal(ens.forces(k), again[k])


def test_gpu_search_ensemble_bitwise(systems):
    """Lists built on the device give the same ensemble, bit for bit."""
    announce("ensemble with device-built lists")
    args = (systems, RC, RTOL, 1.1, 1.04, nb.VDW_CUT_TYPE_TABLE, nb.ACCUM_FIXED64)
    host = nb.Ensemble(*args, max_pairs_per_entry=48)
    dev = nb.Ensemble(*args, max_pairs_per_entry=48, gpu_search=True)
    assert dev.stats() == host.stats()
    assert dev.count_pairs()[0] == host.count_pairs()[0]
    for e in (host, dev):
        e.run(1, nb.MODE_BATCHED, 0.0, 5, True)
        e.sync()
    for k in range(len(systems)):
        assert np.array_equal(dev.forces(k), host.forces(k))
    e1, f1 = host.energies()
    e2, f2 = dev.energies()
    assert np.array_equal(e1, e2) and np.array_equal(f1, f2)
