"""GPU prune kernel and entry ordering: bit-exact against the CPU restatement
of the reference prune on the same list."""
import numpy as np
import pytest

import nbnxm_amd as nb
import oracle
from conftest import announce

pytestmark = pytest.mark.gpu

RC, RLIST_OUTER, RLIST_INNER = 1.0, 1.1, 1.03


@pytest.fixture(scope="module")
def setup():
    system = nb.systems.config_water_small(seed=11)
    ic = nb.interaction_const(RC, 1e-5)
    search, plist, atoms = nb.build(system, RLIST_OUTER, nb.VDW_CUT_COMB_GEOM)
    return system, ic, search, plist, atoms


def test_fresh_prune_masks_and_order(setup):
    system, ic, search, plist, atoms = setup
    announce("fresh prune water9k")
    task = nb.GpuTask(ic, atoms, RLIST_OUTER, RLIST_INNER, nb.ACCUM_FIXED64, rolling_parts=2)
    task.init_pairlist(plist)
    task.prune(1)
    inner, outer = task.masks()
    order, count = task.sci_order()
    o_inner, o_outer, o_count = oracle.prune(plist, atoms.xq, atoms.shift_vec, RLIST_OUTER, RLIST_INNER, True)
    assert np.array_equal(inner, o_inner)
    assert np.array_equal(outer, o_outer)
    assert np.array_equal(count, o_count)
    assert np.array_equal(order, oracle.sci_sort_order(o_count))
    # The inner list is a subset of the outer one, which is a subset of the list.
    assert np.all((inner & ~outer) == 0)
    assert np.all((outer & ~plist.imask) == 0)
    assert (inner != outer).any()


@pytest.mark.parametrize("num_parts", [1, 2, 3])
def test_rolling_prune(setup, num_parts):
    system, ic, search, plist, atoms = setup
    announce(f"rolling prune parts={num_parts}")
    task = nb.GpuTask(ic, atoms, RLIST_OUTER, RLIST_INNER, nb.ACCUM_FIXED64, rolling_parts=num_parts)
    task.init_pairlist(plist)
    task.prune(1)
    inner, outer = task.masks()
    order, _ = task.sci_order()
    rng = np.random.default_rng(3)
    xq = atoms.xq.reshape(-1, 4).copy()
    real = atoms.atom_indices >= 0
    xq[real, :3] += rng.normal(0.0, 0.01, size=(real.sum(), 3)).astype(np.float32)
    xq = xq.reshape(-1)
    task.copy_xq(xq)
    sorted_list = plist.copy()
    sorted_list.sci = plist.sci[order].copy()
    o_inner = inner.copy()
    for part in range(num_parts):
        task.prune(num_parts)
        g_inner, g_outer = task.masks()
        o_inner, _, _ = oracle.prune(sorted_list, xq, atoms.shift_vec, RLIST_OUTER, RLIST_INNER, False,
                                     part, num_parts, imask_outer=outer, imask_inner=o_inner)
        assert np.array_equal(g_outer, outer)
        assert np.array_equal(g_inner, o_inner), f"part {part}"


def test_pruned_list_forces_unchanged(setup):
    """Forces on the rlistInner-pruned list equal those of the full list (rlistInner > rc)."""
    system, ic, search, plist, atoms = setup
    announce("pruned vs unpruned forces")
    full = nb.GpuTask(ic, atoms, RLIST_OUTER, RLIST_OUTER, nb.ACCUM_FP64)
    full.init_pairlist(plist)
    full.launch(True, True)
    _, fd_full, elj_full, eel_full, _ = full.fetch()
    pruned = nb.GpuTask(ic, atoms, RLIST_OUTER, RLIST_INNER, nb.ACCUM_FP64)
    pruned.init_pairlist(plist)
    pruned.launch(True, True)
    _, fd_pruned, elj_p, eel_p, _ = pruned.fetch()
    assert np.allclose(fd_full, fd_pruned, rtol=0, atol=1e-8)
    assert abs(elj_full - elj_p) < 1e-8 and abs(eel_full - eel_p) < 1e-7


@pytest.mark.parametrize("accumulation", [nb.ACCUM_FIXED64, nb.ACCUM_FLOAT])
@pytest.mark.parametrize("energy", [False, True])
def test_combined_prune_kernel(setup, accumulation, energy):
    """A fresh list without dynamic pruning: the interaction kernel prunes it
    to rlistOuter itself (the reference's _prune kernels).  Its masks equal
    the oracle prune with inner = outer = rlistOuter and its entry order the
    oracle sort; its forces, energies and shift forces equal those of the
    two-launch path (prune kernel, then the interaction kernel on the sorted,
    pruned list), bit for bit in fixed point; the next launch, on the list it
    pruned, gives them again."""
    system, ic, search, plist, atoms = setup
    announce(f"combined prune accumulation={accumulation} energy={energy}")
    fused = nb.GpuTask(ic, atoms, RLIST_OUTER, None, accumulation, dynamic_pruning=False)
    fused.init_pairlist(plist)
    fused.launch(energy, True)
    f, _, elj, eel, fs = fused.fetch(want_double=False)
    inner, outer = fused.masks()
    order, count = fused.sci_order()
    o_inner, o_outer, o_count = oracle.prune(plist, atoms.xq, atoms.shift_vec, RLIST_OUTER, RLIST_OUTER, True)
    assert np.array_equal(o_inner, o_outer)
    assert np.array_equal(inner, o_inner) and np.array_equal(outer, o_inner)
    assert np.array_equal(count, o_count)
    assert np.array_equal(order, oracle.sci_sort_order(o_count))
    assert (inner != plist.imask).any()

    two = nb.GpuTask(ic, atoms, RLIST_OUTER, None, accumulation, dynamic_pruning=True)
    two.init_pairlist(plist)
    two.launch(energy, True)
    f2, _, elj2, eel2, fs2 = two.fetch(want_double=False)
    assert np.array_equal(two.masks()[0], inner)
    # Float forces accumulate over launches until gpu_clear_outputs.
    fused.clear_outputs()
    fused.launch(energy, True)
    f3, _, elj3, eel3, fs3 = fused.fetch(want_double=False)
    if accumulation == nb.ACCUM_FIXED64:
        for g, e_lj, e_el, s in ((f2, elj2, eel2, fs2), (f3, elj3, eel3, fs3)):
            assert np.array_equal(f, g)
            assert e_lj == elj and e_el == eel
            assert np.array_equal(fs, s)
    else:
        # Float atomics: the entry order differs between the launches.
        for g in (f2, f3):
            assert np.abs(f - g).max() <= 1e-5 * np.abs(f).max()
        if energy:
            assert elj2 == pytest.approx(elj, rel=1e-6) and eel2 == pytest.approx(eel, rel=1e-6)


def test_outer_masks_of_empty_groups_after_reupload(setup):
    """A list uploaded over an earlier one of the same size with some groups
    emptied (imask 0): the fresh prune leaves their outer masks 0, not what
    the earlier list's prune wrote there, so the rolling prune that follows
    checks nothing in them."""
    system, ic, search, plist, atoms = setup
    announce("outer masks of emptied groups after a list upload")
    task = nb.GpuTask(ic, atoms, RLIST_OUTER, RLIST_INNER, nb.ACCUM_FIXED64, rolling_parts=1)
    task.init_pairlist(plist)
    task.prune(1)
    emptied = plist.copy()
    idx = np.arange(0, len(emptied.cj_packed), 7)
    assert (task.masks()[1][idx] != 0).any()
    emptied.cj_packed["imask"][idx] = 0
    task.init_pairlist(emptied)
    task.prune(1)
    inner, outer = task.masks()
    o_inner, o_outer, _ = oracle.prune(emptied, atoms.xq, atoms.shift_vec, RLIST_OUTER, RLIST_INNER, True)
    assert np.array_equal(outer, o_outer) and np.array_equal(inner, o_inner)
    assert np.all(outer[idx] == 0)
    order, _ = task.sci_order()
    rng = np.random.default_rng(5)
    xq = atoms.xq.reshape(-1, 4).copy()
    real = atoms.atom_indices >= 0
    xq[real, :3] += rng.normal(0.0, 0.01, size=(real.sum(), 3)).astype(np.float32)
    xq = xq.reshape(-1)
    task.copy_xq(xq)
    task.prune(1)
    sorted_list = emptied.copy()
    sorted_list.sci = emptied.sci[order].copy()
    o_rolled, _, _ = oracle.prune(sorted_list, xq, atoms.shift_vec, RLIST_OUTER, RLIST_INNER, False,
                                  0, 1, imask_outer=o_outer, imask_inner=o_inner)
    assert np.array_equal(task.masks()[0], o_rolled)
