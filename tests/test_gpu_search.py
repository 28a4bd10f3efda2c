"""GPU pair search: the device-built grid and list equal the host builder's,
bit for bit (sci, packed groups, exclusion masks, slot -> atom map), for
water boxes of many sizes, the solvated chain system with 1-2/1-3
exclusions, entry splitting and many systems in one build."""
import time

import numpy as np
import pytest

import nbnxm_amd as nb
from conftest import announce

pytestmark = pytest.mark.gpu


def host_list(system, rlist, max_pairs):
    search = nb.PairSearch()
    search.put_atoms_on_grid(system.box, system.x)
    search.construct_pairlist(rlist, system.excl_ranges, system.excl_elements, max_pairs)
    return search.pairlist(), search.atom_indices()


def assert_same(gpu, host):
    (gl, gai), (hl, hai) = gpu, host
    assert np.array_equal(gai, hai)
    assert gl.sci.tobytes() == hl.sci.tobytes()
    assert gl.cj_packed.tobytes() == hl.cj_packed.tobytes()
    assert gl.excl.tobytes() == hl.excl.tobytes()


SYSTEMS = [
    ("water_ensemble_min", lambda: nb.systems.water_box(216, None, seed=7)),     # 648 atoms, 1.86 nm box
    ("water_1500", lambda: nb.systems.water_box(1500, None, seed=8)),
    ("water9k", lambda: nb.systems.config_water_small()),
    ("hetero_chains", lambda: nb.systems.hetero_system(12, 50, 6000, 20, seed=12)),
]


@pytest.mark.parametrize("name,make", SYSTEMS, ids=[s[0] for s in SYSTEMS])
@pytest.mark.parametrize("max_pairs", [0, 64])
def test_gpu_list_equals_host_list(name, make, max_pairs):
    announce(f"gpu search {name} max_pairs={max_pairs}")
    s = make()
    gs = nb.GpuPairSearch()
    gs.build([s], 1.1, max_pairs)
    assert_same(gs.system_list(0), host_list(s, 1.1, max_pairs))


def test_gpu_search_many_systems():
    announce("gpu search 24 ensemble boxes in one build")
    systems = nb.systems.config_ensemble(256)[:24]
    gs = nb.GpuPairSearch()
    gs.build(systems, 1.1, 0)
    for k, s in enumerate(systems):
        assert_same(gs.system_list(k), host_list(s, 1.1, 0))


def test_gpu_search_ensemble_timing():
    """All 256 boxes (1.9 M atoms) in one device build, every list checked (a
    bounding-box distance that ties the cut-off in FP32 decides one cluster
    pair in one of these boxes, so every rounding step must match)."""
    announce("gpu search 256-box ensemble")
    systems = nb.systems.config_ensemble(256)
    gs = nb.GpuPairSearch()
    gs.build(systems, 1.1, 0)  # warm-up (allocations, code objects)
    t0 = time.perf_counter()
    gs.build(systems, 1.1, 0)
    dt = time.perf_counter() - t0
    announce(f"gpu search 256 boxes: {1e3 * dt:.1f} ms, {gs.num_entries()}")
    for k in range(len(systems)):
        assert_same(gs.system_list(k), host_list(systems[k], 1.1, 0))
