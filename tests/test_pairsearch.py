"""Host pair-list builder: layout invariants, exclusion and Newton masks,
half-shift enumeration, entry splitting, error handling (CPU only)."""
import numpy as np
import pytest

import nbnxm_amd as nb
import oracle
from helpers import wrapped_float_system


@pytest.fixture(scope="module")
def water():
    s = nb.systems.water_box(400, None, seed=5)
    search, plist, atoms = nb.build(s, 1.1, nb.VDW_CUT_COMB_GEOM)
    return s, search, plist, atoms


def test_layout_invariants(water):
    s, search, plist, atoms = water
    assert atoms.num_atoms % 64 == 0
    assert np.all(plist.excl[0]["pair"] == 0xFFFFFFFF)
    nclusters = atoms.num_atoms // 8
    for e in plist.sci:
        assert 0 <= e["shift"] <= nb.CENTRAL_SHIFT          # half of the shifts (c_pbcShiftBackward)
        assert e["cjPackedBegin"] < e["cjPackedEnd"]
        for g in range(e["cjPackedBegin"], e["cjPackedEnd"]):
            cj = plist.cj_packed[g]
            assert 0 <= cj["excl_ind"] < len(plist.excl)
            assert np.all((cj["cj"] >= 0) & (cj["cj"] < nclusters))
    # Every real atom appears exactly once in the grid.
    idx = atoms.atom_indices
    assert sorted(idx[idx >= 0].tolist()) == list(range(s.n))
    # Fillers sit far away with zero charge and the dummy type.
    xq = atoms.xq.reshape(-1, 4)
    assert np.all(xq[idx < 0, :3] == -1e6) and np.all(xq[idx < 0, 3] == 0)


def test_central_entries_start_with_diagonal(water):
    s, search, plist, atoms = water
    for e in plist.sci:
        if e["shift"] == nb.CENTRAL_SHIFT:
            first = plist.cj_packed[e["cjPackedBegin"]]
            assert first["cj"][0] == e["sci"] * 8
            assert int(first["imask"]) & 1


def test_newton_masks_on_diagonal(water):
    s, search, plist, atoms = water
    seen = 0
    for e in plist.sci:
        if e["shift"] != nb.CENTRAL_SHIFT:
            continue
        for g in range(e["cjPackedBegin"], e["cjPackedEnd"]):
            cjp = plist.cj_packed[g]
            for jm in range(4):
                cj = int(cjp["cj"][jm])
                if cj // 8 != e["sci"]:
                    continue
                for im in range(8):
                    bit = jm * 8 + im
                    if not (int(cjp["imask"]) >> bit) & 1:
                        continue
                    ci = e["sci"] * 8 + im
                    assert ci <= cj                          # no sub-diagonal cluster pairs
                    if ci == cj:
                        seen += 1
                        pair = plist.excl[cjp["excl_ind"]]["pair"]
                        for tj in range(8):
                            for ti in range(8):
                                on = (int(pair[tj * 8 + ti]) >> bit) & 1
                                assert on == (tj > ti) or not on  # j <= i always masked
                                if tj <= ti:
                                    assert on == 0
    assert seen > 0


def test_topology_exclusions_masked(water):
    """Each intramolecular pair inside a listed cluster pair has its bit cleared
    in exactly the orientation the list holds it (pairlist.cpp:1948-2046)."""
    s, search, plist, atoms = water
    idx = atoms.atom_indices
    slot_of = np.empty(s.n, dtype=np.int64)
    slot_of[idx[idx >= 0]] = np.nonzero(idx >= 0)[0]
    excluded = set()
    for a in range(s.n):
        for b in s.excl_elements[s.excl_ranges[a]:s.excl_ranges[a + 1]]:
            if a != b:
                excluded.add((int(slot_of[a]), int(slot_of[b])))
    found = 0
    for e in plist.sci:
        if e["shift"] != nb.CENTRAL_SHIFT:
            continue
        for g in range(e["cjPackedBegin"], e["cjPackedEnd"]):
            cjp = plist.cj_packed[g]
            pair = plist.excl[cjp["excl_ind"]]["pair"]
            for jm in range(4):
                cj = int(cjp["cj"][jm])
                for im in range(8):
                    bit = jm * 8 + im
                    if not (int(cjp["imask"]) >> bit) & 1:
                        continue
                    ci = e["sci"] * 8 + im
                    for ti in range(8):
                        for tj in range(8):
                            si, sj = ci * 8 + ti, cj * 8 + tj
                            if ci == cj and tj <= ti:
                                continue
                            on = (int(pair[tj * 8 + ti]) >> bit) & 1
                            if (si, sj) in excluded:
                                found += 1
                                assert on == 0
    # Every intramolecular pair is listed once (water molecules are compact).
    assert found == len(excluded) // 2


def test_split_preserves_cluster_pairs():
    s = nb.systems.water_box(700, None, seed=8)
    _, full, _ = nb.build(s, 1.1, nb.VDW_CUT_COMB_GEOM, 0)
    _, split, _ = nb.build(s, 1.1, nb.VDW_CUT_COMB_GEOM, 20)
    assert len(split.sci) > len(full.sci)
    assert split.num_cluster_pairs() == full.num_cluster_pairs()
    assert np.array_equal(split.cj_packed, full.cj_packed)
    # Pieces of one entry are contiguous and tile its groups.
    pieces = {}
    for e in split.sci:
        pieces.setdefault((int(e["sci"]), int(e["shift"])), []).append((int(e["cjPackedBegin"]), int(e["cjPackedEnd"])))
    for e in full.sci:
        ranges = sorted(pieces[(int(e["sci"]), int(e["shift"]))])
        assert ranges[0][0] == e["cjPackedBegin"] and ranges[-1][1] == e["cjPackedEnd"]
        for (a0, a1), (b0, b1) in zip(ranges, ranges[1:]):
            assert a1 == b0


def test_small_box_images():
    """L = 1.86 nm < 2 rc: several images of the same pair are in range; the
    half-shift list holds each (pair, image) once."""
    s = nb.systems.water_box(216, None, seed=3)
    assert s.box[0] < 2.0
    search, plist, atoms = nb.build(s, 1.1, nb.VDW_CUT_TYPE_TABLE)
    shifts = set(int(v) for v in plist.sci["shift"])
    assert len(shifts) > 10


def test_errors():
    s = nb.systems.water_box(216, None, seed=3)
    search = nb.PairSearch()
    search.put_atoms_on_grid(s.box, s.x)
    with pytest.raises(nb.NbnxmError):
        search.construct_pairlist(2.0, s.excl_ranges, s.excl_elements)   # rlist > box edge
    with pytest.raises(nb.NbnxmError):
        search.construct_pairlist(-1.0)
    with pytest.raises(nb.NbnxmError):
        search.put_atoms_on_grid(np.array([0.0, 1.0, 1.0]), s.x)


def test_empty_and_single_atom():
    search = nb.PairSearch()
    search.put_atoms_on_grid(np.array([3.0, 3.0, 3.0]), np.zeros((0, 3)))
    search.construct_pairlist(1.1)
    assert len(search.pairlist().sci) == 0
    search.put_atoms_on_grid(np.array([3.0, 3.0, 3.0]), np.array([[1.0, 1.0, 1.0]]))
    search.construct_pairlist(1.1)
    pl = search.pairlist()
    # One central diagonal entry holding only the self cluster pair.
    assert len(pl.sci) == 1 and pl.sci[0]["shift"] == nb.CENTRAL_SHIFT


def test_search_count_pairs_equals_oracle():
    """The exact useful-pair count of a search equals the oracle's list count
    and the brute-force count."""
    s = wrapped_float_system(nb.systems.water_box(300, None, seed=17))
    search, plist, atoms = nb.build(s, 1.1, nb.VDW_CUT_COMB_GEOM)
    n = search.count_pairs(1.0)
    assert n == oracle.count_pairs(plist, atoms, 1.0)
    assert n == oracle.bruteforce(s, nb.interaction_const(1.0, 1e-5))[3]
    with pytest.raises(nb.NbnxmError):
        search.count_pairs(1.2)
