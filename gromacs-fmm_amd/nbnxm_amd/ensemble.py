"""Batched ensemble of independent systems over the nbnxm_ensemble_* C ABI."""
import ctypes

import numpy as np

from . import _lib
from ._lib import (ACCUM_FIXED64, MODE_BATCHED, MODE_PLAIN, NUM_SHIFTS, VDW_CUT_COMB_GEOM, EnsembleOptions,
                   SystemDesc, check, ptr)

__all__ = ["Ensemble", "MODE_BATCHED", "MODE_PLAIN"]


def system_descs(systems):
    """nbnxm_system_desc_t array for the systems, plus the arrays it points to."""
    keep = []
    descs = (SystemDesc * len(systems))()
    for k, s in enumerate(systems):
        x = np.ascontiguousarray(s.x, dtype=np.float64)
        q = np.ascontiguousarray(s.q, dtype=np.float64)
        t = np.ascontiguousarray(s.types, dtype=np.int32)
        r = np.ascontiguousarray(s.excl_ranges, dtype=np.int32)
        e = np.ascontiguousarray(s.excl_elements, dtype=np.int32)
        keep += [x, q, t, r, e]
        descs[k].numAtoms = s.n
        descs[k].x = ptr(x)
        descs[k].q = ptr(q)
        descs[k].type = ptr(t)
        descs[k].exclRanges = ptr(r)
        descs[k].exclElements = ptr(e) if e.size else ptr(np.zeros(1, dtype=np.int32))
        for d in range(3):
            descs[k].box[d] = float(s.box[d])
    return descs, keep


class Ensemble:
    """All systems must share one (C6, C12) type table."""

    def __init__(self, systems, rc=1.0, ewald_rtol=1e-5, rlist_outer=1.1, rlist_inner=1.04,
                 vdw_type=VDW_CUT_COMB_GEOM, accumulation=ACCUM_FIXED64, max_pairs_per_entry=64,
                 waves_per_block=4, epsilon_r=1.0, gpu_search=False):
        self._lib = _lib.load()
        self.systems = systems
        c6c12 = systems[0].c6c12
        for s in systems[1:]:
            if s.c6c12.shape != c6c12.shape or not np.array_equal(s.c6c12, c6c12):
                raise ValueError("all systems of an ensemble share the type table")
        descs, self._keep = system_descs(systems)
        tab = np.ascontiguousarray(c6c12.reshape(-1), dtype=np.float64)
        self._keep.append(tab)
        opt = EnsembleOptions()
        opt.numTypes = c6c12.shape[0]
        opt.c6c12 = ptr(tab)
        opt.vdwType = vdw_type
        opt.rcoulomb = rc
        opt.ewaldRtol = ewald_rtol
        opt.epsilonR = epsilon_r
        opt.rlistOuter = rlist_outer
        opt.rlistInner = rlist_inner
        opt.accumulation = accumulation
        opt.maxPairsPerEntry = max_pairs_per_entry
        opt.wavesPerBlock = waves_per_block
        opt.numHostThreads = 0
        opt.gpuSearch = 1 if gpu_search else 0
        h = ctypes.c_void_p()
        check(self._lib.nbnxm_ensemble_create(len(systems), ctypes.cast(descs, ctypes.c_void_p), ctypes.byref(opt), ctypes.byref(h)), "ensemble_create")
        self._h = h
        self.accumulation = accumulation
        self.rc = rc
        self._keep = []

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.nbnxm_ensemble_free(self._h)
            self._h = None

    def set_split(self, bucket_edges, max_pairs, waves_per_block=0):
        edges = np.ascontiguousarray(bucket_edges, dtype=np.int32)
        mp = np.ascontiguousarray(max_pairs, dtype=np.int32)
        check(self._lib.nbnxm_ensemble_set_split(self._h, len(mp), ptr(edges) if edges.size else None, ptr(mp), int(waves_per_block)), "ensemble_set_split")

    def autotune(self, bucket_edges, compute_energy=False):
        """Picks the split per size bucket and the waves per block by timing
        the batched launch (F+E with compute_energy, else F)."""
        edges = np.ascontiguousarray(bucket_edges, dtype=np.int32)
        chosen = np.zeros(len(edges) + 1, dtype=np.int32)
        w = ctypes.c_int(0)
        check(self._lib.nbnxm_ensemble_autotune(self._h, len(chosen), ptr(edges) if edges.size else None,
                                                int(compute_energy), ptr(chosen), ctypes.byref(w)), "ensemble_autotune")
        return chosen.tolist(), w.value

    def set_i_cluster_in_lds(self, mode):
        """Batched launch: 1 i-cluster data in LDS, 0 in registers, -1 by entry count."""
        check(self._lib.nbnxm_ensemble_set_i_cluster_in_lds(self._h, int(mode)), "ensemble_set_i_cluster_in_lds")

    def launch_config(self):
        """(waves per block, i-cluster data in LDS 0/1) of the batched launch."""
        w = ctypes.c_int(0)
        lds = ctypes.c_int(0)
        check(self._lib.nbnxm_ensemble_get_launch_config(self._h, ctypes.byref(w), ctypes.byref(lds)), "ensemble_get_launch_config")
        return w.value, lds.value

    def run(self, steps, mode=MODE_BATCHED, sigma=0.0, seed=1, compute_energy=False, use_graph=False, steps_per_graph=0):
        check(self._lib.nbnxm_ensemble_run(self._h, mode, steps, sigma, seed, int(compute_energy), int(use_graph), steps_per_graph), "ensemble_run")

    def set_step(self, step):
        check(self._lib.nbnxm_ensemble_set_step(self._h, int(step)), "ensemble_set_step")

    def sync(self):
        check(self._lib.nbnxm_ensemble_sync(self._h), "ensemble_sync")

    def time(self, steps, mode=MODE_BATCHED, sigma=0.0, seed=1, compute_energy=False, use_graph=False, kernel_time=False):
        total = ctypes.c_double(0.0)
        kern = ctypes.c_double(0.0)
        check(self._lib.nbnxm_ensemble_time(self._h, mode, steps, sigma, seed, int(compute_energy), int(use_graph),
                                            ctypes.byref(total), ctypes.byref(kern) if kernel_time else None), "ensemble_time")
        return total.value, (kern.value if kernel_time else None)

    def forces(self, k):
        f = np.zeros(3 * self.systems[k].n)
        check(self._lib.nbnxm_ensemble_get_forces(self._h, k, ptr(f)), "ensemble_get_forces")
        return f.reshape(-1, 3)

    def positions(self, k):
        x = np.zeros(3 * self.systems[k].n)
        check(self._lib.nbnxm_ensemble_get_positions(self._h, k, ptr(x)), "ensemble_get_positions")
        return x.reshape(-1, 3)

    def energies(self):
        e = np.zeros(2 * len(self.systems))
        fs = np.zeros(3 * NUM_SHIFTS * len(self.systems))
        check(self._lib.nbnxm_ensemble_get_energies(self._h, ptr(e), ptr(fs)), "ensemble_get_energies")
        return e.reshape(-1, 2), fs.reshape(len(self.systems), NUM_SHIFTS, 3)

    def stats(self):
        s = np.zeros(6, dtype=np.int64)
        check(self._lib.nbnxm_ensemble_stats(self._h, ptr(s)), "ensemble_stats")
        keys = ["atoms", "grid_atoms", "work_items", "packed_groups", "cluster_pairs_list", "cluster_pairs_inner"]
        return dict(zip(keys, (int(v) for v in s)))

    def count_pairs(self, rc=None):
        per = np.zeros(len(self.systems), dtype=np.int64)
        tot = ctypes.c_int64(0)
        check(self._lib.nbnxm_ensemble_count_pairs(self._h, self.rc if rc is None else rc, ptr(per), ctypes.byref(tot)), "ensemble_count_pairs")
        return tot.value, per
