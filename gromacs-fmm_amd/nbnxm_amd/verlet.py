"""The nonbonded part of a device-resident MD step (nbnxm_verlet_* C ABI).

Mirrors what nonbonded_verlet_t does for do_force on one GPU: search every
nstlist steps (coordinates put in the box in place), x -> xq, the nonbonded
kernel, the rolling dynamic prune and the force reduction into a device force
array (src/gromacs/mdlib/sim_util.cpp:1619-1680, 925-946;
src/gromacs/nbnxm/pairlistsets.h:91-113).  Coordinates and forces are device
arrays of float32, 3 per atom in atom order, passed as pointers (ints).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import ACCUM_FIXED64, CJ_DTYPE, EXCL_DTYPE, SCI_DTYPE, VerletParams, check, ptr

__all__ = ["NonbondedVerlet"]


class NonbondedVerlet:
    def __init__(self, types, nonbonded_params, charges, ic, nstlist, rlist_outer, rlist_inner=None,
                 nstlist_prune=0, excl_ranges=None, excl_elements=None, accumulation=ACCUM_FIXED64,
                 gpu_search=False):
        """nonbonded_params: (numTypes, numTypes, 2) array of (6*C6, 12*C12)."""
        self._lib = _lib.load()
        self.types = np.ascontiguousarray(types, dtype=np.int32)
        self.n = int(self.types.size)
        nbfp = np.ascontiguousarray(nonbonded_params, dtype=np.float64)
        self.num_types = int(nbfp.shape[0])
        self.nbfp = nbfp.reshape(-1)
        self.q = np.ascontiguousarray(charges, dtype=np.float64)
        self.excl_ranges = None if excl_ranges is None else np.ascontiguousarray(excl_ranges, dtype=np.int32)
        self.excl_elements = None if excl_elements is None else np.ascontiguousarray(excl_elements, dtype=np.int32)
        p = VerletParams()
        p.nstlist = int(nstlist)
        p.rlistOuter = float(rlist_outer)
        p.rlistInner = float(rlist_outer if rlist_inner is None else rlist_inner)
        p.nstlistPrune = int(nstlist_prune)
        p.accumulation = int(accumulation)
        p.gpuSearch = int(bool(gpu_search))
        self.ic = ic
        h = ctypes.c_void_p()
        check(self._lib.nbnxm_verlet_create(self.n, ptr(self.types), self.num_types, ptr(self.nbfp), ptr(self.q),
                                            ptr(self.excl_ranges), ptr(self.excl_elements), ctypes.byref(ic),
                                            ctypes.byref(p), ctypes.byref(h)), "verlet_create")
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.nbnxm_verlet_free(self._h)
            self._h = None

    def step(self, step, box, d_x, d_f, wait_stream=0, compute_energy=False, compute_virial=False, accumulate=False):
        b = np.ascontiguousarray(box, dtype=np.float64)
        check(self._lib.nbnxm_verlet_step(self._h, int(step), ptr(b), ctypes.c_void_p(int(d_x)),
                                          ctypes.c_void_p(int(wait_stream)) if wait_stream else None,
                                          int(compute_energy), int(compute_virial),
                                          ctypes.c_void_p(int(d_f)), int(accumulate)), "verlet_step")

    def energies(self):
        elj = ctypes.c_double(0.0)
        eel = ctypes.c_double(0.0)
        check(self._lib.nbnxm_verlet_get_energies(self._h, ctypes.byref(elj), ctypes.byref(eel)), "verlet_get_energies")
        return elj.value, eel.value

    def virial(self):
        v = np.zeros(9)
        check(self._lib.nbnxm_verlet_get_virial(self._h, ptr(v)), "verlet_get_virial")
        return v.reshape(3, 3)

    def stats(self):
        c = np.zeros(5, dtype=np.int64)
        f = np.zeros(2, dtype=np.int32)
        check(self._lib.nbnxm_verlet_get_stats(self._h, ptr(c), ptr(f)), "verlet_get_stats")
        return dict(searches=int(c[0]), prunes=int(c[1]), steps=int(c[2]), cluster_pairs=int(c[3]),
                    last_search_step=int(c[4]), dynamic_pruning=bool(f[0]), rolling_parts=int(f[1]))

    def current_list(self):
        """(sci, cj_packed with the inner masks, excl, slot -> atom map) of the list in use."""
        sizes = np.zeros(4, dtype=np.int32)
        check(self._lib.nbnxm_verlet_get_list(self._h, None, None, None, None, ptr(sizes)), "verlet_get_list")
        sci = np.zeros(sizes[1], dtype=SCI_DTYPE)
        cj = np.zeros(sizes[2], dtype=CJ_DTYPE)
        excl = np.zeros(sizes[3], dtype=EXCL_DTYPE)
        ai = np.zeros(sizes[0], dtype=np.int32)
        check(self._lib.nbnxm_verlet_get_list(self._h, ptr(sci), ptr(cj), ptr(excl), ptr(ai), ptr(sizes)), "verlet_get_list")
        return sci, cj, excl, ai
