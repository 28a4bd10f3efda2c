"""nblib-style force calculator over the GPU kernel.

Mirrors nblib::GmxNBForceCalculatorCpu (api/nblib/include/nblib/gmxcalculatorcpu.h:67-105)
and nblib::NBKernelOptions (api/nblib/include/nblib/kerneloptions.h:84-102) on
the nbnxm_force_calculator_* C ABI: updatePairlist(coordinates, box) then
compute(coordinates, box, forces[, virial[, energies]]).
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import ACCUM_FIXED64, COULOMB_PME, NUM_ENERGY_TERMS, KernelOptions, check, ptr

__all__ = ["NBKernelOptions", "ForceCalculator", "detect_comb_rule", "NUM_ENERGY_TERMS"]


@dataclass
class NBKernelOptions:
    """The NBKernelOptions fields that apply to the GPU kernel, nblib defaults."""
    useGpu: bool = True
    numOpenMPThreads: int = 1
    pairlistCutoff: float = 1.0
    coulombType: int = COULOMB_PME
    useTabulatedEwaldCorr: bool = False
    ewaldRtol: float = 1e-5
    epsilonR: float = 1.0
    rlist: float = 0.0  # 0: the list radius is pairlistCutoff, as in nblib
    accumulation: int = ACCUM_FIXED64

    def to_c(self):
        o = KernelOptions()
        o.useGpu = int(self.useGpu)
        o.numOpenMPThreads = int(self.numOpenMPThreads)
        o.pairlistCutoff = float(self.pairlistCutoff)
        o.coulombType = int(self.coulombType)
        o.useTabulatedEwaldCorr = int(self.useTabulatedEwaldCorr)
        o.ewaldRtol = float(self.ewaldRtol)
        o.epsilonR = float(self.epsilonR)
        o.rlist = float(self.rlist)
        o.accumulation = int(self.accumulation)
        return o


def detect_comb_rule(non_bonded_params, tol=1e-5):
    """nbnxm_vdw_type detected from (6*C6, 12*C12) per type pair, shape (nt, nt, 2)."""
    p = np.ascontiguousarray(non_bonded_params, dtype=np.float64)
    nt = p.shape[0]
    rc = _lib.load().nbnxm_detect_comb_rule(nt, ptr(p.reshape(-1)), float(tol))
    if rc < 0:
        check(rc, "detect_comb_rule")
    return rc


class ForceCalculator:
    """GmxNBForceCalculator(particleTypeIdOfAllParticles, nonBondedParams, charges,
    exclusionRanges, exclusionElements, options) on the GPU.

    nonBondedParams: (6*C6, 12*C12) per type pair, shape (nt, nt, 2) or flat.
    """

    def __init__(self, particle_type_ids, non_bonded_params, charges, exclusion_ranges=None,
                 exclusion_elements=None, options=None):
        self._lib = _lib.load()
        t = np.ascontiguousarray(particle_type_ids, dtype=np.int32)
        p = np.ascontiguousarray(non_bonded_params, dtype=np.float64).reshape(-1)
        nt = int(round(np.sqrt(p.size // 2)))
        if nt * nt * 2 != p.size:
            raise ValueError("nonBondedParams must hold 2 values per type pair")
        q = np.ascontiguousarray(charges, dtype=np.float64)
        self.n = t.size
        if q.size != self.n:
            raise ValueError("one charge per particle")
        r = None if exclusion_ranges is None else np.ascontiguousarray(exclusion_ranges, dtype=np.int32)
        e = None if exclusion_elements is None else np.ascontiguousarray(exclusion_elements, dtype=np.int32)
        if e is not None and e.size == 0:
            e = np.zeros(1, dtype=np.int32)
        self.options = options if options is not None else NBKernelOptions()
        h = ctypes.c_void_p()
        check(self._lib.nbnxm_force_calculator_create(self.n, ptr(t), nt, ptr(p), ptr(q), ptr(r), ptr(e),
                                                      ctypes.byref(self.options.to_c()), ctypes.byref(h)),
              "force_calculator_create")
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.nbnxm_force_calculator_free(self._h)
            self._h = None

    def update_pairlist(self, coordinates, box):
        """Builds the list; like the reference, puts the coordinates in the box
        in place when given a float64 (n, 3) C-contiguous array, and returns them."""
        x = coordinates if (isinstance(coordinates, np.ndarray) and coordinates.dtype == np.float64
                            and coordinates.flags["C_CONTIGUOUS"]) else np.ascontiguousarray(coordinates, dtype=np.float64)
        b = np.ascontiguousarray(box, dtype=np.float64)
        check(self._lib.nbnxm_force_calculator_update_pairlist(self._h, x.size // 3, ptr(x), ptr(b)),
              "force_calculator_update_pairlist")
        return x

    def compute(self, coordinates, box, forces=None, virial=False, energies=False):
        """Returns (forces (n,3) with the new forces added, virial (3,3) or None,
        energies [CoulombSR, LJSR, BuckinghamSR, Coulomb14, LJ14] or None)."""
        x = np.ascontiguousarray(coordinates, dtype=np.float64)
        b = np.ascontiguousarray(box, dtype=np.float64)
        f = np.zeros((self.n, 3)) if forces is None else forces
        if f.dtype != np.float64 or not f.flags["C_CONTIGUOUS"] or f.size != 3 * self.n:
            raise ValueError("forces must be a C-contiguous float64 (n, 3) array")
        v = np.zeros(9) if virial else None
        en = np.zeros(NUM_ENERGY_TERMS) if energies else None
        check(self._lib.nbnxm_force_calculator_compute(self._h, x.size // 3, ptr(x), ptr(b), ptr(f), ptr(v), ptr(en)),
              "force_calculator_compute")
        return f, (None if v is None else v.reshape(3, 3)), en
