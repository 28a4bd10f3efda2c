"""nbnxm_amd: MI355X (gfx950) nbnxm short-range cluster-pair kernels.

Python host mirror of the reference's nbnxm interfaces over the C ABI of
lib/libnbnxm_amd.so (include/nbnxm_amd.h).  Import with the package directory
``gromacs-fmm_amd`` on sys.path.
"""
from . import _lib
from ._lib import (ACCUM_FIXED64, ACCUM_FLOAT, ACCUM_FP64, CENTRAL_SHIFT, NUM_SHIFTS, VDW_CUT_COMB_GEOM,
                   VDW_CUT_COMB_LB, VDW_CUT_TYPE_TABLE, NbnxmError, load)
from .core import (AtomData, GpuPairSearch, GpuTask, PairSearch, Pairlist, build, comb_params, ewald_coeff,
                   ewald_coeff_lj, ewald_correction_table, interaction_const, interaction_const_ex, is_default_flavour,
                   nbfp_table)
from . import device, systems
from .device import DeviceArray
from .ensemble import MODE_BATCHED, MODE_PLAIN, Ensemble
from .calculator import ForceCalculator, NBKernelOptions, detect_comb_rule
from .verlet import NonbondedVerlet

__all__ = [
    "ACCUM_FIXED64", "ACCUM_FLOAT", "ACCUM_FP64", "CENTRAL_SHIFT", "NUM_SHIFTS",
    "VDW_CUT_COMB_GEOM", "VDW_CUT_COMB_LB", "VDW_CUT_TYPE_TABLE", "NbnxmError", "load",
    "AtomData", "GpuPairSearch", "GpuTask", "PairSearch", "Pairlist", "build", "comb_params", "ewald_coeff",
    "ewald_coeff_lj", "ewald_correction_table", "interaction_const", "interaction_const_ex", "is_default_flavour", "nbfp_table", "systems", "_lib", "Ensemble", "MODE_BATCHED", "MODE_PLAIN",
    "ForceCalculator", "NBKernelOptions", "detect_comb_rule", "NonbondedVerlet", "device", "DeviceArray",
]
