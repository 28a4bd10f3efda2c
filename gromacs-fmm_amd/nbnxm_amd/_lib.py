"""ctypes binding of libnbnxm_amd.so (the C ABI declared in include/nbnxm_amd.h).

The library is built in-tree (``make -C gromacs-fmm_amd``).  Loading fails
loudly when it is missing: there is no fallback implementation.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libnbnxm_amd.so")

CLUSTER_SIZE = 8
NUM_CLUSTER_PER_SUPERCLUSTER = 8
JGROUP_SIZE = 4
EXCL_SIZE = 64
NUM_SHIFTS = 45
CENTRAL_SHIFT = 22
SCI_HISTOGRAM_SIZE = 8192
FIXED_FRAC_BITS = 32

VDW_CUT_TYPE_TABLE = 0
VDW_CUT_COMB_GEOM = 1
VDW_CUT_COMB_LB = 2

ACCUM_FLOAT = 0
ACCUM_FIXED64 = 1
ACCUM_FP64 = 2


class Sci(ctypes.Structure):
    _fields_ = [("sci", ctypes.c_int), ("shift", ctypes.c_int),
                ("cjPackedBegin", ctypes.c_int), ("cjPackedEnd", ctypes.c_int)]


class CjPacked(ctypes.Structure):
    _fields_ = [("cj", ctypes.c_int * 4), ("imask", ctypes.c_uint), ("excl_ind", ctypes.c_int)]


class Excl(ctypes.Structure):
    _fields_ = [("pair", ctypes.c_uint * EXCL_SIZE)]


class PairlistGpu(ctypes.Structure):
    _fields_ = [("na_ci", ctypes.c_int), ("rlist", ctypes.c_float),
                ("nsci", ctypes.c_int), ("sci", ctypes.c_void_p),
                ("ncjPacked", ctypes.c_int), ("cjPacked", ctypes.c_void_p),
                ("nexcl", ctypes.c_int), ("excl", ctypes.c_void_p)]


IC_FIELDS = [("rcoulomb", ctypes.c_double), ("rvdw", ctypes.c_double),
             ("ewaldcoeff_q", ctypes.c_double), ("sh_ewald", ctypes.c_double),
             ("epsfac", ctypes.c_double), ("dispersion_shift_cpot", ctypes.c_double),
             ("repulsion_shift_cpot", ctypes.c_double),
             ("eeltype", ctypes.c_int), ("vdwModifier", ctypes.c_int), ("ljPme", ctypes.c_int),
             ("rvdw_switch", ctypes.c_double), ("ewaldcoeff_lj", ctypes.c_double),
             ("sh_lj_ewald", ctypes.c_double),
             ("dispersion_shift_c2", ctypes.c_double), ("dispersion_shift_c3", ctypes.c_double),
             ("repulsion_shift_c2", ctypes.c_double), ("repulsion_shift_c3", ctypes.c_double),
             ("vdw_switch_c3", ctypes.c_double), ("vdw_switch_c4", ctypes.c_double),
             ("vdw_switch_c5", ctypes.c_double), ("k_rf", ctypes.c_double), ("c_rf", ctypes.c_double),
             ("tabulatedEwald", ctypes.c_int)]


class InteractionConst(ctypes.Structure):
    _fields_ = IC_FIELDS


ELEC_EWALD = 0
ELEC_RF = 1
LJPME_NONE = 0
LJPME_GEOM = 1
LJPME_LB = 2
VDW_MOD_POTSHIFT = 0
VDW_MOD_FORCESWITCH = 1
VDW_MOD_POTSWITCH = 2


class IcOptions(ctypes.Structure):
    _fields_ = [("eeltype", ctypes.c_int), ("vdwModifier", ctypes.c_int), ("ljPme", ctypes.c_int),
                ("rcoulomb", ctypes.c_double), ("rvdw", ctypes.c_double), ("rvdw_switch", ctypes.c_double),
                ("ewaldRtol", ctypes.c_double), ("ewaldRtolLj", ctypes.c_double),
                ("epsilonR", ctypes.c_double), ("epsilonRf", ctypes.c_double), ("plainCutoff", ctypes.c_int),
                ("tabulatedEwald", ctypes.c_int)]


class VerletParams(ctypes.Structure):
    _fields_ = [("nstlist", ctypes.c_int), ("rlistOuter", ctypes.c_double), ("rlistInner", ctypes.c_double),
                ("nstlistPrune", ctypes.c_int), ("accumulation", ctypes.c_int), ("gpuSearch", ctypes.c_int)]


class PairlistParams(ctypes.Structure):
    _fields_ = [("rlistOuter", ctypes.c_double), ("rlistInner", ctypes.c_double),
                ("useDynamicPruning", ctypes.c_int), ("numRollingPruningParts", ctypes.c_int)]


class Atomdata(ctypes.Structure):
    _fields_ = [("numAtoms", ctypes.c_int), ("xq", ctypes.c_void_p),
                ("numTypes", ctypes.c_int), ("type", ctypes.c_void_p),
                ("nbfp", ctypes.c_void_p), ("vdwType", ctypes.c_int),
                ("ljComb", ctypes.c_void_p), ("shiftVec", ctypes.c_void_p)]


class KernelTiming(ctypes.Structure):
    _fields_ = [("t", ctypes.c_double), ("c", ctypes.c_int)]


class GpuTimings(ctypes.Structure):
    _fields_ = [("ktime", (KernelTiming * 2) * 2), ("pruneTime", KernelTiming),
                ("dynamicPruneTime", KernelTiming), ("nb_h2d_t", ctypes.c_double),
                ("nb_d2h_t", ctypes.c_double), ("nb_c", ctypes.c_int),
                ("pl_h2d_t", ctypes.c_double), ("pl_h2d_c", ctypes.c_int)]


class DeviceInputs(ctypes.Structure):
    _fields_ = [("numAtoms", ctypes.c_int), ("xq", ctypes.c_void_p), ("numTypes", ctypes.c_int),
                ("type", ctypes.c_void_p), ("nbfp", ctypes.c_void_p), ("vdwType", ctypes.c_int),
                ("ljComb", ctypes.c_void_p), ("shiftVec", ctypes.c_void_p), ("nsci", ctypes.c_int),
                ("sci", ctypes.c_void_p), ("cjPacked", ctypes.c_void_p), ("excl", ctypes.c_void_p)]


NOALLOC_MAX_ENTRIES = 65536


# numpy dtypes with the same layout as the C structs
SCI_DTYPE = np.dtype([("sci", np.int32), ("shift", np.int32),
                      ("cjPackedBegin", np.int32), ("cjPackedEnd", np.int32)])
CJ_DTYPE = np.dtype([("cj", np.int32, (4,)), ("imask", np.uint32), ("excl_ind", np.int32)])
EXCL_DTYPE = np.dtype([("pair", np.uint32, (EXCL_SIZE,))])

assert SCI_DTYPE.itemsize == ctypes.sizeof(Sci) == 16
assert CJ_DTYPE.itemsize == ctypes.sizeof(CjPacked) == 24
assert EXCL_DTYPE.itemsize == ctypes.sizeof(Excl) == 256

_c_int = ctypes.c_int
_c_double = ctypes.c_double
_vp = ctypes.c_void_p
_PTR = ctypes.POINTER

# (name, restype, argtypes) of every exported function
SIGNATURES = [
    ("nbnxm_last_error", ctypes.c_char_p, []),
    ("nbnxm_version", ctypes.c_char_p, []),
    ("nbnxm_calc_ewaldcoeff_q", _c_double, [_c_double, _c_double]),
    ("nbnxm_init_interaction_const", _c_int, [_c_double, _c_double, _c_double, _PTR(InteractionConst)]),
    ("nbnxm_init_interaction_const_ex", _c_int, [_PTR(IcOptions), _PTR(InteractionConst)]),
    ("nbnxm_calc_ewaldcoeff_lj", _c_double, [_c_double, _c_double]),
    ("nbnxm_ewald_correction_table", _c_int, [_c_double, _c_double, _vp, _vp, _c_int, _PTR(ctypes.c_double)]),
    ("nbnxm_calc_shifts", _c_int, [_vp, _vp]),
    ("nbnxm_search_create", _vp, []),
    ("nbnxm_search_free", None, [_vp]),
    ("nbnxm_search_put_atoms_on_grid", _c_int, [_vp, _vp, _c_int, _vp]),
    ("nbnxm_search_construct_pairlist", _c_int, [_vp, _c_double, _vp, _vp, _c_int]),
    ("nbnxm_search_get_pairlist", _c_int, [_vp, _PTR(PairlistGpu)]),
    ("nbnxm_search_num_grid_atoms", _c_int, [_vp]),
    ("nbnxm_search_atom_indices", _c_int, [_vp, _vp]),
    ("nbnxm_search_shift_vec", _c_int, [_vp, _vp]),
    ("nbnxm_search_fill_xq", _c_int, [_vp, _vp, _vp, _vp]),
    ("nbnxm_search_fill_types", _c_int, [_vp, _vp, _c_int, _vp]),
    ("nbnxm_search_count_pairs", _c_int, [_vp, _c_double, _PTR(ctypes.c_int64)]),
    ("nbnxm_search_fill_ljcomb", _c_int, [_vp, _vp, _vp, _vp]),
    ("nbnxm_gpu_init", _c_int, [_PTR(InteractionConst), _PTR(PairlistParams), _PTR(Atomdata), _c_int, _PTR(_vp)]),
    ("nbnxm_gpu_init_atomdata", _c_int, [_vp, _PTR(Atomdata)]),
    ("nbnxm_gpu_upload_shiftvec", _c_int, [_vp, _vp]),
    ("nbnxm_gpu_init_pairlist", _c_int, [_vp, _PTR(PairlistGpu)]),
    ("nbnxm_gpu_copy_xq_to_gpu", _c_int, [_vp, _vp, _c_int]),
    ("nbnxm_gpu_clear_outputs", _c_int, [_vp, _c_int]),
    ("nbnxm_gpu_launch_kernel", _c_int, [_vp, _c_int, _c_int]),
    ("nbnxm_gpu_launch_kernel_pruneonly", _c_int, [_vp, _c_int]),
    ("nbnxm_gpu_launch_cpyback", _c_int, [_vp, _vp, _vp]),
    ("nbnxm_gpu_wait_finish_task", _c_int, [_vp, _vp, _vp, _vp]),
    ("nbnxm_gpu_try_finish_task", _c_int, [_vp, _vp, _vp, _vp]),
    ("nbnxm_gpu_pme_loadbal_update_param", _c_int, [_vp, _PTR(InteractionConst), _PTR(PairlistParams)]),
    ("nbnxm_gpu_set_timing", _c_int, [_vp, _c_int]),
    ("nbnxm_gpu_get_timings", _c_int, [_vp, _PTR(GpuTimings)]),
    ("nbnxm_gpu_reset_timings", _c_int, [_vp]),
    ("nbnxm_gpu_get_f", _c_int, [_vp, _PTR(_vp)]),
    ("nbnxm_gpu_get_stream", _c_int, [_vp, _PTR(_vp)]),
    ("nbnxm_gpu_launch_kernel_noalloc", _c_int, [_PTR(InteractionConst), _PTR(DeviceInputs), _c_int, _c_int, _c_int,
                                                 _vp, _vp, _vp, _vp, _vp]),
    ("nbnxm_gpu_get_virial", _c_int, [_vp, _vp]),
    ("nbnxm_gpu_init_x_to_nbat_x", _c_int, [_vp, _c_int, _c_int, _vp]),
    ("nbnxm_gpu_x_to_nbat_x", _c_int, [_vp, _vp]),
    ("nbnxm_gpu_add_nbat_f_to_f", _c_int, [_vp, _vp, _vp, _c_int]),
    ("nbnxm_gpu_min_ci_balanced", _c_int, [_vp]),
    ("nbnxm_gpu_is_kernel_ewald_analytical", _c_int, [_vp]),
    ("nbnxm_gpu_get_masks", _c_int, [_vp, _vp, _vp]),
    ("nbnxm_gpu_get_sci_order", _c_int, [_vp, _vp, _vp]),
    ("nbnxm_gpu_set_waves_per_block", _c_int, [_vp, _c_int]),
    ("nbnxm_gpu_set_i_cluster_in_lds", _c_int, [_vp, _c_int]),
    ("nbnxm_gpu_free", None, [_vp]),
    ("nbnxm_ensemble_create", _c_int, [_c_int, _vp, _vp, _PTR(_vp)]),
    ("nbnxm_ensemble_set_split", _c_int, [_vp, _c_int, _vp, _vp, _c_int]),
    ("nbnxm_ensemble_autotune", _c_int, [_vp, _c_int, _vp, _c_int, _vp, _vp]),
    ("nbnxm_ensemble_set_i_cluster_in_lds", _c_int, [_vp, _c_int]),
    ("nbnxm_ensemble_get_launch_config", _c_int, [_vp, _vp, _vp]),
    ("nbnxm_ensemble_run", _c_int, [_vp, _c_int, _c_int, ctypes.c_float, ctypes.c_uint64, _c_int, _c_int, _c_int]),
    ("nbnxm_ensemble_sync", _c_int, [_vp]),
    ("nbnxm_ensemble_set_step", _c_int, [_vp, ctypes.c_uint64]),
    ("nbnxm_ensemble_time", _c_int, [_vp, _c_int, _c_int, ctypes.c_float, ctypes.c_uint64, _c_int, _c_int, _vp, _vp]),
    ("nbnxm_ensemble_get_forces", _c_int, [_vp, _c_int, _vp]),
    ("nbnxm_ensemble_get_energies", _c_int, [_vp, _vp, _vp]),
    ("nbnxm_ensemble_get_positions", _c_int, [_vp, _c_int, _vp]),
    ("nbnxm_ensemble_stats", _c_int, [_vp, _vp]),
    ("nbnxm_ensemble_system_list", _c_int, [_vp, _c_int, _PTR(PairlistGpu), _vp]),
    ("nbnxm_ensemble_count_pairs", _c_int, [_vp, _c_double, _vp, _vp]),
    ("nbnxm_ensemble_free", None, [_vp]),
    ("nbnxm_detect_comb_rule", _c_int, [_c_int, _vp, _c_double]),
    ("nbnxm_force_calculator_create", _c_int, [_c_int, _vp, _c_int, _vp, _vp, _vp, _vp, _vp, _PTR(_vp)]),
    ("nbnxm_force_calculator_update_pairlist", _c_int, [_vp, _c_int, _vp, _vp]),
    ("nbnxm_force_calculator_compute", _c_int, [_vp, _c_int, _vp, _vp, _vp, _vp, _vp]),
    ("nbnxm_force_calculator_free", None, [_vp]),
    ("nbnxm_gpu_search_create", _c_int, [_PTR(_vp)]),
    ("nbnxm_gpu_search_build", _c_int, [_vp, _c_int, _vp, _c_double, _c_int]),
    ("nbnxm_gpu_search_num_entries", _c_int, [_vp, _vp]),
    ("nbnxm_gpu_search_system_list", _c_int, [_vp, _c_int, _vp, _vp, _vp, _vp, _vp]),
    ("nbnxm_gpu_search_free", None, [_vp]),
    ("nbnxm_verlet_create", _c_int, [_c_int, _vp, _c_int, _vp, _vp, _vp, _vp, _PTR(InteractionConst), _PTR(VerletParams), _PTR(_vp)]),
    ("nbnxm_verlet_step", _c_int, [_vp, ctypes.c_longlong, _vp, _vp, _vp, _c_int, _c_int, _vp, _c_int]),
    ("nbnxm_verlet_get_energies", _c_int, [_vp, _PTR(_c_double), _PTR(_c_double)]),
    ("nbnxm_verlet_get_virial", _c_int, [_vp, _vp]),
    ("nbnxm_verlet_get_stats", _c_int, [_vp, _vp, _vp]),
    ("nbnxm_verlet_get_list", _c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    ("nbnxm_verlet_free", None, [_vp]),
]

COULOMB_PME = 0
COULOMB_CUTOFF = 1
COULOMB_REACTION_FIELD = 2
NUM_ENERGY_TERMS = 5


class KernelOptions(ctypes.Structure):
    _fields_ = [("useGpu", ctypes.c_int), ("numOpenMPThreads", ctypes.c_int),
                ("pairlistCutoff", ctypes.c_double), ("coulombType", ctypes.c_int),
                ("useTabulatedEwaldCorr", ctypes.c_int), ("ewaldRtol", ctypes.c_double),
                ("epsilonR", ctypes.c_double), ("rlist", ctypes.c_double), ("accumulation", ctypes.c_int)]


class SystemDesc(ctypes.Structure):
    _fields_ = [("numAtoms", ctypes.c_int), ("x", ctypes.c_void_p), ("q", ctypes.c_void_p),
                ("type", ctypes.c_void_p), ("exclRanges", ctypes.c_void_p),
                ("exclElements", ctypes.c_void_p), ("box", ctypes.c_double * 3)]


class EnsembleOptions(ctypes.Structure):
    _fields_ = [("numTypes", ctypes.c_int), ("c6c12", ctypes.c_void_p), ("vdwType", ctypes.c_int),
                ("rcoulomb", ctypes.c_double), ("ewaldRtol", ctypes.c_double), ("epsilonR", ctypes.c_double),
                ("rlistOuter", ctypes.c_double), ("rlistInner", ctypes.c_double),
                ("accumulation", ctypes.c_int), ("maxPairsPerEntry", ctypes.c_int),
                ("wavesPerBlock", ctypes.c_int), ("numHostThreads", ctypes.c_int), ("gpuSearch", ctypes.c_int)]


MODE_BATCHED = 0
MODE_PLAIN = 1

_lib = None


def load():
    """Loads the library (once) and declares every signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `make -C gromacs-fmm_amd` "
            "(there is no fallback implementation)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, restype, argtypes in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    _lib = lib
    return lib


class NbnxmError(RuntimeError):
    pass


def check(rc, what=""):
    if rc != 0:
        msg = load().nbnxm_last_error().decode()
        raise NbnxmError(f"{what}: error {rc}: {msg}")
    return rc


def ptr(a):
    """Data pointer of a contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data
