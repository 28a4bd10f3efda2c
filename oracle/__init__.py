"""CPU oracle for the nbnxm cluster-pair kernel -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may import
this package, and only as the checker.  It wraps lib/liboracle.so (built from
nbnxm_oracle.c by ``make -C oracle``) and the CPU 4xM baseline kernel
(lib/libcpu4xm.so).  Pinning status of each function is listed in
nbnxm_oracle.h.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "liboracle.so")

NUM_SHIFTS = 45
SCI_HISTOGRAM_SIZE = 8192


class _List(ctypes.Structure):
    _fields_ = [("na_ci", ctypes.c_int), ("rlist", ctypes.c_float),
                ("nsci", ctypes.c_int), ("sci", ctypes.c_void_p),
                ("ncjPacked", ctypes.c_int), ("cjPacked", ctypes.c_void_p),
                ("nexcl", ctypes.c_int), ("excl", ctypes.c_void_p)]


class _Atoms(ctypes.Structure):
    _fields_ = [("numAtoms", ctypes.c_int), ("xq", ctypes.c_void_p),
                ("numTypes", ctypes.c_int), ("type", ctypes.c_void_p),
                ("nbfp", ctypes.c_void_p), ("vdwType", ctypes.c_int),
                ("ljComb", ctypes.c_void_p), ("shiftVec", ctypes.c_void_p)]


class _IC(ctypes.Structure):
    """nbnxm_interaction_const_t of include/nbnxm_amd.h."""
    _fields_ = [("rcoulomb", ctypes.c_double), ("rvdw", ctypes.c_double),
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


_vp = ctypes.c_void_p
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
    lib = ctypes.CDLL(LIB_PATH)
    lib.oracle_calc_ewaldcoeff_q.restype = ctypes.c_double
    lib.oracle_calc_ewaldcoeff_q.argtypes = [ctypes.c_double, ctypes.c_double]
    lib.oracle_ewald_table.restype = ctypes.c_int
    lib.oracle_ewald_table.argtypes = [ctypes.c_double, ctypes.c_double, _vp, _vp, ctypes.c_int, _vp]
    lib.oracle_pmecorrF.restype = ctypes.c_float
    lib.oracle_pmecorrF.argtypes = [ctypes.c_float]
    for name in ("oracle_kernel_f32", "oracle_kernel_f64"):
        fn = getattr(lib, name)
        fn.restype = ctypes.c_int
        fn.argtypes = [_vp, _vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp]
    lib.oracle_bruteforce_f64.restype = ctypes.c_int
    lib.oracle_bruteforce_f64.argtypes = [ctypes.c_int, _vp, _vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    lib.oracle_bruteforce_virial_f64.restype = ctypes.c_int
    lib.oracle_bruteforce_virial_f64.argtypes = [ctypes.c_int, _vp, _vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    lib.oracle_prune.restype = ctypes.c_int
    lib.oracle_prune.argtypes = [_vp, _vp, _vp, ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp]
    lib.oracle_sci_sort.restype = ctypes.c_int
    lib.oracle_sci_sort.argtypes = [ctypes.c_int, _vp, _vp, _vp]
    lib.oracle_count_pairs_f64.restype = ctypes.c_longlong
    lib.oracle_count_pairs_f64.argtypes = [_vp, _vp, ctypes.c_double]
    _lib = lib
    return lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def _ic(ic):
    """Copies any object with the interaction-constant attributes; missing
    flavour fields mean Ewald + LJ potential shift."""
    c = _IC()
    for name, ctype in _IC._fields_:
        v = getattr(ic, name, 0)
        setattr(c, name, int(v) if ctype is ctypes.c_int else float(v))
    return c


def _list(sci, cj, excl, rlist):
    v = _List()
    v.na_ci = 8
    v.rlist = rlist
    v.nsci = len(sci)
    v.sci = _p(sci) if len(sci) else None
    v.ncjPacked = len(cj)
    v.cjPacked = _p(cj) if len(cj) else None
    v.nexcl = len(excl)
    v.excl = _p(excl)
    return v


def _atoms(a):
    v = _Atoms()
    v.numAtoms = a.num_atoms
    v.xq = _p(a.xq)
    v.numTypes = a.num_types
    v.type = _p(a.types)
    v.nbfp = _p(a.nbfp)
    v.vdwType = a.vdw_type
    v.ljComb = _p(a.lj_comb)
    v.shiftVec = _p(a.shift_vec)
    return v


def ewald_coeff(rc, rtol):
    return load().oracle_calc_ewaldcoeff_q(rc, rtol)


def ewald_table(beta, rc):
    """(scale, F table FP32, F table FP64) of the tabulated Ewald correction."""
    lib = load()
    scale = ctypes.c_double(0.0)
    n = lib.oracle_ewald_table(beta, rc, None, None, 0, ctypes.byref(scale))
    t = np.zeros(n, dtype=np.float32)
    td = np.zeros(n, dtype=np.float64)
    if lib.oracle_ewald_table(beta, rc, _p(t), _p(td), n, ctypes.byref(scale)) != n:
        raise RuntimeError("oracle ewald table failed")
    return scale.value, t, td


def pmecorrF(z2):
    return load().oracle_pmecorrF(float(z2))


def kernel(plist, atoms, ic, compute_energy=True, precision="f64"):
    """List-based kernel.  Returns (f grid-order (n,3), fshift (45,3), e_lj, e_el)."""
    lib = load()
    n = atoms.num_atoms
    dt = np.float64 if precision == "f64" else np.float32
    f = np.zeros(3 * n, dtype=dt)
    fs = np.zeros(3 * NUM_SHIFTS, dtype=dt)
    elj = ctypes.c_double(0.0)
    eel = ctypes.c_double(0.0)
    lv = _list(plist.sci, plist.cj_packed, plist.excl, plist.rlist)
    av = _atoms(atoms)
    icv = _ic(ic)
    fn = lib.oracle_kernel_f64 if precision == "f64" else lib.oracle_kernel_f32
    rc = fn(ctypes.byref(lv), ctypes.byref(av), ctypes.byref(icv), int(compute_energy),
            _p(f), _p(fs), ctypes.byref(elj), ctypes.byref(eel))
    if rc != 0:
        raise RuntimeError(f"oracle kernel failed: {rc}")
    return f.reshape(-1, 3), fs.reshape(-1, 3), elj.value, eel.value


def _bruteforce_inputs(system):
    n = system.n
    x = np.ascontiguousarray(system.x, dtype=np.float64)
    q = np.ascontiguousarray(system.q, dtype=np.float64)
    t = np.ascontiguousarray(system.types, dtype=np.int32)
    nt = system.num_types
    nbfp = np.zeros((nt, nt, 2))
    nbfp[:, :, 0] = 6.0 * system.c6c12[:, :, 0]
    nbfp[:, :, 1] = 12.0 * system.c6c12[:, :, 1]
    nbfp = np.ascontiguousarray(nbfp.reshape(-1))
    r = np.ascontiguousarray(system.excl_ranges, dtype=np.int32)
    e = np.ascontiguousarray(system.excl_elements, dtype=np.int32)
    box = np.ascontiguousarray(system.box, dtype=np.float64)
    return n, x, q, t, nt, nbfp, r, e, box


def bruteforce(system, ic):
    """All-image FP64 sum in original atom order: (f (n,3), e_lj, e_el, pairs in range)."""
    lib = load()
    n, x, q, t, nt, nbfp, r, e, box = _bruteforce_inputs(system)
    f = np.zeros(3 * n)
    elj = ctypes.c_double(0.0)
    eel = ctypes.c_double(0.0)
    npairs = ctypes.c_longlong(0)
    icv = _ic(ic)
    rc = lib.oracle_bruteforce_f64(n, _p(x), _p(q), _p(t), nt, _p(nbfp), _p(r), _p(e) if e.size else None,
                                   _p(box), ctypes.byref(icv), _p(f), ctypes.byref(elj), ctypes.byref(eel),
                                   ctypes.byref(npairs))
    if rc != 0:
        raise RuntimeError(f"oracle bruteforce failed: {rc}")
    return f.reshape(-1, 3), elj.value, eel.value, npairs.value


def bruteforce_virial(system, ic):
    """All-image FP64 sum: (f (n,3), pair virial (3,3) = -1/2 sum r_ij (x) F_ij)."""
    lib = load()
    n, x, q, t, nt, nbfp, r, e, box = _bruteforce_inputs(system)
    f = np.zeros(3 * n)
    vir = np.zeros(9)
    icv = _ic(ic)
    rc = lib.oracle_bruteforce_virial_f64(n, _p(x), _p(q), _p(t), nt, _p(nbfp), _p(r), _p(e) if e.size else None,
                                          _p(box), ctypes.byref(icv), _p(f), _p(vir))
    if rc != 0:
        raise RuntimeError(f"oracle bruteforce failed: {rc}")
    return f.reshape(-1, 3), vir.reshape(3, 3)


def single_sum_virial(x, f, shift_vec, fshift):
    """-1/2 (sum_i x_i (x) f_i + sum_s shift_s (x) fshift_s), the single-sum
    virial of calc_vir (src/gromacs/mdlib/calcvir.cpp:89-133) in FP64."""
    x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
    f = np.asarray(f, dtype=np.float64).reshape(-1, 3)
    sv = np.asarray(shift_vec, dtype=np.float64).reshape(-1, 3)
    fs = np.asarray(fshift, dtype=np.float64).reshape(-1, 3)
    return -0.5 * (x.T @ f + sv.T @ fs)


def prune(plist, xq, shift_vec, rlist_outer, rlist_inner, fresh, part=0, num_parts=1,
          imask_outer=None, imask_inner=None):
    """Returns (imask_inner, imask_outer, sci_count) after one prune pass."""
    lib = load()
    ng = len(plist.cj_packed)
    inner = np.ascontiguousarray(plist.cj_packed["imask"] if imask_inner is None else imask_inner, dtype=np.uint32).copy()
    outer = np.zeros(ng, dtype=np.uint32) if imask_outer is None else np.ascontiguousarray(imask_outer, dtype=np.uint32).copy()
    count = np.zeros(len(plist.sci), dtype=np.int32)
    lv = _list(plist.sci, plist.cj_packed, plist.excl, plist.rlist)
    ro = np.float32(rlist_outer) * np.float32(rlist_outer)
    ri = np.float32(rlist_inner) * np.float32(rlist_inner)
    rc = lib.oracle_prune(ctypes.byref(lv), _p(np.ascontiguousarray(xq, dtype=np.float32)),
                          _p(np.ascontiguousarray(shift_vec, dtype=np.float32)), float(ro), float(ri),
                          int(fresh), int(part), int(num_parts), _p(outer), _p(inner), _p(count))
    if rc != 0:
        raise RuntimeError(f"oracle prune failed: {rc}")
    return inner, outer, count


def sci_sort_order(sci_count):
    """Stable counting sort by histogram bin: returns the entry order."""
    lib = load()
    nsci = len(sci_count)
    idx = np.zeros(nsci, dtype=[("sci", np.int32), ("shift", np.int32), ("b", np.int32), ("e", np.int32)])
    idx["sci"] = np.arange(nsci)
    out = np.zeros_like(idx)
    cnt = np.ascontiguousarray(sci_count, dtype=np.int32)
    rc = lib.oracle_sci_sort(nsci, _p(idx), _p(cnt), _p(out))
    if rc != 0:
        raise RuntimeError("oracle sci sort failed")
    return out["sci"].copy()


_cpu4xm = None
CPU4XM_PATH = os.path.join(_HERE, "lib", "libcpu4xm.so")


def _load_cpu4xm():
    global _cpu4xm
    if _cpu4xm is None:
        lib = ctypes.CDLL(CPU4XM_PATH)
        lib.cpu4xm_create.restype = _vp
        lib.cpu4xm_create.argtypes = [_vp, _vp, _vp, ctypes.c_double, ctypes.c_int]
        lib.cpu4xm_compute.restype = ctypes.c_int
        lib.cpu4xm_compute.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp]
        lib.cpu4xm_num_threads.restype = ctypes.c_int
        lib.cpu4xm_num_threads.argtypes = [_vp]
        lib.cpu4xm_num_cluster_pairs.restype = ctypes.c_longlong
        lib.cpu4xm_num_cluster_pairs.argtypes = [_vp]
        lib.cpu4xm_free.restype = None
        lib.cpu4xm_free.argtypes = [_vp]
        _cpu4xm = lib
    return _cpu4xm


def cpu4xm_available():
    return os.path.exists(CPU4XM_PATH)


class Cpu4xm:
    """The 4xM SIMD CPU baseline kernel on one system's list (geometric LJ)."""

    def __init__(self, plist, atoms, ic, r_prune, threads=0):
        self._lib = _load_cpu4xm()
        self._lv = _list(plist.sci, plist.cj_packed, plist.excl, plist.rlist)
        self._av = _atoms(atoms)
        self._ic = _ic(ic)
        self._keep = (plist, atoms)
        self._h = self._lib.cpu4xm_create(ctypes.byref(self._lv), ctypes.byref(self._av), ctypes.byref(self._ic), float(r_prune), int(threads))
        if not self._h:
            raise RuntimeError("cpu4xm_create failed (geometric LJ parameters required)")
        self.n = atoms.num_atoms

    @property
    def threads(self):
        return self._lib.cpu4xm_num_threads(self._h)

    def compute(self, compute_energy=True):
        f = np.zeros(3 * self.n, dtype=np.float32)
        vlj = ctypes.c_double(0.0)
        vc = ctypes.c_double(0.0)
        rc = self._lib.cpu4xm_compute(self._h, int(compute_energy), _p(f), ctypes.byref(vlj), ctypes.byref(vc))
        if rc != 0:
            raise RuntimeError("cpu4xm_compute failed")
        return f.reshape(-1, 3), vlj.value, vc.value

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.cpu4xm_free(self._h)
            self._h = None


def count_pairs(plist, atoms, rc):
    lv = _list(plist.sci, plist.cj_packed, plist.excl, plist.rlist)
    av = _atoms(atoms)
    return int(load().oracle_count_pairs_f64(ctypes.byref(lv), ctypes.byref(av), rc))
