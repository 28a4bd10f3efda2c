"""nbnxm_gpu_launch_kernel_noalloc, the entry point that allocates nothing,
against the task path (nbnxm_gpu_init + launch_kernel + cpyback + wait) on the
same constants, pruned list and atom data: fixed-point forces, energies and
shift forces equal bit for bit, on two calls in a row (the second with fewer
entries, so the reused scratch holds the first call's partials beyond them).
Device memory for inputs and outputs is allocated here, by the test, through
the HIP runtime the library is bound to."""
import ctypes

import numpy as np
import pytest

import nbnxm_amd as nb
from conftest import announce

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Hip:
    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize"):
            getattr(self.lib, f).restype = ctypes.c_int
        self.ptrs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        self.ptrs.append(p.value)
        if a.nbytes:
            assert self.lib.hipMemcpy(p.value, a.ctypes.data, a.nbytes, H2D) == 0
        return p.value

    def download(self, p, like):
        out = np.empty_like(like)
        assert self.lib.hipDeviceSynchronize() == 0
        assert self.lib.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.lib.hipFree(ctypes.c_void_p(p))
        self.ptrs = []


def _task_reference(ic, atoms, plist):
    """Forces (double, from fixed point), energies, shift forces and the pruned
    list of the task path."""
    task = nb.GpuTask(ic, atoms, 1.1, 1.04, nb.ACCUM_FIXED64)
    task.init_pairlist(plist)
    task.launch(compute_energy=True, compute_virial=True)
    f, fd, elj, eel, fshift = task.fetch()
    inner, _ = task.masks()
    pruned = plist.copy()
    pruned.cj_packed["imask"] = inner
    return fd.reshape(-1, 3), elj, eel, fshift, pruned


def _noalloc(hip, ic, atoms, plist):
    lib = nb.load()
    n = atoms.num_atoms
    sv = np.zeros((nb.NUM_SHIFTS, 4), dtype=np.float32)
    sv[:, :3] = atoms.shift_vec.reshape(-1, 3)
    inp = nb._lib.DeviceInputs()
    inp.numAtoms = n
    inp.xq = hip.upload(atoms.xq)
    inp.numTypes = atoms.num_types
    inp.type = None
    inp.nbfp = None
    inp.vdwType = atoms.vdw_type
    inp.ljComb = hip.upload(atoms.lj_comb)
    inp.shiftVec = hip.upload(sv)
    inp.nsci = len(plist.sci)
    inp.sci = hip.upload(plist.sci)
    inp.cjPacked = hip.upload(plist.cj_packed)
    inp.excl = hip.upload(plist.excl)
    f_fixed = np.zeros(3 * n, dtype=np.uint64)
    d_f = hip.upload(f_fixed)
    d_e = hip.upload(np.zeros(2))
    d_fs = hip.upload(np.zeros(3 * nb.NUM_SHIFTS))
    rc = lib.nbnxm_gpu_launch_kernel_noalloc(ctypes.byref(ic), ctypes.byref(inp), nb.ACCUM_FIXED64, 1, 1,
                                             ctypes.c_void_p(d_f), None, ctypes.c_void_p(d_e), ctypes.c_void_p(d_fs), None)
    nb._lib.check(rc, "gpu_launch_kernel_noalloc")
    fixed = hip.download(d_f, f_fixed).view(np.int64)
    e = hip.download(d_e, np.zeros(2))
    fs = hip.download(d_fs, np.zeros(3 * nb.NUM_SHIFTS))
    return (fixed.astype(np.float64) * 2.0 ** -32).reshape(-1, 3), e[0], e[1], fs


def test_noalloc_equals_task_path_twice():
    ic = nb.interaction_const(1.0, 1e-5)
    hip = _Hip()
    try:
        for name, system in (("water 9k", nb.systems.config_water_small()),
                             ("water 3k", nb.systems.water_box(1000, None, seed=61))):
            _, plist, atoms = nb.build(system, 1.1, nb.VDW_CUT_COMB_GEOM)
            f_ref, elj_ref, eel_ref, fs_ref, pruned = _task_reference(ic, atoms, plist)
            announce(f"noalloc {name}: {len(pruned.sci)} entries")
            f, elj, eel, fs = _noalloc(hip, ic, atoms, pruned)
            assert np.array_equal(f, f_ref), name
            assert elj == elj_ref and eel == eel_ref, name
            assert np.array_equal(fs, fs_ref), name
            hip.free()
    finally:
        hip.free()


def test_noalloc_rejects_beyond_scratch():
    lib = nb.load()
    ic = nb.interaction_const(1.0, 1e-5)
    inp = nb._lib.DeviceInputs()
    inp.nsci = nb._lib.NOALLOC_MAX_ENTRIES + 1
    rc = lib.nbnxm_gpu_launch_kernel_noalloc(ctypes.byref(ic), ctypes.byref(inp), nb.ACCUM_FIXED64, 1, 1,
                                             ctypes.c_void_p(8), None, ctypes.c_void_p(8), ctypes.c_void_p(8), None)
    assert rc == -1
    assert b"NBNXM_NOALLOC_MAX_ENTRIES" in lib.nbnxm_last_error()
    ic2 = nb.interaction_const_ex(1.0, coulomb="rf")
    inp.nsci = 1
    assert lib.nbnxm_gpu_launch_kernel_noalloc(ctypes.byref(ic2), ctypes.byref(inp), nb.ACCUM_FIXED64, 1, 1,
                                               ctypes.c_void_p(8), None, ctypes.c_void_p(8), ctypes.c_void_p(8), None) == -4
