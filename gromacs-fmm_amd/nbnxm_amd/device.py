"""Device arrays for callers that hand the library device pointers (the MD
step, the buffer ops): memory from the HIP runtime the library itself is
linked against, so a process that drives the library holds one HIP runtime.

The library is linked against libamdhip64.so.7 of /opt/rocm-7.2.0 (RUNPATH).
torch's wheel bundles another libamdhip64.so.7 (ROCm 7.0) under the same
SONAME, and whichever is loaded first serves both, so device memory is not
taken from torch here.  Plumbing only: no compute.
"""
import ctypes

import numpy as np

from ._lib import load

_H2D, _D2H = 1, 2
_hip = None


def _runtime():
    """The libamdhip64 the library loaded (dlopen by SONAME returns the
    object already mapped)."""
    global _hip
    if _hip is None:
        load()
        hip = ctypes.CDLL("libamdhip64.so.7")
        hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        hip.hipFree.argtypes = [ctypes.c_void_p]
        hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        hip.hipGetErrorString.argtypes = [ctypes.c_int]
        hip.hipGetErrorString.restype = ctypes.c_char_p
        for name in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize"):
            getattr(hip, name).restype = ctypes.c_int
        _hip = hip
    return _hip


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: HIP error {rc}: {_runtime().hipGetErrorString(rc).decode()}")


def synchronize():
    """Waits for all work on the current device."""
    _check(_runtime().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """A device copy of a numpy array of fixed shape and dtype.  `ptr` is the
    device address to pass through the C ABI."""

    def __init__(self, host):
        host = np.ascontiguousarray(host)
        self.shape, self.dtype, self.nbytes = host.shape, host.dtype, host.nbytes
        p = ctypes.c_void_p()
        _check(_runtime().hipMalloc(ctypes.byref(p), max(self.nbytes, 8)), "hipMalloc")
        self.ptr = p.value
        self.upload(host)

    @classmethod
    def zeros(cls, shape, dtype=np.float32):
        return cls(np.zeros(shape, dtype=dtype))

    def upload(self, host):
        """Copies `host` (same shape and dtype) to the device, after all
        queued device work."""
        host = np.ascontiguousarray(host, dtype=self.dtype)
        if host.shape != self.shape:
            raise ValueError(f"shape {host.shape} != {self.shape}")
        synchronize()
        if self.nbytes:
            _check(_runtime().hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, _H2D), "hipMemcpy H2D")

    def download(self):
        """A host copy, after all queued device work."""
        out = np.empty(self.shape, dtype=self.dtype)
        synchronize()
        if self.nbytes:
            _check(_runtime().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, _D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr is not None and _hip is not None:
            _hip.hipFree(ctypes.c_void_p(self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
