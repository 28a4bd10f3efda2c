"""Shared helpers of the test suite (no GPU use at import time)."""
import numpy as np


def rel_rms(f, ref):
    f = np.asarray(f, dtype=np.float64).reshape(-1, 3)
    ref = np.asarray(ref, dtype=np.float64).reshape(-1, 3)
    return float(np.sqrt(np.sum((f - ref) ** 2) / np.sum(ref ** 2)))


def real_mask(atoms):
    return np.repeat(atoms.atom_indices >= 0, 1)


def wrapped_float_system(system):
    """Copy of a system whose coordinates are the FP32-rounded wrapped ones the
    kernels see, so brute-force FP64 sums and list kernels share inputs."""
    import copy
    s = copy.copy(system)
    x = np.asarray(system.x, dtype=np.float64)
    box = np.asarray(system.box, dtype=np.float32).astype(np.float64)
    s.box = box
    xw = x - np.floor(x / box) * box
    xw = xw.astype(np.float32)
    xw[xw >= box.astype(np.float32)] = 0.0
    s.x = xw.astype(np.float64)
    s.q = np.asarray(system.q, dtype=np.float32).astype(np.float64)
    # LJ parameters as the FP32 type table holds them (6*C6, 12*C12 in FP32).
    c = np.array(system.c6c12, dtype=np.float64)
    c[..., 0] = (6.0 * c[..., 0]).astype(np.float32).astype(np.float64) / 6.0
    c[..., 1] = (12.0 * c[..., 1]).astype(np.float32).astype(np.float64) / 12.0
    s.c6c12 = c
    return s
