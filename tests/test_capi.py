"""The C-ABI library loads, exports every function include/nbnxm_amd.h
declares, and rejects bad arguments before touching the GPU (CPU only)."""
import ctypes
import os
import re

import pytest

import nbnxm_amd as nb
from nbnxm_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "nbnxm_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(nbnxm_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_functions():
    names = header_functions()
    assert "nbnxm_gpu_launch_kernel" in names and "nbnxm_ensemble_run" in names
    assert len(names) >= 40


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert missing == []


def test_binding_covers_header():
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert set(header_functions()) == bound


def test_struct_layouts():
    assert ctypes.sizeof(_lib.Sci) == 16
    assert ctypes.sizeof(_lib.CjPacked) == 24
    assert ctypes.sizeof(_lib.Excl) == 256


def test_error_conventions_without_gpu():
    lib = nb.load()
    assert lib.nbnxm_gpu_init(None, None, None, 1, None) == -1
    assert b"null" in lib.nbnxm_last_error()
    assert lib.nbnxm_gpu_launch_kernel(None, 1, 1) == -1
    assert lib.nbnxm_ensemble_create(0, None, None, None) == -1
    assert lib.nbnxm_search_construct_pairlist(None, 1.0, None, None, 0) == -1
    ic = _lib.InteractionConst()
    assert lib.nbnxm_init_interaction_const(-1.0, 1e-5, 1.0, ctypes.byref(ic)) == -1
    assert lib.nbnxm_init_interaction_const(1.0, 1e-5, 1.0, ctypes.byref(ic)) == 0
    assert ic.ewaldcoeff_q == pytest.approx(3.1234132743, abs=1e-9)
    assert ic.epsfac == pytest.approx(138.935457644, rel=1e-9)
    assert ic.sh_ewald == pytest.approx(1e-5, rel=1e-9)
    assert ic.dispersion_shift_cpot == -1.0 and ic.repulsion_shift_cpot == -1.0


def test_shift_vectors():
    import numpy as np
    lib = nb.load()
    box = np.array([2.0, 3.0, 4.0])
    sv = np.zeros(3 * 45, dtype=np.float32)
    assert lib.nbnxm_calc_shifts(box.ctypes.data, sv.ctypes.data) == 0
    sv = sv.reshape(45, 3)
    assert sv[nb.CENTRAL_SHIFT].tolist() == [0.0, 0.0, 0.0]
    # index 5*(3*(z+1)+y+1)+x+2 (ishift.h:58-62)
    assert sv[5 * (3 * 2 + 0) + 4].tolist() == [4.0, -3.0, 4.0]
