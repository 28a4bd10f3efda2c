"""The erfc approximation of the tuned kernel's exclusion-free path,
erfc(z) = exp(-z^2) R(1/(1 + a z)), checked against scipy's erfcx over the
range the kernel uses (z = beta r <= 4.5).  The coefficients and the scale a
are read from the kernel source, and R is evaluated as the kernel does: FP32
Horner steps with one rounding per fused multiply-add."""
import pathlib
import re

import numpy as np
import pytest

scipy_special = pytest.importorskip("scipy.special")

SOURCE = pathlib.Path(__file__).resolve().parents[1] / "gromacs-fmm_amd" / "csrc" / "flavour_math.h"
FLOAT = r"(-?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?)f"


def kernel_polynomial():
    text = SOURCE.read_text()
    scale = float(re.search(r"c_erfcxPolyScale = " + FLOAT, text).group(1))
    body = text[text.index("float erfcxPolyT(float t)"):]
    body = body[:body.index("return p;")]
    first = float(re.search(r"float p = " + FLOAT, body).group(1))
    rest = [float(x) for x in re.findall(r"__builtin_fmaf\(p, t, " + FLOAT + r"\)", body)]
    return scale, np.array([first] + rest, dtype=np.float32)


def horner_f32(coef, t):
    p = np.full_like(t, coef[0])
    for k in coef[1:]:
        # FP32 fma: the product of two floats is exact in FP64, one rounding.
        p = (p.astype(np.float64) * t.astype(np.float64) + np.float64(k)).astype(np.float32)
    return p


def test_erfcx_polynomial_relative_error():
    scale, coef = kernel_polynomial()
    assert len(coef) == 7
    z = np.linspace(0.0, 4.5, 200001)
    t = (np.float32(1.0) / (np.float32(1.0) + np.float32(scale) * z.astype(np.float32))).astype(np.float32)
    rel = horner_f32(coef, t).astype(np.float64) / scipy_special.erfcx(z) - 1.0
    assert np.abs(rel).max() < 6e-7
    assert np.sqrt(np.mean(rel ** 2)) < 2.5e-7


def test_erfc_near_cutoff_keeps_relative_accuracy():
    """erfc(beta rc) for rtol 1e-5 (z = 3.123) is 1e-5: the product
    exp(-z^2) R keeps its relative accuracy where 1 - erf(z) would not."""
    scale, coef = kernel_polynomial()
    z = np.array([3.0, 3.1234132743, 3.5, 4.0, 4.5])
    t = (np.float32(1.0) / (np.float32(1.0) + np.float32(scale) * z.astype(np.float32))).astype(np.float32)
    approx = np.exp(-z * z) * horner_f32(coef, t).astype(np.float64)
    assert np.all(np.abs(approx / scipy_special.erfc(z) - 1.0) < 6e-7)
