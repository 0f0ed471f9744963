"""The far-field form of the smoothing kernel, checked on the CPU against the
oracle's long-double evaluation of the reference's closed forms
(highorder.h:20-60, gaussian.h:20-58; oracle orc_kernel_ld).

Far pairs (t = rho^2/sigma^2 >= 1024 in the binary64 passes, >= 128 in the
binary32 ones, >= 64 in binary32 far tiles with one term more) take q and F
from a truncated series in u = 1/t with Z dropped:
  HIGHORDER  (t + 1.5)(t + 1)^-2 = u g(u),  g(u) = sum_k (-1)^k (1 - k/2) u^k
             q = (0.25/pi) g(u) u^2 t^-1/2 / sigma^3,  F = -3 q / rho^2
  GAUSSIAN   q = (0.25/pi) t^-1.5 / sigma^3,  F = -3 q / rho^2
This test evaluates those series in binary64 at the truncation the kernels use
(u^4 for binary64, u^3 for binary32, u^4 past t = 64) and bounds their distance from the full
kernel where each threshold lets them in, and the size of the dropped Z; it
also checks the series' coefficients against the expansion of the closed form
by exact rational arithmetic, and that a series truncated one term earlier
fails the bound (so the check can tell a dropped term).  The GPU side of the
same form is checked against the oracle's rates in tests/test_gpu_far.py."""
import ctypes as C
from fractions import Fraction

import numpy as np
import pytest

import oracle as orc

T64, T32, T32B = 1024.0, 128.0, 64.0
ZBOUND = {T64: 1e-14, T32: 1e-9, T32B: 1e-8}  # |Z| / |q| where Z is dropped


def _kernel_ld(kind, t, sigma=0.05):
    L = orc.load()
    L.orc_kernel_ld.argtypes = [C.c_int, C.c_longdouble, C.c_longdouble] + \
        [C.POINTER(C.c_longdouble)] * 3
    q, F, Z = C.c_longdouble(), C.c_longdouble(), C.c_longdouble()
    L.orc_kernel_ld(kind, np.sqrt(t) * sigma, sigma, C.byref(q), C.byref(F), C.byref(Z))
    return float(q.value), float(F.value), float(Z.value)


def _series(kind, t, degree, sigma=0.05):
    """q, F of the far form at t (binary64 arithmetic)."""
    u = 1.0 / t
    rt = 1.0 / np.sqrt(t)
    if kind == orc.HIGHORDER:
        g = sum((-1) ** k * (1.0 - 0.5 * k) * u ** k for k in range(degree + 1))
        q = 0.25 / np.pi * g * u * u * rt / sigma ** 3
    else:
        q = 0.25 / np.pi * u * rt / sigma ** 3
    F = -3.0 * q / (t * sigma * sigma)
    return q, F


def test_series_coefficients_exact():
    """(1 + 3u/2)(1 + u)^-2 expanded with exact rationals: the coefficient of
    u^k is (-1)^k (1 - k/2)."""
    # (1 + u)^-2 = sum (-1)^k (k + 1) u^k
    inv2 = [Fraction((-1) ** k * (k + 1)) for k in range(12)]
    prod = [inv2[k] + (Fraction(3, 2) * inv2[k - 1] if k else 0) for k in range(12)]
    assert prod == [Fraction((-1) ** k) * (1 - Fraction(k, 2)) for k in range(12)]


@pytest.mark.parametrize("kind", [orc.HIGHORDER, orc.GAUSSIAN])
@pytest.mark.parametrize("thresh,degree,tol", [(T64, 4, 4e-15), (T32, 3, 1e-8), (T32B, 4, 2e-9)])
def test_far_series_matches_full_kernel(kind, thresh, degree, tol):
    ts = thresh * np.exp(np.linspace(0.0, np.log(1e10), 400))
    eq = ef = zr = 0.0
    for t in ts:
        q, F, Z = _kernel_ld(kind, t)
        qs, Fs = _series(kind, t, degree)
        eq = max(eq, abs(qs - q) / abs(q))
        ef = max(ef, abs(Fs - F) / abs(F))
        # the dropped Z against the pair's q (the PSE weight vs the
        # stretching weight, both per sigma^3)
        zr = max(zr, abs(Z) / abs(q))
    print(f"  kind {kind} t >= {thresh:g}: q {eq:.2e} F {ef:.2e}, Z/q {zr:.2e}")
    assert eq < tol and ef < tol
    assert zr < ZBOUND[thresh]


@pytest.mark.parametrize("thresh,degree,tol", [(T64, 3, 4e-15), (T32, 2, 1e-8), (T32B, 3, 2e-9)])
def test_one_term_fewer_fails(thresh, degree, tol):
    """Each series truncated one term earlier misses its bound at the
    threshold: the test above can tell a dropped term."""
    q, F, _ = _kernel_ld(orc.HIGHORDER, thresh)
    qs, _ = _series(orc.HIGHORDER, thresh, degree)
    assert abs(qs - q) / abs(q) > tol
