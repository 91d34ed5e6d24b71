"""The FP64 exponential of the pair kernels (exp2_neg in
lddmm_amd/csrc/common.hpp): 2^-t = 2^f 2^-n with n = rint(t), f = n - t and
2^f by a degree-10 polynomial on [-1/2, 1/2].  The coefficients are read
from the source and checked against 2^f in 50-digit arithmetic: relative
error below 2.8e-16 (the Remez-levelled 2.12e-16 of the exact coefficients
plus their rounding to FP64, mostly c0 = 1.00000000000000007 -> 1).
Evaluated with FP64 Horner steps the result stays within 3 ulp of 2^f;
the ldexp part is exact."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def coefficients():
    src = open(os.path.join(ROOT, "lddmm_amd", "csrc", "common.hpp")).read()
    body = src[src.index("double exp2_neg(double t)"):]
    body = body[:body.index("return")]
    first = re.search(r"double p = ([0-9.e+-]+);", body).group(1)
    rest = re.findall(r"p = fma\(p, f, ([0-9.e+-]+)\);", body)
    return [first] + rest   # highest degree first, as strings


def test_exp2_polynomial_is_minimax_accurate():
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 50
    cs = coefficients()
    assert len(cs) == 11
    c64 = [mp.mpf(float(c)) for c in cs]   # the FP64 values the device uses
    worst = mp.mpf(0)
    for k in range(2001):
        f = mp.mpf(-0.5) + mp.mpf(k) / 2000
        p = mp.mpf(0)
        for c in c64:
            p = p * f + c
        worst = max(worst, abs(p / mp.power(2, f) - 1))
    assert worst <= 2.8e-16, worst


def test_exp2_fp64_horner_within_ulps():
    """FP64 Horner with rounding after each step (an upper bound on the fused
    steps' error) against 2^f in long double."""
    cs = [float(c) for c in coefficients()]
    f = np.linspace(-0.5, 0.5, 100001)
    p = np.full_like(f, cs[0])
    for c in cs[1:]:
        p = p * f + c
    ref = np.exp2(f.astype(np.longdouble))
    ulp = np.spacing(np.abs(ref.astype(np.float64)))
    assert np.max(np.abs((p - ref) / ulp)) <= 3.0
