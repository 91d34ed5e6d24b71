"""The FP32 Bessel factors of the 2D Matern kernels (pair_driver.hpp,
bessel_p01_f32): the constants in the source are read back, the device
arithmetic is emulated in numpy FP32 operation by operation, and the result
is compared with scipy's k0/k1; the series coefficients are checked against
their closed forms and the fitted ones against lddmm_amd/fit_bessel_f32.py."""
import os
import re
from math import factorial

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "lddmm_amd", "csrc", "pair_driver.hpp")
f32 = np.float32


def _function_text():
    s = open(SRC).read()
    a = s.index("void bessel_p01_f32(")
    return s[a:s.index("\n}\n", a)]


def _horner_coefs(text, name):
    """Coefficients (ascending powers) of `const float name = c0 + v * (c1 + ...)`."""
    m = re.search(r"const float %s\s*=(.*?);" % name, text, re.S)
    return [float(c[:-1]) for c in re.findall(r"[-+]?\d+\.\d+(?:e[-+]\d+)?f", m.group(1))]


def coefs():
    t = _function_text()
    return {k: _horner_coefs(t, k) for k in ("i0", "s0", "i1", "s1", "f0", "f1")}


def horner(c, x):
    acc = np.full_like(x, f32(c[-1]))
    for a in c[-2::-1]:
        acc = (acc * x + f32(a)).astype(f32)
    return acc


def p01_emulated(x, c):
    x = np.asarray(x, dtype=f32)
    p0, p1 = np.zeros_like(x), np.ones_like(x)
    sm = x <= 1
    xs = x[sm]
    t = (f32(0.25) * xs * xs).astype(f32)
    with np.errstate(divide="ignore", invalid="ignore"):
        lg = ((np.log2(xs).astype(f32) - f32(1)) * f32(0.693147181)).astype(f32)
        i0, s0, i1, s1 = (horner(c[k], t) for k in ("i0", "s0", "i1", "s1"))
        a0 = (f32(4) * t * (s0 - lg * i0)).astype(f32)
        a1 = (f32(1) + t * (f32(2) * lg * i1 - s1)).astype(f32)
    tiny = ~(xs >= f32(8.673617380e-19))
    p0[sm] = np.where(tiny, f32(0), a0)
    p1[sm] = np.where(tiny, f32(1), a1)
    xl = x[~sm]
    u = (f32(1) / xl).astype(f32)
    se = (np.sqrt(xl).astype(f32) * np.exp2(-(xl * f32(1.442695041)).astype(f32)).astype(f32)).astype(f32)
    p1[~sm] = (se * horner(c["f1"], u)).astype(f32)
    p0[~sm] = (xl * se * horner(c["f0"], u)).astype(f32)
    return p0, p1


def test_series_coefficients_are_the_closed_forms():
    c = coefs()
    g = 0.57721566490153286061
    H = np.concatenate([[0.0], np.cumsum(1.0 / np.arange(1, 12))])
    psi = lambda k: -g + H[k - 1]   # noqa: E731  (psi(k), k >= 1)
    want = {"i0": [1 / factorial(k) ** 2 for k in range(6)],
            "s0": [psi(k + 1) / factorial(k) ** 2 for k in range(6)],
            "i1": [1 / (factorial(k) * factorial(k + 1)) for k in range(6)],
            "s1": [(psi(k + 1) + psi(k + 2)) / (factorial(k) * factorial(k + 1)) for k in range(6)]}
    for k, w in want.items():
        assert len(c[k]) == 6, k
        np.testing.assert_allclose(c[k], w, rtol=2e-9, err_msg=k)
    # the first omitted terms are below 1e-9 at t = 1/4 (x = 1)
    assert 0.25 ** 6 / factorial(6) ** 2 < 1e-9 and abs(psi(7)) * 0.25 ** 6 / factorial(6) ** 2 < 1e-9


def test_fitted_coefficients_match_the_fit_script():
    pytest.importorskip("scipy")
    from lddmm_amd import fit_bessel_f32 as F
    c = coefs()
    for order, key in ((0, "f0"), (1, "f1")):
        coef, e_fit, e_f32 = F.fit(order)
        assert e_fit < 1e-8 and e_f32 < 1.2e-7
        np.testing.assert_allclose(c[key], coef.astype(np.float64), rtol=2e-8, atol=1e-10, err_msg=key)


def test_emulated_fp32_factors_against_scipy():
    sp = pytest.importorskip("scipy.special")
    c = coefs()
    x = np.concatenate([[0.0, 1e-30, 1e-19, 1e-12], np.geomspace(1e-8, 1, 20001), np.linspace(1, 90, 40001),
                        np.nextafter(f32(1), f32(2), dtype=f32)[None]]).astype(f32)
    p0, p1 = p01_emulated(x, c)
    xd = x.astype(np.float64)
    with np.errstate(invalid="ignore"):
        r0 = np.where(xd > 0, xd * xd * sp.k0(xd), 0.0)
        r1 = np.where(xd > 0, xd * sp.k1(xd), 1.0)
    # p1(0) = 1 is the scale of both factors (p0 <= 0.6)
    assert np.abs(p0 - r0).max() <= 2e-7 and np.abs(p1 - r1).max() <= 2e-7
    big = r1 > 1e-3
    assert np.max(np.abs(p1[big] / r1[big] - 1)) <= 2e-6
    assert p0[0] == 0 and p1[0] == 1 and p0[1] == 0 and p1[1] == 1
