"""The CPU oracle against the reference's own kernels.

Golden vectors in tests/golden were produced by the reference's kernel
sources (Cpp/src/{computeKernel,multiplyKernel,dqKernel,polybesselk,
varifold2D,varifold3D}.cpp) compiled in place (oracle/Makefile `ref`).
When oracle/_ref is present (the build container) the oracle is also
compared with it directly on fresh inputs.
"""
import numpy as np
import pytest

import oracle
from conftest import rel_err

ORDERS = (-1, 0, 1, 2, 3, 4)


def _tag(dim, order):
    return "d%d_o%d" % (dim, order if order >= 0 else 9)


@pytest.mark.parametrize("dim", (2, 3))
@pytest.mark.parametrize("order", ORDERS)
def test_pair_operators_match_reference(golden_kernels, dim, order):
    g = golden_kernels
    t = _tag(dim, order)
    q, qj, a, aj, l, w = g[t + "_q"], g[t + "_qj"], g[t + "_a"], g[t + "_aj"], g[t + "_l"], float(g[t + "_w"])
    tol = 1e-14 if dim == 3 or order == -1 else 1e-13   # 2D Matern: independent Bessel evaluation
    assert rel_err(oracle.multiply_kernel(q, a, order, w), g[t + "_mult"]) <= tol
    assert rel_err(oracle.multiply_kernel_cross(q, qj, aj, order, w), g[t + "_mult_cross"]) <= tol
    assert rel_err(oracle.compute_kernel(q, order, w), g[t + "_K"]) <= tol
    assert rel_err(oracle.compute_kernel_cross(q, qj, order, w), g[t + "_K_cross"]) <= tol
    dq = oracle.dq_kernel(q, l, a, order, w)
    dqi, dqj = oracle.dq_kernel_cross(q, qj, l, aj, order, w)
    if order == 0:   # no derivative for order 0 in the reference: zeros
        assert not np.any(g[t + "_dq"]) and not np.any(dq)
        assert not np.any(dqi) and not np.any(dqj)
    else:
        assert rel_err(dq, g[t + "_dq"]) <= tol
        assert rel_err(dqi, g[t + "_dqi"]) <= tol
        assert rel_err(dqj, g[t + "_dqj"]) <= tol


def test_bessel_factors_match_reference(golden_kernels):
    g = golden_kernels
    for x, p0, p1 in zip(g["bessel_x"], g["bessel_p0"], g["bessel_p1"]):
        o0, o1 = oracle.bessel_p01(x)
        assert abs(o0 - p0) <= 2e-15 * max(abs(p0), 1e-300) + 1e-300
        assert abs(o1 - p1) <= 2e-15 * max(abs(p1), 1e-300) + 1e-300


def test_bessel_factors_match_scipy():
    sp = pytest.importorskip("scipy.special")
    for x in np.concatenate([np.geomspace(1e-8, 1.0, 40), np.linspace(1.0, 3.0, 41), np.geomspace(3.0, 600, 40)]):
        p0, p1 = oracle.bessel_p01(x)
        e0, e1 = x * x * sp.k0(x), x * sp.k1(x)
        assert abs(p0 - e0) <= 1e-14 * e0
        assert abs(p1 - e1) <= 1e-14 * e1


@pytest.mark.parametrize("dim", (2, 3))
@pytest.mark.parametrize("ct", "GC")
@pytest.mark.parametrize("dt", "BLOU")
def test_varifold_matches_reference(golden_varifold, dim, ct, dt):
    g = golden_varifold
    d = "d%d" % dim
    t = "%s_%s%s" % (d, ct, dt)
    cw, dw = g[t + "_w"]
    args = (g[d + "_q"], g[d + "_elm"], g[d + "_ifo"], g[d + "_tc"], g[d + "_tu"], g[d + "_ta"], ct, cw, dt, dw)
    val, dq = oracle.varifold(*args, grad=True)
    assert abs(val - float(g[t + "_val"])) <= 1e-13 * abs(float(g[t + "_val"]))
    assert rel_err(dq, g[t + "_dq"]) <= 1e-13
    vv = oracle.varifold(*args, grad=False)
    assert abs(vv - float(g[t + "_valonly"])) <= 1e-13 * abs(float(g[t + "_valonly"]))


@pytest.mark.skipif(oracle.ref(3) is None, reason="reference kernels not built here (no reference tree)")
@pytest.mark.parametrize("dim", (2, 3))
def test_oracle_matches_compiled_reference_fresh_inputs(dim):
    from lddmm_amd import geometry as G
    L = oracle.ref(dim)
    for seed, n in ((1, 1), (2, 7), (3, 131)):
        q = (0.5 * G.normal(seed, n * dim)).reshape(n, dim)
        a = G.normal(seed + 50, n * dim).reshape(n, dim)
        l = G.normal(seed + 90, n * dim).reshape(n, dim)
        for order in (-1, 1, 3) if dim == 3 else (-1, 2):
            v = np.empty_like(q)
            L.ref_multiply_kernel(v, q, a, order, 0.35, n, dim)
            assert rel_err(oracle.multiply_kernel(q, a, order, 0.35), v) <= 1e-13
            g = np.empty_like(q)
            L.ref_dq_kernel(g, q, l, a, order, 0.35, n, dim)
            if n > 1:
                assert rel_err(oracle.dq_kernel(q, l, a, order, 0.35), g) <= 1e-13
