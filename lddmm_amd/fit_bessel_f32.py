"""Coefficients of the FP32 Bessel factors of the 2D Matern kernels.

The FP32 policy evaluates p0 = x^2 K0(x) and p1 = x K1(x) (pair_driver.hpp,
bessel_p01_f32) for x > 1 from the exponentially scaled functions
    f_n(u) = sqrt(x) e^x K_n(x),   u = 1/x in (0, 1],   f_n(0) = sqrt(pi/2),
each a polynomial of degree 10 in u.  This script makes those polynomials:
a Chebyshev least-squares fit on [0, 1] with Lawson's iterative reweighting
toward the minimax fit (relative error), reference values from scipy's
k0e/k1e.  It prints the coefficients (power basis, FP32-rounded) and the
relative error of the fit and of its FP32 Horner evaluation.

    python -m lddmm_amd.fit_bessel_f32
"""
import numpy as np
from numpy.polynomial import chebyshev as C

DEGREE = 10


def fit(order, degree=DEGREE, samples=200001, sweeps=80):
    from scipy.special import k0e, k1e
    fn = k0e if order == 0 else k1e
    u = np.linspace(0.0, 1.0, samples)
    x = 1.0 / np.maximum(u, 1e-300)
    f = np.where(u > 0, np.sqrt(x) * fn(x), np.sqrt(np.pi / 2))
    w = np.ones_like(u)
    best = None
    for _ in range(sweeps):
        c = C.chebfit(2 * u - 1, f, degree, w=w)
        r = (C.chebval(2 * u - 1, c) - f) / f
        m = np.abs(r).max()
        if best is None or m < best[0]:
            best = (m, c)
        w = w * (1 + np.abs(r) / m)
    p = np.polynomial.Polynomial(C.cheb2poly(best[1]))(np.polynomial.Polynomial([-1.0, 2.0]))
    coef = p.coef.astype(np.float32)
    uf = u.astype(np.float32)
    acc = np.full_like(uf, coef[-1])
    for a in coef[-2::-1]:
        acc = (acc * uf + a).astype(np.float32)
    return coef, best[0], float(np.abs(acc / f - 1).max())


if __name__ == "__main__":
    for order in (0, 1):
        coef, e_fit, e_f32 = fit(order)
        print("f%d: fit %.2e, FP32 Horner %.2e" % (order, e_fit, e_f32))
        print("    " + ", ".join("%.9ef" % c for c in coef))
