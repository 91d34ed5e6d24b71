"""Writes the golden vectors in this folder from the reference's own kernel
sources, compiled in place by oracle/Makefile (`make -C oracle ref`) into
oracle/_ref/libref{2,3}d.so, and from the reference's example data files.

Run in the build container (the reference tree is not on the GPU box):
    python tests/golden/make_golden.py

Inputs are seeded (lddmm_amd.geometry's splitmix64 stream); every fixture
stores its inputs and the reference's outputs, nothing else.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from lddmm_amd import geometry as G  # noqa: E402


def pts(seed, n, d, scale=0.4):
    return (scale * G.normal(seed, n * d)).reshape(n, d)


def kernels():
    out = {}
    for dim in (2, 3):
        L = oracle.ref(dim)
        for order in (-1, 0, 1, 2, 3, 4):
            tag = "d%d_o%d" % (dim, order if order >= 0 else 9)
            n, m = 37, 29
            q = pts(100 + dim * 10 + order, n, dim)
            qj = pts(200 + dim * 10 + order, m, dim)
            a = pts(300 + dim * 10 + order, n, dim, 1.0)
            aj = pts(400 + dim * 10 + order, m, dim, 1.0)
            l = pts(500 + dim * 10 + order, n, dim, 1.0)
            w = 0.3 if dim == 3 else 0.25
            v = np.empty_like(q)
            L.ref_multiply_kernel(v, q, a, order, w, n, dim)
            vc = np.empty_like(q)
            L.ref_multiply_kernel_cross(vc, q, qj, aj, order, w, n, m, dim)
            g = np.empty_like(q)
            L.ref_dq_kernel(g, q, l, a, order, w, n, dim)
            gi, gj = np.empty_like(q), np.empty_like(qj)
            L.ref_dq_kernel_cross(gi, gj, q, qj, l, aj, order, w, n, m, dim)
            K = np.empty((n, n))
            L.ref_compute_kernel(K, q, order, w, n, dim)
            Kc = np.empty((m, n))
            L.ref_compute_kernel_cross(Kc, q, qj, order, w, n, m, dim)
            out.update({tag + "_q": q, tag + "_qj": qj, tag + "_a": a, tag + "_aj": aj, tag + "_l": l,
                        tag + "_w": np.array(w), tag + "_mult": v, tag + "_mult_cross": vc, tag + "_dq": g,
                        tag + "_dqi": gi, tag + "_dqj": gj, tag + "_K": K, tag + "_K_cross": Kc})
    xs = np.concatenate([[0.0, 1e-12, 1e-6, 0.01, 0.3, 0.999, 1.0, 1.001, 1.7, 2.0, 2.5, 5.0, 12.0, 40.0, 150.0,
                          650.0, 699.0, 701.0]])
    p0, p1 = np.empty_like(xs), np.empty_like(xs)
    L = oracle.ref(2)
    for k, x in enumerate(xs):
        a, b = C.c_double(), C.c_double()
        L.ref_p01(float(x), C.byref(a), C.byref(b))
        p0[k], p1[k] = a.value, b.value
    out.update(bessel_x=xs, bessel_p0=p0, bessel_p1=p1)
    np.savez_compressed(os.path.join(HERE, "kernels.npz"), **out)


def varifolds():
    out = {}
    # 2D: jittered circle against the square polyline
    q2, e2 = G.circle_curve(48)
    q2 = q2 + 0.02 * G.normal(11, q2.size).reshape(q2.shape)
    tc2, tu2, ta2 = G.polyline_varifold(G.square_polyline(40, 0.9))
    # 3D: jittered icosphere-1 against a superquadric
    q3, e3 = G.icosphere(1)
    q3 = q3 + 0.03 * G.normal(12, q3.size).reshape(q3.shape)
    sq, st = G.superquadric(2, 5)
    tc3, tu3, ta3 = G.triangle_varifold(sq, st)
    cases = [(2, q2, e2, tc2, tu2, ta2), (3, q3, e3, tc3, tu3, ta3)]
    for dim, q, e, tc, tu, ta in cases:
        L = oracle.ref(dim)
        ifo = oracle.get_elm_ifo(e, q.shape[0])
        out["d%d_q" % dim], out["d%d_elm" % dim], out["d%d_ifo" % dim] = q, e.astype(np.int32), ifo
        out["d%d_tc" % dim], out["d%d_tu" % dim], out["d%d_ta" % dim] = tc, tu, ta
        for ct in "GC":
            for dt in "BLOU":
                cw, dw = (0.1 if dim == 2 else 0.3), 0.8
                val, dq = oracle.varifold(q, e, ifo, tc, tu, ta, ct, cw, dt, dw, grad=True, L=L, prefix="ref")
                vv = oracle.varifold(q, e, ifo, tc, tu, ta, ct, cw, dt, dw, grad=False, L=L, prefix="ref")
                tag = "d%d_%s%s" % (dim, ct, dt)
                out[tag + "_val"], out[tag + "_dq"], out[tag + "_valonly"] = np.array(val), dq, np.array(vv)
                out[tag + "_w"] = np.array([cw, dw])
    np.savez_compressed(os.path.join(HERE, "varifold.npz"), **out)


def example_sphere2cube():
    """The reference's own sphere->cube example data (Cpp/examples/sphere2cube/
    sphereData.mat, cubeData.mat; identical copies under CUDA/ and MATLAB/),
    read with scipy.io.loadmat (plain arrays, nothing executed), plus the
    compiled reference's varifold value and gradient at the source with the
    example's matching parameters ('C' 0.1 x 'B', run_sphere2cube.m:11-19)."""
    import scipy.io as sio
    src = os.path.join(REF_ROOT, "Cpp", "examples", "sphere2cube")
    out = {}
    for name in ("sphereData.mat", "cubeData.mat"):
        d = sio.loadmat(os.path.join(src, name))
        out.update({k: np.asarray(v) for k, v in d.items() if not k.startswith("__")})
    q = np.ascontiguousarray(out["iniNdeMat"].T)
    e = np.ascontiguousarray(out["elmVtxMat"].T.astype(np.int32) - 1)
    ifo = oracle.get_elm_ifo(e, q.shape[0])
    tc = np.ascontiguousarray(out["tgtCenPosMat"].T)
    tu = np.ascontiguousarray(out["tgtUniDirMat"].T)
    ta = np.ascontiguousarray(out["tgtElmVolVec"].ravel())
    val, dq = oracle.varifold(q, e, ifo, tc, tu, ta, "C", 0.1, "B", 0.0, grad=True, L=oracle.ref(3), prefix="ref")
    out["ref_vf_val"], out["ref_vf_dq"] = np.array(val), dq
    np.savez_compressed(os.path.join(HERE, "sphere2cube.npz"), **out)


REF_ROOT = "/root/reference"

if __name__ == "__main__":
    oracle.build()
    if oracle.ref(2) is None or oracle.ref(3) is None:
        raise SystemExit("oracle/_ref was not built (needs the reference tree)")
    kernels()
    varifolds()
    example_sphere2cube()
    print("golden vectors written to", HERE)
