"""GPU parity of every operator and of the objective against the FP64 CPU
oracle, through the C-ABI (lddmm_amd is a ctypes mirror of it).

Tolerance for the FP64 kernels: 1e-12 relative (max-norm) for the pair
operators and 1e-10 for objective energy and gradient, the bound the
design sets for FP64; differences come only from summation order and FMA
contraction."""
import numpy as np
import pytest

import oracle
from conftest import rel_err
from lddmm_amd import geometry as G

pytestmark = pytest.mark.gpu

TOL_OP = 1e-12
TOL_OBJ = 1e-10


def pts(seed, n, d, scale=0.5):
    return (scale * G.normal(seed, n * d)).reshape(n, d)


CASES = [(d, o) for d in (2, 3) for o in (-1, 0, 1, 2, 3, 4)]


@pytest.mark.parametrize("dim,order", CASES)
@pytest.mark.parametrize("n", [1, 7, 300, 1237])
def test_multiply_kernel(gpu, dim, order, n):
    q, a = pts(n + 1, n, dim), pts(n + 2, n, dim, 1.0)
    w = 0.3
    print("multiply_kernel dim=%d order=%d n=%d" % (dim, order, n), flush=True)
    assert rel_err(gpu.multiplyKernel(q, a, order, w), oracle.multiply_kernel(q, a, order, w)) <= TOL_OP


@pytest.mark.parametrize("dim,order", CASES)
@pytest.mark.parametrize("ni,nj", [(5, 1), (300, 77), (64, 2000)])
def test_multiply_kernel_cross(gpu, dim, order, ni, nj):
    qi, qj, a = pts(3, ni, dim), pts(4, nj, dim), pts(5, nj, dim, 1.0)
    print("multiply_kernel_cross dim=%d order=%d %d x %d" % (dim, order, ni, nj), flush=True)
    assert rel_err(gpu.multiplyKernel(qi, qj, a, order, 0.4),
                   oracle.multiply_kernel_cross(qi, qj, a, order, 0.4)) <= TOL_OP


@pytest.mark.parametrize("dim,order", CASES)
@pytest.mark.parametrize("n", [2, 300, 1237])
def test_dq_kernel(gpu, dim, order, n):
    q, l, r = pts(n + 6, n, dim), pts(n + 7, n, dim, 1.0), pts(n + 8, n, dim, 1.0)
    print("dq_kernel dim=%d order=%d n=%d" % (dim, order, n), flush=True)
    got = gpu.dqKernel(q, l, r, order, 0.3)
    exp = oracle.dq_kernel(q, l, r, order, 0.3)
    if order == 0:
        assert not np.any(got)
    else:
        assert rel_err(got, exp) <= TOL_OP


@pytest.mark.parametrize("dim,order", [(2, -1), (2, 2), (2, 4), (3, -1), (3, 1), (3, 4)])
def test_dq_kernel_cross(gpu, dim, order):
    qi, qj = pts(9, 250, dim), pts(10, 333, dim)
    l, r = pts(11, 250, dim, 1.0), pts(12, 333, dim, 1.0)
    print("dq_kernel_cross dim=%d order=%d" % (dim, order), flush=True)
    gi, gj = gpu.dqKernel(qi, qj, l, r, order, 0.35)
    ei, ej = oracle.dq_kernel_cross(qi, qj, l, r, order, 0.35)
    assert rel_err(gi, ei) <= TOL_OP and rel_err(gj, ej) <= TOL_OP


def coincident_points(dim):
    """Points with exact repeats (r2 = 0 between distinct indices) and a
    cluster at the origin whose partners lie 1e-120, 1e-155 and 1e-160 away:
    r2 = 1e-240 (below 2^-767), 1e-310 and 1e-320 (subnormal), the range
    where the FP64 Matern distance adds 1e-300 to r2 instead of scaling it."""
    q = pts(21, 40, dim)
    tiny = np.zeros((4, dim))
    tiny[1, 0], tiny[2, 0], tiny[3, dim - 1] = 1e-120, 1e-160, 1e-155
    return np.concatenate([q, q[:12], tiny])


def test_coincident_points_reach_tiny_r2():
    """The fixture really holds squared distances in (0, 2^-767), subnormal
    ones included (0.5 N(0,1) coordinates plus 1e-120 would round away)."""
    for dim in (2, 3):
        q = coincident_points(dim)
        r2 = np.sum((q[:, None, :] - q[None, :, :]) ** 2, axis=-1)
        pos = r2[r2 > 0]
        assert np.any(pos < 2.0 ** -767) and np.any(pos < np.finfo(float).tiny)
        assert np.any(r2[~np.eye(len(q), dtype=bool)] == 0)


@pytest.mark.parametrize("dim,order", CASES)
def test_coincident_points(gpu, dim, order):
    q = coincident_points(dim)
    n = len(q)
    a, l, r = pts(22, n, dim, 1.0), pts(23, n, dim, 1.0), pts(24, n, dim, 1.0)
    print("coincident points dim=%d order=%d" % (dim, order), flush=True)
    assert rel_err(gpu.multiplyKernel(q, a, order, 0.3), oracle.multiply_kernel(q, a, order, 0.3)) <= TOL_OP
    assert rel_err(gpu.computeKernel(q, order, 0.3), oracle.compute_kernel(q, order, 0.3)) <= TOL_OP
    got = gpu.dqKernel(q, l, r, order, 0.3)
    if order == 0:
        assert not np.any(got)
    else:
        assert rel_err(got, oracle.dq_kernel(q, l, r, order, 0.3)) <= TOL_OP


@pytest.mark.parametrize("dim,order", [(2, -1), (2, 0), (2, 3), (3, -1), (3, 2)])
def test_compute_kernel(gpu, dim, order):
    q, qj = pts(13, 90, dim), pts(14, 41, dim)
    print("compute_kernel dim=%d order=%d" % (dim, order), flush=True)
    assert rel_err(gpu.computeKernel(q, order, 0.3), oracle.compute_kernel(q, order, 0.3)) <= TOL_OP
    assert rel_err(gpu.computeKernel(q, qj, order, 0.3), oracle.compute_kernel_cross(q, qj, order, 0.3)) <= TOL_OP


def test_golden_kernels_on_gpu(gpu, golden_kernels):
    """The GPU operators against the reference's own outputs directly."""
    g = golden_kernels
    for dim, order in CASES:
        t = "d%d_o%d" % (dim, order if order >= 0 else 9)
        q, qj, a, aj, l, w = g[t + "_q"], g[t + "_qj"], g[t + "_a"], g[t + "_aj"], g[t + "_l"], float(g[t + "_w"])
        assert rel_err(gpu.multiplyKernel(q, a, order, w), g[t + "_mult"]) <= TOL_OP
        assert rel_err(gpu.multiplyKernel(q, qj, aj, order, w), g[t + "_mult_cross"]) <= TOL_OP
        if order != 0:
            assert rel_err(gpu.dqKernel(q, l, a, order, w), g[t + "_dq"]) <= TOL_OP


@pytest.mark.parametrize("order", [0, 1, 2, 3, 4])
def test_2d_matern_over_the_whole_distance_range(gpu, order):
    """2D Matern kernel values at distances d = |q_i - q_j|/sigma from 0 and
    below machine epsilon through both sides of the series / continued
    fraction switch (d = 2) to past the cut-off (d = 700), against the oracle
    (pinned to the reference's polybesselk.cpp and scipy's k0/k1)."""
    w = 0.3
    d = np.concatenate([[0.0, 1e-17, 1e-12, 1e-6], np.logspace(-4, 2.9, 200), [1.999999, 2.0, 2.000001],
                        [650.0, 699.9, 700.0, 700.1, 800.0]])
    qi = np.zeros((1, 2))
    qj = np.stack([d * w, np.zeros_like(d)], axis=1)
    got = gpu.computeKernel(qi, qj, order, w)
    exp = oracle.compute_kernel_cross(qi, qj, order, w)
    print("2d matern order %d: max rel %.2e" % (order, np.max(np.abs(got - exp) / np.maximum(np.abs(exp), 1e-300))),
          flush=True)
    # relative agreement down to the subnormal range (e^-700 ~ 1e-304 at the cut-off)
    assert np.all(np.abs(got - exp) <= 1e-13 * np.abs(exp) + 1e-300)
    assert np.all(got[d > 700.0] == 0.0) and got[0] == exp[0]


@pytest.mark.parametrize("dim", (2, 3))
@pytest.mark.parametrize("ct", "GC")
@pytest.mark.parametrize("dt", "BLOU")
def test_varifold_golden(gpu, golden_varifold, dim, ct, dt):
    g = golden_varifold
    d = "d%d" % dim
    t = "%s_%s%s" % (d, ct, dt)
    cw, dw = g[t + "_w"]
    print("varifold", t, flush=True)
    val, dq = gpu.varifold(g[d + "_q"], g[d + "_elm"], g[d + "_ifo"], g[d + "_tc"], g[d + "_tu"], g[d + "_ta"],
                           ct, cw, dt, dw)
    assert abs(val - float(g[t + "_val"])) <= TOL_OP * max(1.0, abs(float(g[t + "_val"])))
    assert rel_err(dq, g[t + "_dq"]) <= 1e-11
    # value-only mode (the MATLAB varifold MEX with one output)
    vv = gpu.varifold(g[d + "_q"], g[d + "_elm"], g[d + "_ifo"], g[d + "_tc"], g[d + "_tu"], g[d + "_ta"],
                      ct, cw, dt, dw, grad=False)
    assert abs(vv - float(g[t + "_valonly"])) <= TOL_OP * max(1.0, abs(float(g[t + "_valonly"])))


def test_varifold_larger_and_without_target(gpu):
    q, tri = G.icosphere(3)
    q = q + 0.01 * G.normal(3, q.size).reshape(q.shape)
    ifo = oracle.get_elm_ifo(tri, q.shape[0])
    cv, ct = G.cube_surface(12)
    tc, tu, ta = G.triangle_varifold(cv, ct)
    print("varifold ico3 vs cube12", flush=True)
    v, dq = gpu.varifold(q, tri, ifo, tc, tu, ta, "G", 0.15, "B", 0.0)
    ev, edq = oracle.varifold(q, tri, ifo, tc, tu, ta, "G", 0.15, "B", 0.0)
    assert abs(v - ev) <= 1e-11 * abs(ev) and rel_err(dq, edq) <= 1e-11
    empty = np.zeros((0, 3))
    v, dq = gpu.varifold(q, tri, ifo, empty, empty, np.zeros(0), "C", 0.2, "U", 0.6)
    ev, edq = oracle.varifold(q, tri, ifo, empty, empty, np.zeros(0), "C", 0.2, "U", 0.6)
    assert abs(v - ev) <= 1e-11 * abs(ev) and rel_err(dq, edq) <= 1e-11


def make_problem(kind):
    if kind.startswith("circle_square"):
        w = G.circle_square(200)
        w["q0"] = w["q0"] * 0.95
        if kind.startswith("circle_square_matern"):
            # 2D Matern nu = p + 1 (Bessel factors), order p = 2 unless the kind names it
            w["knl_order"] = int(kind[-1]) if kind[-1].isdigit() else 2
            if kind.endswith("example3"):
                # the reference's circle->square example: order 3, width 0.025, 'C' x 'B'
                # (Cpp/examples/circle2square/run_circle2square.m:31-39)
                w["knl_width"], w["cen_type"], w["cen_width"] = 0.025, "C", 0.1
    elif kind == "sphere_cube":
        w = G.sphere_cube(subdivisions=2, cells=10)
    else:
        w = G.superquadric_pair(4, 17)
    w["time_num"] = 6
    w["time_stp"] = 0.2
    return w


@pytest.mark.parametrize("kind", ["circle_square", "circle_square_matern", "sphere_cube", "superquadric"])
def test_objgrd_matches_oracle(gpu, kind):
    w = make_problem(kind)
    n, d = w["q0"].shape
    ifo = oracle.get_elm_ifo(w["elm"], n)
    args = (w["q0"], w["elm"], ifo, w["tgt_cen"], w["tgt_vol"], w["tgt_dir"], w["cen_type"], w["cen_width"],
            w["dir_type"], w["dir_width"], w["dis_wgt"], w["knl_order"], w["knl_width"], w["time_stp"], w["time_num"])
    obj = gpu.Objective(*args)
    P = oracle.OracleProblem(w["q0"], w["elm"], w["tgt_cen"], w["tgt_dir"], w["tgt_vol"], w["cen_type"],
                             w["cen_width"], w["dir_type"], w["dir_width"], w["dis_wgt"], w["knl_order"],
                             w["knl_width"], w["time_stp"], w["time_num"], ifo=ifo)
    print("objgrd", kind, "n=%d" % n, flush=True)
    for seed, scale in ((0, 0.0), (5, 0.05), (6, 0.3)):
        a = scale * G.normal(seed + 100, obj.varnum)
        f, g = obj(a)
        ef, eg, traj = P.objgrd(a, want_traj=True)
        assert abs(f - ef) <= TOL_OBJ * abs(ef), (seed, f, ef)
        assert rel_err(g, eg) <= TOL_OBJ, (seed, rel_err(g, eg))
        assert rel_err(obj.trajectory(), traj) <= 1e-12
    tt = oracle.varifold_terms(w["q0"], w["elm"], w["tgt_cen"], w["tgt_dir"], w["tgt_vol"], w["cen_type"],
                               w["cen_width"], w["dir_type"], w["dir_width"])[2]
    assert abs(obj.target_term() - tt) <= 1e-12 * abs(tt)
    obj.close()


def test_objgrd_is_deterministic(gpu):
    w = make_problem("sphere_cube")
    n = w["q0"].shape[0]
    ifo = oracle.get_elm_ifo(w["elm"], n)
    obj = gpu.Objective(w["q0"], w["elm"], ifo, w["tgt_cen"], w["tgt_vol"], w["tgt_dir"], w["cen_type"],
                        w["cen_width"], w["dir_type"], w["dir_width"], w["dis_wgt"], w["knl_order"], w["knl_width"],
                        w["time_stp"], w["time_num"])
    a = 0.1 * G.normal(9, obj.varnum)
    f1, g1 = obj(a)
    f2, g2 = obj(a)
    assert f1 == f2 and np.array_equal(g1, g2)


def test_compute_node_position_matches_trajectory(gpu):
    w = make_problem("circle_square")
    n, d = w["q0"].shape
    a = 0.2 * G.normal(44, d * n * (w["time_num"] - 1)).reshape(w["time_num"] - 1, n, d)
    nde, vlc = gpu.computeNodePosition(w["q0"], a, w["knl_order"], w["knl_width"], w["time_stp"], w["time_num"])
    q = w["q0"].copy()
    assert np.array_equal(nde[0], q)
    for t in range(w["time_num"] - 1):
        v = oracle.multiply_kernel(q, a[t], w["knl_order"], w["knl_width"])
        assert rel_err(vlc[t], v) <= 1e-12
        q = q + w["time_stp"] * v
        assert rel_err(nde[t + 1], q) <= 1e-12


@pytest.mark.parametrize("kind", ["circle_square", "circle_square_matern", "circle_square_matern0",
                                  "circle_square_matern1", "circle_square_matern3", "circle_square_matern4",
                                  "circle_square_matern_example3", "sphere_cube", "superquadric"])
def test_objgrd_fp32_policy_error(gpu, kind):
    """FP32 pair arithmetic (FP64 differences and cross-tile sums) against
    the FP64 oracle: relative error of the energy and of the gradient
    (max norm) stated and bounded."""
    w = make_problem(kind)
    n = w["q0"].shape[0]
    ifo = oracle.get_elm_ifo(w["elm"], n)
    args = (w["q0"], w["elm"], ifo, w["tgt_cen"], w["tgt_vol"], w["tgt_dir"], w["cen_type"], w["cen_width"],
            w["dir_type"], w["dir_width"], w["dis_wgt"], w["knl_order"], w["knl_width"], w["time_stp"], w["time_num"])
    obj = gpu.Objective(*args, precision="fp32")
    P = oracle.OracleProblem(w["q0"], w["elm"], w["tgt_cen"], w["tgt_dir"], w["tgt_vol"], w["cen_type"],
                             w["cen_width"], w["dir_type"], w["dir_width"], w["dis_wgt"], w["knl_order"],
                             w["knl_width"], w["time_stp"], w["time_num"], ifo=ifo)
    for seed, scale in ((0, 0.0), (6, 0.3)):
        a = scale * G.normal(seed + 100, obj.varnum)
        f, g = obj(a)
        ef, eg = P.objgrd(a)
        e_rel, g_rel = abs(f - ef) / abs(ef), rel_err(g, eg)
        print("fp32 %s scale %.2f: energy rel err %.3e, gradient rel err %.3e" % (kind, scale, e_rel, g_rel),
              flush=True)
        assert e_rel <= 1e-6 and g_rel <= 1e-4
    obj.close()
