"""Pins of the CPU oracle by what the mathematics fixes: closed forms,
finite differences (the discrete adjoint is the exact gradient of the
discrete objective), invariants, a dense BFGS update and crafted line-search
objectives.  These cover the parts of the reference that cannot be compiled
here (Cpp/src/objgrd.cpp, LBFGS.cpp, getDirection.cpp, lineSearch.cpp need
MATLAB's blas.h)."""
import numpy as np
import pytest

import oracle
from conftest import rel_err
from lddmm_amd import geometry as G


def small_2d(n=16, seed=3):
    q, e = G.circle_curve(n)
    q = q + 0.05 * G.normal(seed, q.size).reshape(q.shape)
    tc, tu, ta = G.polyline_varifold(G.square_polyline(12, 0.8))
    return q, e, tc, tu, ta


def small_3d(seed=4):
    q, tri = G.icosphere(0)
    q = q + 0.05 * G.normal(seed, q.size).reshape(q.shape)
    sq, st = G.superquadric(1, 9)
    tc, tu, ta = G.triangle_varifold(sq, st)
    return q, tri, tc, tu, ta


def problem(dim, order, ct="G", dt="B", T=4, w=10.0, width=0.4):
    q, e, tc, tu, ta = small_2d() if dim == 2 else small_3d()
    return oracle.OracleProblem(q, e, tc, tu, ta, ct, 0.3, dt, 0.7, w, order, width, 1.0 / (T - 1), T)


def test_multiply_kernel_brute_force():
    q = G.normal(1, 60).reshape(20, 3)
    a = G.normal(2, 60).reshape(20, 3)
    d2 = ((q[:, None, :] - q[None, :, :]) ** 2).sum(-1)
    K = np.exp(-d2 / 0.7 ** 2)
    assert rel_err(oracle.multiply_kernel(q, a, -1, 0.7), K @ a) <= 1e-14
    d = np.sqrt(d2) / 0.7
    K1 = (1 + d) * np.exp(-d)
    assert rel_err(oracle.multiply_kernel(q, a, 1, 0.7), K1 @ a) <= 1e-14


@pytest.mark.parametrize("dim,order", [(2, -1), (3, -1), (3, 1), (3, 2), (3, 4), (2, 3)])
def test_dq_kernel_is_gradient_of_bilinear_form(dim, order):
    n = 9
    q = (0.3 * G.normal(5, n * dim)).reshape(n, dim)
    l = G.normal(6, n * dim).reshape(n, dim)
    r = G.normal(7, n * dim).reshape(n, dim)
    w = 0.5

    def form(x):
        return float(np.sum(l * oracle.multiply_kernel(x, r, order, w)))

    g = oracle.dq_kernel(q, l, r, order, w)
    u = G.normal(8, n * dim).reshape(n, dim)
    eps = 1e-6
    fd = (form(q + eps * u) - form(q - eps * u)) / (2 * eps)
    assert abs(np.sum(g * u) - fd) <= 1e-7 * max(1.0, abs(fd))


@pytest.mark.parametrize("dim,order,ct,dt", [(2, -1, "G", "B"), (2, -1, "C", "L"), (3, -1, "G", "B"),
                                             (3, 1, "C", "B"), (3, 3, "G", "O"), (3, 2, "G", "U")])
def test_objgrd_matches_finite_differences(dim, order, ct, dt):
    P = problem(dim, order, ct, dt)
    a = 0.3 * G.normal(21, P.varnum)
    f, g = P.objgrd(a)
    for k in range(3):
        u = G.normal(30 + k, P.varnum)
        eps = 1e-6
        fp, _ = P.objgrd(a + eps * u)
        fm, _ = P.objgrd(a - eps * u)
        fd = (fp - fm) / (2 * eps)
        assert abs(np.dot(g, u) - fd) <= 1e-6 * max(1.0, abs(fd)), (k, np.dot(g, u), fd)


def test_objgrd_zero_momentum_closed_form():
    for dim in (2, 3):
        P = problem(dim, -1)
        f, g = P.objgrd(np.zeros(P.varnum))
        v = oracle.varifold(P.q0, P.elm, P.ifo, P.tc, P.tu, P.ta, "G", 0.3, "B", 0.7, grad=False)
        assert abs(f - 10.0 * v) <= 1e-13 * abs(f)
        # at alpha = 0 the flow is the identity: g_t = h K(q0) (0 - p) with p = -w dV, the same for all t
        _, dq = oracle.varifold(P.q0, P.elm, P.ifo, P.tc, P.tu, P.ta, "G", 0.3, "B", 0.7)
        h = 1.0 / (P.T - 1)
        g_last = g.reshape(P.T - 1, P.n, P.dim)[-1]
        expect = h * oracle.multiply_kernel(P.q0, 10.0 * dq, -1, 0.4)
        assert rel_err(g_last, expect) <= 1e-12


@pytest.mark.parametrize("dim", (2, 3))
@pytest.mark.parametrize("ct", "GC")
@pytest.mark.parametrize("dt", "BLOU")
def test_varifold_gradient_finite_differences(dim, ct, dt):
    q, e, tc, tu, ta = small_2d() if dim == 2 else small_3d()
    ifo = oracle.get_elm_ifo(e, q.shape[0])
    args = (e, ifo, tc, tu, ta, ct, 0.3, dt, 0.7)
    v, dq = oracle.varifold(q, *args)
    u = G.normal(41, q.size).reshape(q.shape)
    eps = 1e-6
    fd = (oracle.varifold(q + eps * u, *args, grad=False) - oracle.varifold(q - eps * u, *args, grad=False)) / (2 * eps)
    assert abs(np.sum(dq * u) - fd) <= 1e-7 * max(1.0, abs(fd))


def test_varifold_vanishes_on_its_own_target():
    q, tri, *_ = small_3d()
    c, u, a = G.triangle_varifold(q, tri)
    ifo = oracle.get_elm_ifo(tri, q.shape[0])
    v, dq = oracle.varifold(q, tri, ifo, c, u, a, "G", 0.3, "B", 0.0)
    assert abs(v) <= 1e-13 and np.max(np.abs(dq)) <= 1e-12


def test_varifold_translation_invariance_without_target():
    q, tri, *_ = small_3d()
    ifo = oracle.get_elm_ifo(tri, q.shape[0])
    empty = np.zeros((0, 3))
    v, dq = oracle.varifold(q, tri, ifo, empty, empty, np.zeros(0), "C", 0.3, "U", 0.7)
    assert np.max(np.abs(dq.sum(0))) <= 1e-12 * np.max(np.abs(dq))
    v2 = oracle.varifold(q + 0.37, tri, ifo, empty, empty, np.zeros(0), "C", 0.3, "U", 0.7, grad=False)
    assert abs(v - v2) <= 1e-12 * abs(v)


def test_unoriented_kernels_ignore_target_orientation():
    q, tri, tc, tu, ta = small_3d()
    ifo = oracle.get_elm_ifo(tri, q.shape[0])
    flip = np.where(np.arange(len(ta)) % 3 == 0, -1.0, 1.0)[:, None]
    for dt in "BU":
        v1 = oracle.varifold(q, tri, ifo, tc, tu, ta, "G", 0.3, dt, 0.7, grad=False)
        v2 = oracle.varifold(q, tri, ifo, tc, tu * flip, ta, "G", 0.3, dt, 0.7, grad=False)
        assert abs(v1 - v2) <= 1e-13 * abs(v1)
    v1 = oracle.varifold(q, tri, ifo, tc, tu, ta, "G", 0.3, "L", 0.7, grad=False)
    v2 = oracle.varifold(q, tri, ifo, tc, tu * flip, ta, "G", 0.3, "L", 0.7, grad=False)
    assert abs(v1 - v2) > 1e-6


def _get_elm_ifo_by_definition(elm0, nq):
    """Direct reading of Cpp/src/getElmIfo.m: for each vertex, the elements
    containing it in ascending order, with (elm, sign) in 2D and
    (elm, v3, v2) / (elm, v1, v3) / (elm, v2, v1) in 3D."""
    vpe = elm0.shape[1]
    cols = []
    for n in range(nq):
        col = []
        for e in range(elm0.shape[0]):
            for k in range(vpe):
                if elm0[e, k] == n:
                    if vpe == 2:
                        col += [e, -1 if k == 0 else 1]
                    else:
                        ev = elm0[e]
                        col += [e] + [[ev[2], ev[1]], [ev[0], ev[2]], [ev[1], ev[0]]][k]
        cols.append([len(col) // vpe] + col)
    rows = max(len(c) for c in cols)
    out = np.zeros((nq, rows), dtype=np.int32)
    for n, c in enumerate(cols):
        out[n, :len(c)] = c
    return out


def test_get_elm_ifo_matches_definition():
    q, tri = G.icosphere(1)
    assert np.array_equal(oracle.get_elm_ifo(tri, q.shape[0]), _get_elm_ifo_by_definition(tri, q.shape[0]))
    q2, seg = G.circle_curve(11)
    assert np.array_equal(oracle.get_elm_ifo(seg, 11), _get_elm_ifo_by_definition(seg, 11))


def test_two_loop_equals_dense_bfgs_inverse():
    n, m = 6, 5
    rng_s = G.normal(61, n * 4).reshape(4, n)
    A = np.diag(np.linspace(1, 5, n)) + 0.1
    S = np.zeros((m, n))
    Y = np.zeros((m, n))
    for k in range(4):
        S[k] = rng_s[k]
        Y[k] = A @ rng_s[k]
    g = G.normal(62, n)
    new = 3
    gamma = S[new] @ Y[new] / (Y[new] @ Y[new])
    d = oracle.get_direction(gamma, g, S, Y, new, 4, m)
    H = gamma * np.eye(n)
    for k in range(4):
        rho = 1.0 / (Y[k] @ S[k])
        V = np.eye(n) - rho * np.outer(Y[k], S[k])
        H = V.T @ H @ V + rho * np.outer(S[k], S[k])
    assert rel_err(d, -H @ g) <= 1e-12


def test_line_search_codes():
    # 1: not a descent direction
    quad = lambda x: (0.5 * float(x @ x), x.copy())
    x = np.array([1.0, -2.0])
    code, *_ = oracle.line_search(quad, x, x, x, 2.5, 0.0, 0.5)
    assert code == 1
    # 2: unbounded below along d -> the bracket never closes
    c = np.array([1.0, 2.0])
    lin = lambda x: (float(c @ x), c.copy())
    code, *_ = oracle.line_search(lin, x, c, -c, float(c @ x), 0.0, 0.5)
    assert code == 2
    # 3: f and g inconsistent: g claims descent, f rises -> bracketed, never satisfied
    bad = lambda x: (float(x @ x), np.array([-1.0, 0.0]))
    z = np.zeros(2)
    code, *_ = oracle.line_search(bad, z, np.array([-1.0, 0.0]), np.array([1.0, 0.0]), 0.0, 0.5, 0.9)
    assert code == 3
    # 0 on a quadratic: the unit step of the Newton direction satisfies weak Wolfe
    code, xt, gt, ft, step, cnt = oracle.line_search(quad, x, x, -x, 2.5, 1e-4, 0.9)
    assert code == 0 and cnt == 1 and np.allclose(xt, 0.0)


def test_lbfgs_converges_on_rosenbrock():
    def rosen(x):
        f = 100.0 * (x[1] - x[0] ** 2) ** 2 + (1 - x[0]) ** 2
        g = np.array([-400.0 * x[0] * (x[1] - x[0] ** 2) - 2 * (1 - x[0]), 200.0 * (x[1] - x[0] ** 2)])
        return f, g
    x, f, tr = oracle.lbfgs(rosen, np.array([-1.2, 1.0]), itr_max=500, tol=1e-9, wolfe1=1e-4, wolfe2=0.9, mem=5)
    assert tr["status"] == 0 and np.allclose(x, 1.0, atol=1e-6)


def test_lddmm_run_decreases_energy():
    P = problem(2, -1, T=5)
    alp, f, tr = P.run(itr_max=15, tol=0.0, wolfe1=0.0, wolfe2=0.5, mem=5)
    assert tr["iters"] == 15 and tr["status"] == 0
    assert np.all(np.diff(tr["f"]) < 0)
    f2, _ = P.objgrd(alp)
    assert f2 == f


def l2_problem(dim, order, n=11, T=5):
    q = (0.4 * G.normal(61 + dim, n * dim)).reshape(n, dim)
    t = q + (0.2 * G.normal(71 + dim, n * dim)).reshape(n, dim)
    none = np.zeros((0, dim))
    P = oracle.OracleProblem(q, np.zeros((0, dim), dtype=np.int32), none, none, np.zeros(0), "G", 0.3, "B", 0.0,
                             7.0, order, 0.45, 1.0 / (T - 1), T, ifo=np.zeros((n, 1), dtype=np.int32), tgt_nde=t)
    return P, q, t


def test_l2_dist_closed_form():
    q = G.normal(1, 12).reshape(4, 3)
    t = G.normal(2, 12).reshape(4, 3)
    f, g = oracle.l2_dist(q, t, 3.0)
    assert abs(f - 1.5 * np.sum((q - t) ** 2)) <= 1e-15 * f
    assert np.array_equal(g, 3.0 * (q - t))


@pytest.mark.parametrize("dim,order", [(2, -1), (3, -1), (3, 2), (2, 1)])
def test_l2_objective_matches_finite_differences(dim, order):
    """The objective with the l2Dist matching term (MATLAB/src/objgrd.m with
    disHdl = l2Dist): its discrete adjoint is the exact gradient."""
    P, q, t = l2_problem(dim, order)
    a = 0.3 * G.normal(81, P.varnum)
    f, g = P.objgrd(a)
    for k in range(3):
        u = G.normal(90 + k, P.varnum)
        eps = 1e-6
        fp, _ = P.objgrd(a + eps * u)
        fm, _ = P.objgrd(a - eps * u)
        fd = (fp - fm) / (2 * eps)
        assert abs(np.dot(g, u) - fd) <= 1e-6 * max(1.0, abs(fd)), (k, np.dot(g, u), fd)


def test_l2_objective_zero_momentum_closed_form():
    P, q, t = l2_problem(3, -1)
    f, g = P.objgrd(np.zeros(P.varnum))
    assert abs(f - 3.5 * np.sum((q - t) ** 2)) <= 1e-14 * f
    h = 1.0 / (P.T - 1)
    expect = h * oracle.multiply_kernel(q, 7.0 * (q - t), -1, 0.45)
    for s in g.reshape(P.T - 1, P.n, P.dim):
        assert rel_err(s, expect) <= 1e-13


def _spd(seed, n, lo, hi):
    """A symmetric matrix with eigenvalues spread over (lo, hi)."""
    Q, _ = np.linalg.qr(G.normal(seed, n * n).reshape(n, n))
    return Q @ np.diag(np.linspace(lo, hi, n)) @ Q.T


def _dense_bfgs_iterates(A, x0, mem, iters):
    """L-BFGS iterates on f = x^T A x / 2 with every unit step accepted,
    built from the dense inverse-Hessian update: H0 = gamma I with gamma =
    s^T y / y^T y of the newest pair (1 before any pair), then the BFGS
    update H <- V^T H V + rho s s^T applied for the newest min(k-1, mem)
    pairs, oldest first (Nocedal & Wright eq. 7.19)."""
    xs, pairs, x = [x0.copy()], [], x0.copy()
    for _ in range(iters):
        g = A @ x
        use = pairs[-mem:] if pairs else []
        gamma = 1.0 if not use else float(use[-1][0] @ use[-1][1] / (use[-1][1] @ use[-1][1]))
        H = gamma * np.eye(x.size)
        for s, y in use:
            rho = 1.0 / float(y @ s)
            V = np.eye(x.size) - rho * np.outer(y, s)
            H = V.T @ H @ V + rho * np.outer(s, s)
        d = -H @ g
        xn = x + d
        pairs.append((xn - x, A @ xn - g))
        xs.append(xn)
        x = xn
    return xs


@pytest.mark.parametrize("mem", (1, 2, 3))
def test_lbfgs_iterates_equal_dense_bfgs_on_a_quadratic(mem):
    """Cpp/src/LBFGS.cpp:63-103: H0 from the newest pair (:69-81), the
    history count min(k-1, m) (:83-86) and the ring index (:97).  On a
    quadratic with eigenvalues in (0.5, 1.5) every unit step satisfies weak
    Wolfe at (0, 0.5) (asserted below), so the line search accepts it with one
    evaluation and the iterates are those of the dense BFGS recursion with
    the same memory.  Five iterations at memory 1-3 wrap the ring."""
    n, iters = 4, 5
    A = _spd(5, n, 0.5, 1.5)
    x0 = G.normal(6, n)
    quad = lambda x: (0.5 * float(x @ A @ x), A @ x)
    xs = _dense_bfgs_iterates(A, x0, mem, iters)
    for k in range(iters):
        x, xn = xs[k], xs[k + 1]
        d = xn - x
        f, g = quad(x)
        fn, gn = quad(xn)
        assert fn < f and gn @ d > 0.5 * (g @ d), k   # weak Wolfe at (0, 0.5) for the unit step
    for k in range(1, iters + 1):
        x, f, tr = oracle.lbfgs(quad, x0, itr_max=k, tol=0.0, wolfe1=0.0, wolfe2=0.5, mem=mem)
        assert tr["status"] == 0 and tr["iters"] == k
        assert np.all(tr["evals"][1:] == 1)
        assert rel_err(x, xs[k]) <= 1e-12, (mem, k)
    # a ring of the wrong size is not the same method: the memory-1 and
    # memory-2 sequences part from the third iterate on
    assert rel_err(_dense_bfgs_iterates(A, x0, 1, iters)[3], _dense_bfgs_iterates(A, x0, 2, iters)[3]) > 1e-6


def test_line_search_limits_truncate_toward_zero():
    """Cpp/src/lineSearch.cpp:52-53: bisMax = max(50, log2(1e5 |d|)) and
    epdMax = max(10, log2(1e5 / |d|)) are doubles stored in an int, so the
    logarithm is truncated.  The loop (:57) runs while bisCnt <= bisMax and
    epdCnt <= epdMax, one evaluation per pass.
    Expansion: on f = c.x along d = -c every step satisfies Armijo with
    c1 = 0 and fails the curvature test, so the step doubles until
    epdCnt > epdMax: |d| = 1 gives log2(1e5) = 16.61, epdMax = 16, 17
    evaluations, code 2.
    Bisection: on f = x^2/2 from x = R along d = -R with (c1, c2) = (0.9,
    0.1) no step meets both conditions (Armijo needs t < 0.2, the curvature
    test t > 0.9), so every pass bisects: R = 3e10 gives log2(3e15) = 51.41,
    bisMax = 51, 52 evaluations, code 3; R = 1 gives bisMax = 50, 51."""
    c = np.array([0.6, 0.8])
    lin = lambda x: (float(c @ x), c.copy())
    code, _, _, _, _, cnt = oracle.line_search(lin, np.zeros(2), c, -c, 0.0, 0.0, 0.5)
    assert (code, cnt) == (2, 17)
    c2 = 1e-3 * c   # |d| = 1e-3: log2(1e8) = 26.58 -> epdMax 26
    lin2 = lambda x: (float(c2 @ x), c2.copy())
    code, _, _, _, _, cnt = oracle.line_search(lin2, np.zeros(2), c2, -c2, 0.0, 0.0, 0.5)
    assert (code, cnt) == (2, 27)
    quad = lambda x: (0.5 * float(x @ x), x.copy())
    for R, expect in ((3e10, 52), (1.0, 51)):
        x = np.array([R])
        code, _, _, _, _, cnt = oracle.line_search(quad, x, x, -x, 0.5 * R * R, 0.9, 0.1)
        assert (code, cnt) == (3, expect), (R, code, cnt)


@pytest.mark.parametrize("dim", (2, 3))
@pytest.mark.parametrize("ct", "GC")
@pytest.mark.parametrize("dt", "BLOU")
def test_varifold_terms_combine_to_the_reference_value(golden_varifold, dim, ct, dt):
    """DD - 2 DT + TT of orc_varifold_terms (the three pair sums the
    objective caches separately) equals the compiled reference's varifold
    value (Cpp/src/varifold3D.cpp:259-357, varifold2D.cpp:126-216)."""
    g = golden_varifold
    d = "d%d" % dim
    t = "%s_%s%s" % (d, ct, dt)
    cw, dw = g[t + "_w"]
    dd, dtt, tt = oracle.varifold_terms(g[d + "_q"], g[d + "_elm"], g[d + "_tc"], g[d + "_tu"], g[d + "_ta"],
                                        ct, cw, dt, dw)
    ref = float(g[t + "_val"])
    assert abs(dd - 2 * dtt + tt - ref) <= 1e-13 * max(abs(dd), abs(tt), abs(ref))
    assert dd > 0 and tt > 0 or dt in "LO"
