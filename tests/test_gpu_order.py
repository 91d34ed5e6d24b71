"""The spatial order and what rests on it.  Objectives and batches hold
every problem in their own spatial order; relabelling the vertices,
elements and target elements of a problem must therefore leave energies
unchanged and permute gradients, trajectories and optimised momenta in the
same way, through every entry point (host arrays, device arrays, the
registration driver, the batch).  The FP32 policy's culling of tile pairs
whose Gaussian values are all below 2^-48 must agree with computing every
pair to a few FP64 ulps (in the expanded form of the Gaussian terms, the
default, and in the direct form)."""
import numpy as np
import pytest

import oracle
from conftest import rel_err
from lddmm_amd import geometry as G

pytestmark = pytest.mark.gpu

K = dict(G.BATCH_KERNELS)
K.update(time_num=6, time_stp=0.2)


def relabel(p, seed):
    """New vertex k is old vertex vp[k]; element and target rows shuffled,
    each element's vertex order (its orientation) kept."""
    rng = np.random.default_rng(seed)
    n = len(p["q0"])
    vp = rng.permutation(n)
    inv = np.empty(n, dtype=np.int64)
    inv[vp] = np.arange(n)
    ep = rng.permutation(len(p["elm"]))
    tp = rng.permutation(len(p["tgt_vol"]))
    q = dict(p)
    q["q0"] = p["q0"][vp]
    q["elm"] = inv[p["elm"][ep]].astype(np.int32)
    q["tgt_cen"], q["tgt_dir"], q["tgt_vol"] = p["tgt_cen"][tp], p["tgt_dir"][tp], p["tgt_vol"][tp]
    return q, vp


def objective(gpu, p, precision="fp64"):
    ifo = oracle.get_elm_ifo(p["elm"], len(p["q0"]))
    return gpu.Objective(p["q0"], p["elm"], ifo, p["tgt_cen"], p["tgt_vol"], p["tgt_dir"], K["cen_type"],
                         K["cen_width"], K["dir_type"], K["dir_width"], K["dis_wgt"], K["knl_order"], p["knl_width"],
                         K["time_stp"], K["time_num"], precision=precision)


@pytest.fixture(scope="module")
def problem():
    return G.batch_problems(1, seed=17, nmin=1500, nmax=1600)[0]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_objective_is_invariant_under_relabelling(gpu, problem, precision):
    print("relabelled objective, %s" % precision, flush=True)
    q, vp = relabel(problem, 3)
    n, T = len(problem["q0"]), K["time_num"]
    alp = 0.05 * G.normal(5, (T - 1) * n * 3).reshape(T - 1, n, 3)
    o1, o2 = objective(gpu, problem, precision), objective(gpu, q, precision)
    f1, g1 = o1(alp)
    f2, g2 = o2(alp[:, vp, :])
    tol = (1e-13, 1e-12) if precision == "fp64" else (1e-6, 1e-4)
    assert abs(f2 - f1) <= tol[0] * abs(f1)
    assert rel_err(g2.reshape(T - 1, n, 3), g1.reshape(T - 1, n, 3)[:, vp, :]) <= tol[1]
    if precision == "fp64":
        tr1, tr2 = o1.trajectory(), o2.trajectory()
        assert np.max(np.abs(tr2 - tr1[:, vp, :])) <= 1e-13
        assert np.array_equal(tr1[0], problem["q0"]) and np.array_equal(tr2[0], q["q0"])
    o1.close()
    o2.close()


def test_device_entry_takes_the_callers_order(gpu, problem):
    """lddmm_objgrd_dev on device arrays (device layout, caller's order)
    gives the host entry's result bitwise."""
    n, T = len(problem["q0"]), K["time_num"]
    alp = 0.05 * G.normal(6, (T - 1) * n * 3).reshape(T - 1, n, 3)
    obj = objective(gpu, problem)
    f, g = obj(alp)
    d_alp = gpu.DeviceArray(np.ascontiguousarray(alp.transpose(0, 2, 1)))
    d_grd = gpu.DeviceArray(nbytes=alp.nbytes)
    d_f = gpu.DeviceArray(nbytes=8)
    obj.eval_device(d_f.ptr, d_grd.ptr, d_alp.ptr)
    obj.synchronize()
    gd = d_grd.get().reshape(T - 1, 3, n).transpose(0, 2, 1)
    assert d_f.get()[0] == f and np.array_equal(gd.reshape(-1), g)
    obj.close()


def test_registration_is_invariant_under_relabelling(gpu, problem):
    print("relabelled registration", flush=True)
    q, vp = relabel(problem, 4)
    out = []
    for p in (problem, q):
        ifo = gpu.getElmIfo(p["elm"], len(p["q0"]), one_based=False)
        a, info = gpu.run(p["q0"], None, p["elm"], ifo, p["tgt_cen"], p["tgt_vol"], p["tgt_dir"], K["cen_type"],
                          K["cen_width"], K["dir_type"], K["dir_width"], K["dis_wgt"], K["knl_order"],
                          p["knl_width"], K["time_stp"], K["time_num"], 6, 0.0, 0.0, 0.5, 5)
        out.append((a, info))
    (a1, i1), (a2, i2) = out
    assert i1["iterations"] == i2["iterations"] == 6
    assert abs(i2["fcn"] - i1["fcn"]) <= 1e-11 * abs(i1["fcn"])
    assert rel_err(a2, a1[:, vp, :]) <= 1e-9


def test_batch_is_invariant_under_relabelling(gpu):
    print("relabelled batch", flush=True)
    probs = G.batch_problems(3, seed=23, nmin=300, nmax=2000)
    rel = [relabel(p, 10 + i) for i, p in enumerate(probs)]
    T = K["time_num"]
    args = (K["cen_type"], K["cen_width"], K["dir_type"], K["dir_width"], K["dis_wgt"], K["knl_order"],
            K["time_stp"], T)
    B1 = gpu.Batch(probs, *args, chunk=256)
    B2 = gpu.Batch([r[0] for r in rel], *args, chunk=256)
    alps = [0.05 * G.normal(40 + i, B1.sizes[i]).reshape(T - 1, B1.n[i], 3) for i in range(B1.P)]
    f1, g1 = B1.objgrd(alps)
    f2, g2 = B2.objgrd([a[:, vp, :] for a, (_, vp) in zip(alps, rel)])
    for i, (_, vp) in enumerate(rel):
        assert abs(f2[i] - f1[i]) <= 1e-13 * abs(f1[i])
        assert rel_err(g2[i], g1[i][:, vp, :]) <= 1e-12
    a1, i1, _ = B1.run(4, 0.0, 0.0, 0.5, 5)
    a2, i2, _ = B2.run(4, 0.0, 0.0, 0.5, 5)
    for i, (_, vp) in enumerate(rel):
        assert abs(i2[i]["fcn"] - i1[i]["fcn"]) <= 1e-11 * abs(i1[i]["fcn"])
        assert rel_err(a2[i], a1[i][:, vp, :]) <= 1e-9
    B1.close()
    B2.close()


def two_clusters(seed, nmin=1200, nmax=1300):
    """Two copies of a batch problem 6 apart: most tile pairs across the
    gap are beyond the FP32 Gaussian's range and are culled."""
    p = G.batch_problems(1, seed=seed, nmin=nmin, nmax=nmax)[0]
    n = len(p["q0"])
    shift = np.array([6.0, 0.0, 0.0])
    return dict(q0=np.concatenate([p["q0"], p["q0"] + shift]),
                elm=np.concatenate([p["elm"], p["elm"] + n]).astype(np.int32),
                tgt_cen=np.concatenate([p["tgt_cen"], p["tgt_cen"] + shift]),
                tgt_dir=np.concatenate([p["tgt_dir"], p["tgt_dir"]]),
                tgt_vol=np.concatenate([p["tgt_vol"], p["tgt_vol"]]), knl_width=p["knl_width"])


def culled_close(f0, f1, g0, g1):
    """Culled against full sums: the terms left out are below 2^-48 of the
    terms' other factors (at most N 2^-48 of a sum, measured at FP64
    rounding), so the sums agree far inside these bounds."""
    return abs(f1 - f0) <= 1e-13 * abs(f0) and rel_err(g1, g0) <= 1e-12


def test_culling_is_negligible(gpu, problem):
    """FP32 objective with and without culling, on a single shape (the
    varifold's width-0.1 Gaussian culls its far tiles) and on two shapes far
    apart (the deformation kernel culls too)."""
    print("culling on/off, objective", flush=True)
    for p in (problem, two_clusters(8)):
        n, T = len(p["q0"]), K["time_num"]
        alp = 0.05 * G.normal(7, (T - 1) * n * 3).reshape(T - 1, n, 3)
        obj = objective(gpu, p, "fp32")
        f64, g64 = objective(gpu, p)(alp)
        for expanded in (True, False):
            try:
                gpu.set_pair_expansion(expanded)
                gpu.set_pair_culling(False)
                f0, g0 = obj(alp)
                gpu.set_pair_culling(True)
                f1, g1 = obj(alp)
            finally:
                gpu.set_pair_culling(True)
                gpu.set_pair_expansion(True)
            assert culled_close(f0, f1, g0, g1)
            assert abs(f1 - f64) <= 1e-6 * abs(f64) and rel_err(g1, g64) <= 1e-4
        obj.close()


@pytest.mark.parametrize("shape", [None, (1024, 2, 4)])
def test_culling_is_negligible_in_a_batch(gpu, shape):
    """Batched FP32 evaluation with and without culling, at the default launch
    shape and with the matrix-core adjoint (adjoint lanes 4).  With culling
    off the matrix-core adjoint also stages the far source groups, whose
    offsets widen a tile's bound and lower its column scale, so the FP16
    halves of the near columns keep fewer bits: its gradient then agrees to
    FP32 precision (as two launch shapes do), not to the culled terms' size."""
    print("culling on/off, batch, shape %s" % (shape,), flush=True)
    probs = G.batch_problems(3, seed=29, nmin=400, nmax=2500) + [two_clusters(9)]
    T = K["time_num"]
    B = gpu.Batch(probs, K["cen_type"], K["cen_width"], K["dir_type"], K["dir_width"], K["dis_wgt"],
                  K["knl_order"], K["time_stp"], T, precision="fp32", chunk=256)
    if shape is not None:
        B.set_tuning(*shape)
    alps = [0.05 * G.normal(60 + i, B.sizes[i]).reshape(T - 1, B.n[i], 3) for i in range(B.P)]
    f1, g1 = B.objgrd(alps)
    try:
        gpu.set_pair_culling(False)
        f0, g0 = B.objgrd(alps)
    finally:
        gpu.set_pair_culling(True)
    f2, g2 = B.objgrd(alps)
    assert np.array_equal(f1, f2)
    for i in range(B.P):
        assert np.array_equal(g1[i], g2[i])   # repeatable after the switch
        if shape is None:
            assert culled_close(f0[i], f1[i], g0[i], g1[i])
        else:
            assert abs(f1[i] - f0[i]) <= 1e-13 * abs(f0[i]) and rel_err(g1[i], g0[i]) <= 1e-5, \
                (i, rel_err(g1[i], g0[i]))
    B.close()


def test_expanded_form(gpu):
    """FP32 Gaussian terms in the expanded form (r2 = |x_i|^2 + |x_j|^2 -
    2 x_i.x_j, the factor exp2(-|x_i|^2) applied per target) against the
    direct form and FP64, on one shape and on two shapes 6 apart whose tiles
    are small enough for it (about 9,500 vertices), and on a small shape
    whose tiles are too wide (the direct form is kept there: both switches
    give bitwise the same result)."""
    print("expanded form, objective", flush=True)
    big = G.batch_problems(1, seed=19, nmin=9000, nmax=10000)[0]
    small = dict(G.batch_problems(1, seed=31, nmin=40, nmax=50)[0])
    small["q0"], small["tgt_cen"] = 3.0 * small["q0"], 3.0 * small["tgt_cen"]   # tile radius 3 > 8 widths
    for p, expanded in ((big, True), (two_clusters(8, 9000, 10000), True), (small, False)):
        n, T = len(p["q0"]), K["time_num"]
        alp = 0.05 * G.normal(8, (T - 1) * n * 3).reshape(T - 1, n, 3)
        obj = objective(gpu, p, "fp32")
        try:
            gpu.set_pair_expansion(False)
            f0, g0 = obj(alp)
        finally:
            gpu.set_pair_expansion(True)
        f1, g1 = obj(alp)
        f64, g64 = objective(gpu, p)(alp)
        for f, g in ((f0, g0), (f1, g1)):
            assert abs(f - f64) <= 1e-6 * abs(f64) and rel_err(g, g64) <= 1e-4
        assert abs(f1 - f0) <= 1e-7 * abs(f64) and rel_err(g1, g0) <= 2e-5
        assert (not np.array_equal(g0, g1)) if expanded else (f0 == f1 and np.array_equal(g0, g1))
        obj.close()


@pytest.mark.parametrize("shape", [None, (1024, 2, 4)])
def test_expanded_form_in_a_batch(gpu, shape):
    """Batched FP32 evaluation with the expanded form switched off and on, at
    the default launch shape and with the matrix-core adjoint (which takes
    the vector path when the expanded form is off)."""
    print("expanded form, batch, shape %s" % (shape,), flush=True)
    probs = G.batch_problems(3, seed=37, nmin=40, nmax=2500) + [two_clusters(10)] + \
        G.batch_problems(1, seed=41, nmin=9000, nmax=10000)
    T = K["time_num"]
    args = (K["cen_type"], K["cen_width"], K["dir_type"], K["dir_width"], K["dis_wgt"], K["knl_order"],
            K["time_stp"], T)
    B = gpu.Batch(probs, *args, precision="fp32", chunk=256)
    if shape is not None:
        B.set_tuning(*shape)
    B64 = gpu.Batch(probs, *args, chunk=256)
    alps = [0.05 * G.normal(70 + i, B.sizes[i]).reshape(T - 1, B.n[i], 3) for i in range(B.P)]
    f1, g1 = B.objgrd(alps)
    try:
        gpu.set_pair_expansion(False)
        f0, g0 = B.objgrd(alps)
    finally:
        gpu.set_pair_expansion(True)
    f2, g2 = B.objgrd(alps)
    f64, g64 = B64.objgrd(alps)
    assert np.array_equal(f1, f2)
    assert not np.array_equal(g0[-1], g1[-1])   # the largest problem takes the expanded form
    for i in range(B.P):
        assert np.array_equal(g1[i], g2[i])
        for f, g in ((f0[i], g0[i]), (f1[i], g1[i])):
            assert abs(f - f64[i]) <= 1e-6 * abs(f64[i]) and rel_err(g, g64[i]) <= 1e-4
    B.close()
    B64.close()


def test_expanded_form_2d(gpu):
    """The expanded form on curves: a 1,600-point circle -> square, whose
    128-point tiles are arcs small enough for it, against the direct form and
    FP64 at two momentum scales."""
    print("expanded form, 2D", flush=True)
    w = G.circle_square(1600)
    w["q0"] = w["q0"] * 0.95
    w["time_num"], w["time_stp"] = 6, 0.2
    n = len(w["q0"])
    ifo = oracle.get_elm_ifo(w["elm"], n)
    args = (w["q0"], w["elm"], ifo, w["tgt_cen"], w["tgt_vol"], w["tgt_dir"], w["cen_type"], w["cen_width"],
            w["dir_type"], w["dir_width"], w["dis_wgt"], w["knl_order"], w["knl_width"], w["time_stp"], w["time_num"])
    o32, o64 = gpu.Objective(*args, precision="fp32"), gpu.Objective(*args)
    for seed, scale in ((0, 0.0), (3, 0.02)):
        a = scale * G.normal(seed + 200, o32.varnum)
        f64, g64 = o64(a)
        try:
            gpu.set_pair_expansion(False)
            f0, g0 = o32(a)
        finally:
            gpu.set_pair_expansion(True)
        f1, g1 = o32(a)
        assert not np.array_equal(g0, g1)
        for f, g in ((f0, g0), (f1, g1)):
            assert abs(f - f64) <= 1e-6 * abs(f64) and rel_err(g, g64) <= 1e-4
    o32.close()
    o64.close()


@pytest.mark.parametrize("dir_type", ["L", "O", "U"])
def test_expanded_varifold_orientation_kernels(gpu, dir_type):
    """The expanded form of the varifold's 'G' position kernel with the
    non-Binet orientation kernels (its second pair-term branch), on a shape
    whose element tiles take it, against the direct form and FP64."""
    print("expanded varifold, G x %s" % dir_type, flush=True)
    p = G.batch_problems(1, seed=43, nmin=6000, nmax=7000)[0]
    k = dict(K, dir_type=dir_type, dir_width=0.0 if dir_type == "L" else 0.6, time_num=4, time_stp=1.0 / 3.0)
    n, T = len(p["q0"]), k["time_num"]
    ifo = oracle.get_elm_ifo(p["elm"], n)
    args = (p["q0"], p["elm"], ifo, p["tgt_cen"], p["tgt_vol"], p["tgt_dir"], k["cen_type"], k["cen_width"],
            k["dir_type"], k["dir_width"], k["dis_wgt"], k["knl_order"], p["knl_width"], k["time_stp"], T)
    o32, o64 = gpu.Objective(*args, precision="fp32"), gpu.Objective(*args)
    alp = 0.05 * G.normal(11, (T - 1) * n * 3).reshape(T - 1, n, 3)
    f64, g64 = o64(alp)
    try:
        gpu.set_pair_expansion(False)
        f0, g0 = o32(alp)
    finally:
        gpu.set_pair_expansion(True)
    f1, g1 = o32(alp)
    assert not np.array_equal(g0, g1)
    for f, g in ((f0, g0), (f1, g1)):
        assert abs(f - f64) <= 1e-6 * abs(f64) and rel_err(g, g64) <= 1e-4
    o32.close()
    o64.close()
