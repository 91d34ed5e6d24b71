"""Edge cases of the GPU path: batches of curves and of Matern problems,
tiny and target-free problems in a batch, the shortest time grid,
concurrent evaluations from several host threads, and argument errors
(reported as negative statuses with a message, nothing launched)."""
import threading

import numpy as np
import pytest

import oracle
from conftest import rel_err
from lddmm_amd import geometry as G

pytestmark = pytest.mark.gpu


def curve_problems(count, seed):
    out = []
    u = G.uniform(seed, 2 * count)
    for p in range(count):
        n = int(40 + 400 * u[2 * p])
        q, seg = G.circle_curve(n)
        q = q * (0.8 + 0.3 * u[2 * p + 1]) + 0.01 * G.normal(seed + p, q.size).reshape(q.shape)
        tc, tu, ta = G.polyline_varifold(G.square_polyline(max(8, n // 4 * 4), 0.9))
        out.append(dict(q0=q, elm=seg, tgt_cen=tc, tgt_dir=tu, tgt_vol=ta, knl_width=0.2 + 0.1 * u[2 * p]))
    return out


def tetrahedron():
    q = np.array([[1.0, 1.0, 1.0], [1.0, -1.0, -1.0], [-1.0, 1.0, -1.0], [-1.0, -1.0, 1.0]]) * 0.5
    tri = np.array([[0, 1, 2], [0, 3, 1], [0, 2, 3], [1, 3, 2]], dtype=np.int32)
    return q, tri


def batch_vs_single(gpu, probs, k, precision="fp64", tol_f=1e-12, tol_g=1e-11):
    B = gpu.Batch(probs, k["cen_type"], k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"], k["knl_order"],
                  k["time_stp"], k["time_num"], precision=precision, chunk=256)
    alps = [0.05 * G.normal(500 + i, B.sizes[i]).reshape(k["time_num"] - 1, B.n[i], B.dim) for i in range(B.P)]
    f, g = B.objgrd(alps)
    for i, p in enumerate(probs):
        ifo = oracle.get_elm_ifo(p["elm"], len(p["q0"]))
        obj = gpu.Objective(p["q0"], p["elm"], ifo, p["tgt_cen"], p["tgt_vol"], p["tgt_dir"], k["cen_type"],
                            k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"], k["knl_order"],
                            p["knl_width"], k["time_stp"], k["time_num"], precision=precision)
        fs, gs = obj(alps[i])
        assert abs(f[i] - fs) <= tol_f * abs(fs), (i, f[i], fs)
        assert rel_err(g[i].reshape(-1), gs) <= tol_g, (i, rel_err(g[i].reshape(-1), gs))
        obj.close()
    B.close()


@pytest.mark.parametrize("order,precision", [(-1, "fp64"), (2, "fp64"), (3, "fp32")])
def test_batch_of_curves(gpu, order, precision):
    k = dict(cen_type="G", cen_width=0.1, dir_type="B", dir_width=0.0, dis_wgt=100.0, knl_order=order,
             time_stp=0.2, time_num=6)
    print("batch of 2D curves, order %d, %s" % (order, precision), flush=True)
    # FP32 (2D Matern: FP32 Bessel factors): the batch's source chunks differ from
    # the single-problem split, so the FP32 error budget applies
    tol = (1e-12, 1e-11) if precision == "fp64" else (1e-6, 1e-4)
    batch_vs_single(gpu, curve_problems(5, 31), k, precision=precision, tol_f=tol[0], tol_g=tol[1])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_of_matern_surfaces(gpu, precision):
    k = dict(G.BATCH_KERNELS)
    k.update(knl_order=1, time_num=4, time_stp=1.0 / 3.0, cen_type="C", dir_type="U", dir_width=0.7)
    print("batch of Matern surfaces, %s" % precision, flush=True)
    probs = G.batch_problems(4, seed=12, nmin=50, nmax=900)
    # FP32: the batch's source chunks differ from the single-problem split, so
    # the FP32 tile sums round differently; the FP32 error budget applies
    tol = (1e-12, 1e-11) if precision == "fp64" else (1e-6, 1e-4)
    batch_vs_single(gpu, probs, k, precision=precision, tol_f=tol[0], tol_g=tol[1])


def test_batch_with_tiny_and_target_free_problems(gpu):
    k = dict(G.BATCH_KERNELS)
    k.update(time_num=2, time_stp=1.0)   # the shortest time grid: one Euler step
    q, tri = tetrahedron()
    tc, tu, ta = G.triangle_varifold(1.2 * q, tri)
    empty3 = np.zeros((0, 3))
    probs = [dict(q0=q, elm=tri, tgt_cen=tc, tgt_dir=tu, tgt_vol=ta, knl_width=0.3),
             dict(q0=q.copy(), elm=tri, tgt_cen=empty3, tgt_dir=empty3, tgt_vol=np.zeros(0), knl_width=0.25)]
    probs += G.batch_problems(2, seed=5, nmin=100, nmax=300)
    print("batch with a tetrahedron and a target-free problem, T=2", flush=True)
    batch_vs_single(gpu, probs, k)


def test_concurrent_objectives_from_host_threads(gpu):
    """Handles are independent: evaluations from four host threads at once
    give the sequential results bitwise."""
    probs = G.batch_problems(4, seed=21, nmin=300, nmax=1500)
    k = dict(G.BATCH_KERNELS)
    k.update(time_num=5, time_stp=0.25)
    objs, alps = [], []
    for i, p in enumerate(probs):
        ifo = oracle.get_elm_ifo(p["elm"], len(p["q0"]))
        objs.append(gpu.Objective(p["q0"], p["elm"], ifo, p["tgt_cen"], p["tgt_vol"], p["tgt_dir"], k["cen_type"],
                                  k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"], k["knl_order"],
                                  p["knl_width"], k["time_stp"], k["time_num"]))
        alps.append(0.05 * G.normal(700 + i, objs[-1].varnum))
    seq = [o(a) for o, a in zip(objs, alps)]
    out = [None] * len(objs)

    def work(i):
        for _ in range(3):
            out[i] = objs[i](alps[i])

    th = [threading.Thread(target=work, args=(i,)) for i in range(len(objs))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for (f0, g0), (f1, g1) in zip(seq, out):
        assert f0 == f1 and np.array_equal(g0, g1)
    for o in objs:
        o.close()


def test_argument_errors_are_reported(gpu):
    q, tri = tetrahedron()
    tc, tu, ta = G.triangle_varifold(q, tri)
    ifo = oracle.get_elm_ifo(tri, 4)
    bad = tri.copy()
    bad[0, 0] = 4   # vertex index out of range
    cases = [
        lambda: gpu.Objective(q, bad, ifo, tc, ta, tu, "G", 0.1, "B", 0.0, 1.0, -1, 0.3, 0.5, 3),
        lambda: gpu.Objective(q, tri, ifo, tc, ta, tu, "G", 0.1, "B", 0.0, 1.0, 7, 0.3, 0.5, 3),
        lambda: gpu.Objective(q, tri, ifo, tc, ta, tu, "X", 0.1, "B", 0.0, 1.0, -1, 0.3, 0.5, 3),
        lambda: gpu.Objective(q, tri, ifo, tc, ta, tu, "G", 0.1, "B", 0.0, 1.0, -1, 0.3, 0.5, 1),
        lambda: gpu.Objective(q, tri, ifo, tc, ta, tu, "G", 0.1, "O", 0.0, 1.0, -1, 0.3, 0.5, 3),
        lambda: gpu.Batch([dict(q0=q, elm=bad, tgt_cen=tc, tgt_dir=tu, tgt_vol=ta, knl_width=0.3)], "G", 0.1, "B",
                          0.0, 1.0, -1, 0.5, 3),
        # several problems: the set-up runs on host threads, the error comes from one of them
        lambda: gpu.Batch([dict(q0=q, elm=bad if i == 3 else tri, tgt_cen=tc, tgt_dir=tu, tgt_vol=ta, knl_width=0.3)
                           for i in range(6)], "G", 0.1, "B", 0.0, 1.0, -1, 0.5, 3),
        lambda: gpu.l2Dist(q, q[:3], 1.0),
    ]
    for c in cases:
        with pytest.raises((gpu.LddmmError, ValueError)) as ei:
            c()
        if isinstance(ei.value, gpu.LddmmError):
            assert ei.value.code == -1 and str(ei.value)
