"""The learned warm start on the GPU: the BF16 MFMA GEMM against numpy,
the network's forward pass, gradient and Adam step against the FP64 numpy
restatement (with the GPU's BF16 operand rounding emulated), the point
features against their dense numpy version, and the warm start's effect on
held-out registrations."""
import numpy as np
import pytest

import mlp_ref as R
from conftest import rel_err
from lddmm_amd import geometry as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ws(gpu):
    from lddmm_amd import warmstart
    return warmstart


@pytest.mark.parametrize("ta,tb,split", [(0, 0, 1), (1, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 4), (0, 0, 3)])
def test_gemm_bf16_matches_numpy(ws, ta, tb, split):
    M, N, K = 160, 96, 272
    rng = np.random.default_rng(10 + 4 * ta + 2 * tb + split)
    A = ws.to_bf16_bits(rng.normal(size=(K, M) if ta else (M, K)))
    B = ws.to_bf16_bits(rng.normal(size=(N, K) if tb else (K, N)))
    print("gemm M=%d N=%d K=%d ta=%d tb=%d split=%d" % (M, N, K, ta, tb, split), flush=True)
    C = ws.gemm_bf16(A, B, bool(ta), bool(tb), split)
    a = ws.from_bf16_bits(A).astype(np.float64)
    b = ws.from_bf16_bits(B).astype(np.float64)
    ref = (a.T if ta else a) @ (b.T if tb else b)
    # products of BF16 values are exact in FP32; only the FP32 sums round
    assert np.max(np.abs(C - ref)) <= 1e-5 * np.max(np.abs(ref)) * np.sqrt(K)


def test_gemm_bf16_lane_layout_single_tile(ws):
    """One 32 x 32 x 16 tile with one-hot operands pins the MFMA lane
    mapping of A rows, B columns and accumulator elements."""
    A = np.zeros((32, 16), dtype=np.float32)
    B = np.zeros((16, 32), dtype=np.float32)
    for i in range(32):
        A[i, i % 16] = i + 1
    for j in range(32):
        B[(j * 5) % 16, j] = 1.0 / (j + 1) if j % 3 else 2.0
    C = ws.gemm_bf16(ws.to_bf16_bits(A), ws.to_bf16_bits(B))
    ref = ws.from_bf16_bits(ws.to_bf16_bits(A)).astype(np.float64) @ ws.from_bf16_bits(ws.to_bf16_bits(B))
    assert np.array_equal(C, ref.astype(np.float32))


def make_data(rows, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(rows, 32)).astype(np.float32)
    Wt = rng.normal(size=(32, 3)) / 6.0
    Y = (np.sin(X @ Wt) + 0.1 * X[:, :3]).astype(np.float32)
    return X, Y


@pytest.mark.parametrize("rows", [128, 300, 5000])
def test_mlp_forward_and_gradient_match_numpy(ws, rows):
    net = ws.Mlp(hidden=128, layers=4, seed=5)
    p = net.get_params()
    rng = np.random.default_rng(rows)
    p = (p + 0.05 * rng.normal(size=p.size)).astype(np.float32)
    net.set_params(p)
    X, Y = make_data(rows, rows + 1)
    print("mlp rows=%d params=%d" % (rows, net.nparam), flush=True)
    layers = R.unpack(p, net.shapes())
    out_ref, _ = R.forward(layers, X, emulate=True)
    out = net.forward(X)
    # FP32 sums differ from the FP64 emulation in the last bits, which now
    # and then moves a hidden activation across a BF16 rounding boundary
    # (one ulp, 2^-8 relative): rare large differences, tiny mean
    err = np.abs(out - out_ref[:, :3]) / np.max(np.abs(out_ref[:, :3]))
    assert err.max() <= 2e-3 and err.mean() <= 1e-5, (err.max(), err.mean())
    loss, g = net.gradient(X, Y)
    loss_ref, grads = R.loss_grad(layers, X, Y, emulate=True)
    assert abs(loss - loss_ref) <= 1e-5 * loss_ref
    gg = R.unpack(g, net.shapes())
    for l, ((gW, gb), (rW, rb)) in enumerate(zip(gg, grads)):
        eW, eb = rel_err(gW, rW), rel_err(gb, rb)
        print("layer %d: dW %.2e db %.2e" % (l, eW, eb), flush=True)
        assert eW <= 2e-3 and eb <= 2e-3
    assert np.all(gg[-1][0][:, 3:] == 0) and np.all(gg[-1][1][3:] == 0)
    # the gradient call left the parameters alone
    assert np.array_equal(net.get_params(), p)
    net.close()


def test_mlp_adam_step_and_training(ws):
    net = ws.Mlp(hidden=128, layers=4, seed=7)
    X, Y = make_data(2048, 3)
    p0 = net.get_params().astype(np.float64)
    _, g = net.gradient(X, Y)
    losses = net.train(X, Y, 1, 2048, 1e-3)
    p1 = net.get_params().astype(np.float64)
    ref, _, _ = R.adam_step(p0, g.astype(np.float64), np.zeros_like(p0), np.zeros_like(p0), 1, 1e-3)
    assert np.max(np.abs(p1 - ref)) <= 1e-6
    losses = net.train(X, Y, 400, 512, 2e-3)
    print("training loss %.4e -> %.4e" % (losses[0], losses[-5:].mean()), flush=True)
    assert losses[-5:].mean() < 0.2 * losses[0]
    net.close()


def test_point_features_match_numpy(ws):
    q, tri = G.geodesic_sphere(4)
    q = q + 0.01 * G.normal(3, q.size).reshape(q.shape)
    tq, ttri = G.superquadric(4, 11)
    tc, tu, ta = G.triangle_varifold(tq, ttri)
    f = ws.point_features(q, tri, tc, tu, ta, 0.25)
    ref = R.point_features(q, tri, tc, tu, ta, 0.25)
    for c in range(32):
        assert rel_err(f[:, c], ref[:, c]) <= 1e-5, c


def test_warm_start_fit_learns_momentum_density(gpu, ws):
    """WarmStart.fit on a few solved problems (T = 6): the training loss of
    the time-mean momentum density falls, and the prediction is a
    constant-in-time stack of the problem's shape."""
    k = dict(G.BATCH_KERNELS)
    T = 6
    train = G.batch_problems(8, seed=101, nmin=100, nmax=600)
    B = gpu.Batch(train, k["cen_type"], k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"],
                  k["knl_order"], 1.0 / (T - 1), T, chunk=256)
    sol, _, _ = B.run(60, 0.0, 0.0, 0.5, 5, rel_energy_tol=1e-6)
    B.close()
    w = ws.WarmStart(hidden=128, layers=4, seed=3)
    losses = w.fit(train, sol, steps=400, batch_rows=1024, lr=1e-3)
    print("warm-start fit: loss %.3e -> %.3e" % (losses[0], losses[-20:].mean()), flush=True)
    assert losses[-20:].mean() < 0.7 * losses[0]
    a = w.predict(train[0], T)
    assert a.shape == (T - 1, len(train[0]["q0"]), 3) and np.all(a[0] == a[-1]) and np.all(np.isfinite(a))
    w.close()


def test_trained_warm_start_gain_at_full_time_grid(gpu, ws):
    """The network shipped in lddmm_amd/data/warmstart_c4.npz (trained by
    lddmm_amd.train_warmstart on 1,792 problems of the batched configuration's
    generator, solved by the plain FP64 method) on 12 held-out problems of
    another seed at T = 21: the batched plain FP64 run sets each energy; the
    batched FP32 policy reaches it from zero momentum and from the warm start
    (at most 375 iterations, then a zero-start fallback).  The warm start
    must lower the starting objective for most problems and reach the
    targets in fewer iterations at the median, fallbacks counted."""
    import os
    k = G.BATCH_KERNELS
    T = k["time_num"]
    weights = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lddmm_amd", "data",
                           "warmstart_c4.npz")
    w = ws.WarmStart.load(weights)
    assert w.meta["problems"] >= 1792
    test = G.batch_problems(12, seed=5, nmin=500, nmax=3000)
    B = gpu.Batch(test, k["cen_type"], k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"], k["knl_order"],
                  k["time_stp"], T)
    _, i0, _ = B.run(300, 0.0, 0.0, 0.5, 5, rel_energy_tol=1e-6)
    targets = np.array([r["fcn"] for r in i0]) * (1 + 1e-6)
    _, iz, _ = B.run(900, 0.0, 0.0, 0.5, 5, precision="fp32", targets=targets)
    base = w.start_vectors(test, threads=4)
    for p, u in zip(test, base):   # threaded features and one pass over all rows: the per-problem result
        assert np.array_equal(u, w.start_vector(p))
    scale, fs, f0 = ws.choose_start(B, base)
    _, iw, _ = B.run_scaled(base, scale, 375, 0.0, 0.0, 0.5, 5, precision="fp32", targets=targets)
    itz = np.array([r["iterations"] for r in iz])
    reached = np.array([r["reached_target"] for r in iw])
    costw = np.array([r["iterations"] for r in iw]) + np.where(reached, 0, itz)
    print("warm start at T=21: start energy / zero start %s; iterations zero %s (median %d), warm %s "
          "(median %d), reached %d of %d" % (np.round(fs / f0, 3).tolist(), itz.tolist(), np.median(itz),
                                            costw.tolist(), np.median(costw), reached.sum(), len(test)), flush=True)
    assert all(r["reached_target"] for r in iz)
    assert np.sum(scale > 0) >= len(test) // 2 and np.all(fs <= f0)
    assert reached.sum() >= len(test) // 2
    assert np.median(costw) < np.median(itz)
    B.close()
    w.close()
