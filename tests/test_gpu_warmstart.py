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


def test_warm_start_reduces_iterations_on_held_out_problems(gpu, ws):
    """Train on registrations the method solved, then register held-out
    problems from the warm start and from zero momentum: starting from the
    warm start lowers the objective at a small iteration budget for most
    problems and reaches the zero start's final energy (within 1e-6) in
    fewer L-BFGS iterations at the median, fallbacks counted."""
    k = dict(G.BATCH_KERNELS)
    T = 6
    k["time_num"], k["time_stp"] = T, 1.0 / (T - 1)

    def batch(probs):
        return gpu.Batch(probs, k["cen_type"], k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"],
                         k["knl_order"], k["time_stp"], k["time_num"], chunk=256)

    train = G.batch_problems(48, seed=101, nmin=100, nmax=800)
    Bt = batch(train)
    sol, _, _ = Bt.run(300, 0.0, 0.0, 0.5, 5, precision="fp32", rel_energy_tol=1e-6)
    Bt.close()
    w = ws.WarmStart(hidden=128, layers=4, seed=3)
    losses = w.fit(train, sol, k["knl_order"], k["time_stp"], T, steps=3000, batch_rows=4096, lr=1e-3)
    print("warm-start training loss %.3e -> %.3e" % (losses[0], losses[-20:].mean()), flush=True)
    assert losses[-20:].mean() < 0.1 * losses[0]
    test = G.batch_problems(16, seed=202, nmin=100, nmax=800)
    B = batch(test)
    _, i0, _ = B.run(300, 0.0, 0.0, 0.5, 5, rel_energy_tol=1e-6)
    targets = np.array([r["fcn"] for r in i0]) * (1 + 1e-6)
    warm = [w.predict(p, k["knl_order"], k["time_stp"], T) for p in test]
    starts, scale = ws.choose_start(B, warm)
    f0, _ = B.objgrd([np.zeros_like(a) for a in warm])
    fs, _ = B.objgrd(starts)
    assert np.all(fs <= f0)
    _, z10, _ = B.run(10, 0.0, 0.0, 0.5, 5)
    _, w10, _ = B.run(10, 0.0, 0.0, 0.5, 5, alp0=starts)
    better10 = sum(b["fcn"] < a["fcn"] for a, b in zip(z10, w10))
    _, iw, _ = B.run(300, 0.0, 0.0, 0.5, 5, alp0=starts, rel_energy_tol=1e-6, targets=targets)
    its0 = np.array([r["iterations"] for r in i0])
    reached = np.array([r["reached_target"] for r in iw])
    # a warm-started run that misses the zero start's energy falls back to
    # the zero-start run: it costs both runs' iterations
    costw = np.array([r["iterations"] for r in iw]) + np.where(reached, 0, its0)
    print("scales %s; lower energy after 10 iterations in %d of %d; iterations to the zero start's energy: "
          "zero %s (median %d), warm %s (median %d), reached %d of %d" %
          (np.round(scale, 2).tolist(), better10, len(test), its0.tolist(), np.median(its0), costw.tolist(),
           np.median(costw), reached.sum(), len(test)), flush=True)
    assert np.sum(scale > 0) >= len(test) // 2
    assert better10 >= (2 * len(test)) // 3
    assert reached.sum() >= len(test) // 2
    assert np.median(costw) < 0.8 * np.median(its0)
    B.close()
