"""CPU checks of the numpy warm-start restatements the GPU tests compare
against: backprop against central finite differences, BF16 rounding
against known bit patterns, Adam against its closed first step."""
import numpy as np

import mlp_ref as R
from lddmm_amd import geometry as G
from lddmm_amd.warmstart import from_bf16_bits, to_bf16_bits

SHAPES = [(32, 64), (64, 64), (64, 32)]


def random_model(seed, shapes=SHAPES):
    rng = np.random.default_rng(seed)
    return [(rng.normal(0, np.sqrt(2.0 / a), (a, b)), rng.normal(0, 0.1, b)) for a, b in shapes]


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(0)
    layers = random_model(1)
    X = rng.normal(size=(40, 32))
    Y = rng.normal(size=(40, 3))
    loss, grads = R.loss_grad(layers, X, Y, emulate=False)
    flat = R.pack([(W, b) for W, b in layers])
    g = R.pack(grads)
    eps = 1e-6
    for k in rng.choice(flat.size, 60, replace=False):
        fp, fm = flat.copy(), flat.copy()
        fp[k] += eps
        fm[k] -= eps
        lp, _ = R.loss_grad(R.unpack(fp, SHAPES), X, Y, emulate=False)
        lm, _ = R.loss_grad(R.unpack(fm, SHAPES), X, Y, emulate=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_padded_output_columns_get_zero_gradient():
    rng = np.random.default_rng(2)
    layers = random_model(3)
    _, grads = R.loss_grad(layers, rng.normal(size=(10, 32)), rng.normal(size=(10, 3)), emulate=True)
    gW, gb = grads[-1]
    assert np.all(gW[:, 3:] == 0) and np.all(gb[3:] == 0)


def test_bf16_rounding_bit_patterns():
    x = np.array([1.0, -2.0, 0.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, 3.0e38, 1e-40], dtype=np.float32)
    bits = to_bf16_bits(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1.0) wins;
    # 1 + 3*2^-8 ties between 1 + 2^-7 and 1 + 2^-6: even (1 + 2^-6) wins
    assert list(bits[:6]) == [0x3F80, 0xC000, 0x0000, 0x3F80, 0x3F82, 0x3F80]
    assert np.array_equal(from_bf16_bits(bits), R.bf16(x).astype(np.float32))
    assert np.all(np.abs(R.bf16(x[:7]) - x[:7]) <= np.abs(x[:7]) * 2 ** -8)


def test_adam_first_step_is_signed_lr():
    p = np.array([1.0, -1.0, 0.5])
    g = np.array([0.3, -2.0, 1e-3])
    p1, _, _ = R.adam_step(p, g, np.zeros(3), np.zeros(3), 1, 0.01)
    assert np.allclose(p1, p - 0.01 * np.sign(g), atol=1e-7)


def test_point_features_invariants():
    q, tri = G.geodesic_sphere(3)
    tc, tu, ta = G.triangle_varifold(q, tri)
    f = R.point_features(q, tri, tc, tu, ta, 0.2)
    assert f.shape == (len(q), 32)
    # outward unit normals of a sphere are the positions
    assert np.allclose(f[:, 3:6], q / np.linalg.norm(q, axis=1, keepdims=True), atol=0.02)
    # the target is the source itself: the orientation moment (binormal
    # kernel (u.n) u summed) is along the normal, and the mass is positive
    for k in range(3):
        mu = f[:, 6 + 7 * k + 4:6 + 7 * k + 7]
        assert np.all(f[:, 6 + 7 * k] > 0)
        cos = np.sum(mu * f[:, 3:6], axis=1) / np.linalg.norm(mu, axis=1)
        assert np.all(cos > 0.99)
