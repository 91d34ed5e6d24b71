"""numpy restatements used to check the GPU warm-start code: the
multilayer perceptron (forward, hand-written backprop, Adam) in FP64 with
optional emulation of the GPU's BF16 operand rounding, and the point
features.  Test infrastructure only."""
import numpy as np


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(flat, shapes):
    out, o = [], 0
    for din, dout in shapes:
        W = np.asarray(flat[o:o + din * dout], dtype=np.float64).reshape(din, dout)
        o += din * dout
        b = np.asarray(flat[o:o + dout], dtype=np.float64)
        o += dout
        out.append((W, b))
    return out


def pack(grads):
    return np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])


def forward(layers, X, emulate=True):
    """Returns (output, cache).  With emulate, GEMM operands (inputs,
    weights, hidden activations) are rounded to BF16 as on the GPU."""
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    h = r(X)
    ins, acts = [h], []
    for l, (W, b) in enumerate(layers):
        z = h @ r(W) + b
        if l < len(layers) - 1:
            a = np.maximum(z, 0.0)
            acts.append(a)
            h = r(a)
            ins.append(h)
        else:
            out = z
    return out, (ins, acts)


def loss_grad(layers, X, Y, emulate=True):
    """Mean squared error over the first 3 output columns and its gradient
    by backpropagation (list of (dW, db))."""
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    out, (ins, acts) = forward(layers, X, emulate)
    rows = len(X)
    d = out[:, :3] - Y
    loss = float(np.mean(d ** 2))
    dz = np.zeros_like(out)
    dz[:, :3] = 2.0 * d / (3.0 * rows)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        dzb = r(dz)
        grads[l] = (ins[l].T @ dzb, dz.sum(axis=0))
        if l == 0:
            break
        dh = dzb @ r(W).T
        dz = dh * (acts[l - 1] > 0)
    return loss, grads


def adam_step(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return p, m, v


WIDTHS = (0.1, 0.25, 0.5)


def point_features(q, tri, tc, tu, ta, knl_width):
    """Dense numpy version of lddmm_point_features (3D, small N)."""
    q = np.asarray(q, dtype=np.float64)
    n = len(q)
    nrm = np.zeros_like(q)
    c = np.cross(q[tri[:, 1]] - q[tri[:, 0]], q[tri[:, 2]] - q[tri[:, 0]])
    for k in range(3):
        np.add.at(nrm, tri[:, k], c)
    ln = np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm = np.where(ln > 0, nrm / np.where(ln > 0, ln, 1.0), 0.0)
    cols = [q, nrm]
    for s in WIDTHS:
        df = tc[None, :, :] - q[:, None, :]
        w = ta[None, :] * np.exp(-np.sum(df ** 2, axis=2) / s ** 2)
        un = nrm @ tu.T
        m0 = w.sum(axis=1, keepdims=True)
        cols += [m0, np.einsum("ij,ijc->ic", w, df) / (m0 + 1e-3 * s * s), (w * un) @ tu]
    for s in WIDTHS:
        r2 = np.sum((q[:, None, :] - q[None, :, :]) ** 2, axis=2)
        cols.append(np.exp(-r2 / s ** 2).sum(axis=1, keepdims=True) / n)
    cols += [np.full((n, 1), knl_width), np.full((n, 1), 1.0 / np.sqrt(n))]
    return np.concatenate(cols, axis=1)
