"""Learned warm start for batched registrations (no counterpart in the
reference, which always starts L-BFGS from zero momentum: Cpp/src/LDDMM.m
passes iniAlpStk, normally zeros; the start enters at Cpp/src/LDDMM.cpp:87).

A per-vertex multilayer perceptron (32 point features -> 3, hidden layers
of 128, ReLU) is trained on the GPU with the library's BF16 MFMA kernels
(lddmm_mlp_*, FP32 master weights, Adam) to predict the time-mean momentum
density of solved registrations: (1/(T-1)) sum_t alpha_t[i] divided by the
vertex area A_i (a third of the adjacent triangles' areas), which is
comparable across vertex counts.  Its prediction times A_i is every time
slice of the initial momentum stack; `choose_start` scales it per problem by
the multiple with the lowest objective, zero included, so the plain start
remains the fallback.

Why the momentum itself: registrations started from the exact solutions'
time-mean momentum (constant in time) need 56 L-BFGS iterations to the
plain run's energy instead of 298 (median over 24 problems of 500-3,000
vertices, T = 21), whereas the momentum that a smooth field built from the
exact end-point displacement gives (displacement / (h (T-1)) divided by the
kernel mass sum_j k(q_i, q_j)) still needs 274.  The features are the
library's point features (lddmm_point_features) with the three source-
density columns replaced by the varifold gradient at the start, -dV/dq_i /
A_i (the end-point force of a registration that does not move).
"""
import ctypes as C

import numpy as np

from . import _check, _dp, _f64, _i32, _ip, _p, lib

__all__ = ["point_features", "Mlp", "gemm_bf16", "WarmStart", "to_bf16_bits", "from_bf16_bits"]

_fp = C.POINTER(C.c_float)
_u16p = C.POINTER(C.c_uint16)
NFEAT = 32


def _setup():
    L = lib()
    if not hasattr(L.lddmm_mlp_create, "_set"):
        L.lddmm_point_features.argtypes = [_fp, C.c_int, _dp, C.c_int, _ip, C.c_int, _dp, _dp, _dp, C.c_int,
                                           C.c_double]
        L.lddmm_mlp_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_uint64]
        L.lddmm_mlp_param_count.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.lddmm_mlp_get_params.argtypes = [C.c_void_p, _fp]
        L.lddmm_mlp_set_params.argtypes = [C.c_void_p, _fp]
        L.lddmm_mlp_forward.argtypes = [C.c_void_p, _fp, C.c_int, _fp]
        L.lddmm_mlp_train.argtypes = [C.c_void_p, _fp, _fp, C.c_int, C.c_int, C.c_int, C.c_float, _dp, _fp]
        L.lddmm_mlp_destroy.argtypes = [C.c_void_p]
        L.lddmm_mlp_destroy.restype = None
        L.lddmm_gemm_bf16.argtypes = [_fp, _u16p, _u16p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.lddmm_mlp_create._set = True
    return L


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _pf(a):
    return None if a is None else a.ctypes.data_as(_fp)


def to_bf16_bits(x):
    """float32 -> bfloat16 bit patterns, round to nearest even."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def from_bf16_bits(b):
    return (np.asarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def gemm_bf16(A, B, trans_a=False, trans_b=False, split_k=1):
    """C = op(A) op(B) on the library's MFMA kernel; A, B are bf16 bit
    arrays (uint16), C float32."""
    L = _setup()
    A = np.ascontiguousarray(A, dtype=np.uint16)
    B = np.ascontiguousarray(B, dtype=np.uint16)
    M, K = (A.shape[1], A.shape[0]) if trans_a else A.shape
    N = B.shape[0] if trans_b else B.shape[1]
    Cm = np.empty((M, N), dtype=np.float32)
    _check(L.lddmm_gemm_bf16(_pf(Cm), A.ctypes.data_as(_u16p), B.ctypes.data_as(_u16p), M, N, K, int(trans_a),
                             int(trans_b), int(split_k)))
    return Cm


def point_features(q0, elm, tgt_cen, tgt_dir, tgt_vol, knl_width):
    """(N, 32) float32 features of every source vertex (lddmm_point_features)."""
    L = _setup()
    q = _f64(q0)
    e = _i32(elm)
    tc, tu, ta = _f64(tgt_cen), _f64(tgt_dir), _f64(tgt_vol)
    out = np.empty((len(q), NFEAT), dtype=np.float32)
    _check(L.lddmm_point_features(_pf(out), q.shape[1], _p(q), len(q), _p(e), len(e), _p(tc), _p(tu), _p(ta),
                                  len(ta), float(knl_width)))
    return out


class Mlp:
    """The GPU multilayer perceptron (lddmm_mlp_*)."""

    def __init__(self, hidden=128, layers=4, seed=1):
        L = _setup()
        self.hidden, self.layers = hidden, layers
        h = C.c_void_p()
        _check(L.lddmm_mlp_create(C.byref(h), hidden, layers, seed))
        self._h = h
        n = C.c_int()
        _check(L.lddmm_mlp_param_count(self._h, C.byref(n)))
        self.nparam = n.value

    def shapes(self):
        """[(din, dout)] per layer, the last one padded to 32 outputs."""
        dims = [NFEAT] + [self.hidden] * (self.layers - 1) + [32]
        return list(zip(dims[:-1], dims[1:]))

    def unpack(self, flat):
        out, o = [], 0
        for din, dout in self.shapes():
            W = flat[o:o + din * dout].reshape(din, dout)
            o += din * dout
            b = flat[o:o + dout]
            o += dout
            out.append((W, b))
        return out

    def get_params(self):
        out = np.empty(self.nparam, dtype=np.float32)
        _check(lib().lddmm_mlp_get_params(self._h, _pf(out)))
        return out

    def set_params(self, flat):
        flat = _f32(flat)
        assert flat.size == self.nparam
        _check(lib().lddmm_mlp_set_params(self._h, _pf(flat)))

    def forward(self, feat):
        feat = _f32(feat)
        out = np.empty((len(feat), 3), dtype=np.float32)
        _check(lib().lddmm_mlp_forward(self._h, _pf(feat), len(feat), _pf(out)))
        return out

    def train(self, feat, tgt, steps, batch_rows, lr):
        feat, tgt = _f32(feat), _f32(tgt)
        losses = np.empty(max(steps, 1))
        _check(lib().lddmm_mlp_train(self._h, _pf(feat), _pf(tgt), len(feat), steps, batch_rows, lr,
                                     _p(losses), None))
        return losses[:steps]

    def gradient(self, feat, tgt):
        """(loss, gradient) of the full batch, parameters unchanged."""
        feat, tgt = _f32(feat), _f32(tgt)
        loss = np.empty(1)
        g = np.empty(self.nparam, dtype=np.float32)
        _check(lib().lddmm_mlp_train(self._h, _pf(feat), _pf(tgt), len(feat), 1, len(feat), 0.0, _p(loss), _pf(g)))
        return loss[0], g

    def close(self):
        if getattr(self, "_h", None):
            lib().lddmm_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def vertex_area(q0, elm):
    """A_i: a third of the areas of the triangles at each vertex."""
    q, e = np.asarray(q0, dtype=np.float64), np.asarray(elm)
    a = 0.5 * np.linalg.norm(np.cross(q[e[:, 1]] - q[e[:, 0]], q[e[:, 2]] - q[e[:, 0]]), axis=1)
    out = np.zeros(len(q))
    for c in range(3):
        out += np.bincount(e[:, c], weights=a / 3.0, minlength=len(q))
    return out


# columns of the point features that the warm start replaces by the start's
# varifold gradient per vertex area
GRAD_COLS = slice(27, 30)


class WarmStart:
    """Feature standardisation, target scale and the GPU network.  The
    network maps the point features of a source vertex to its time-mean
    momentum density; `predict` turns that into a constant-in-time momentum
    stack.  cen_type ... dir_width are the varifold kernels of the problems
    (the gradient feature)."""

    def __init__(self, hidden=128, layers=4, seed=1, cen_type="G", cen_width=0.1, dir_type="B", dir_width=0.0):
        self.net = Mlp(hidden, layers, seed)
        self.fmean = np.zeros(NFEAT, dtype=np.float32)
        self.fstd = np.ones(NFEAT, dtype=np.float32)
        self.yscale = 1.0
        self.vf = (cen_type, float(cen_width), dir_type, float(dir_width))
        self.meta = {}

    def features(self, p):
        """(N, 32) float32 features and the vertex areas of one problem."""
        from . import getElmIfo, varifold
        f = point_features(p["q0"], p["elm"], p["tgt_cen"], p["tgt_dir"], p["tgt_vol"], p["knl_width"])
        A = vertex_area(p["q0"], p["elm"])
        ifo = getElmIfo(np.asarray(p["elm"]), len(p["q0"]), one_based=False)
        _, dq = varifold(p["q0"], p["elm"], ifo, p["tgt_cen"], p["tgt_dir"], p["tgt_vol"], *self.vf)
        f[:, GRAD_COLS] = -dq / A[:, None]
        return f, A

    def fit(self, problems, solutions, steps=6000, batch_rows=8192, lr=1e-3, seed=0):
        """problems: batch problem dicts; solutions: their momentum stacks
        (T-1, N, 3) from the plain method.  Returns the per-step training
        losses (mean squared error of the standardised targets)."""
        feats, ys = [], []
        for p, a in zip(problems, solutions):
            f, A = self.features(p)
            feats.append(f)
            ys.append(np.asarray(a).mean(axis=0) / A[:, None])
        X = np.concatenate(feats)
        Y = np.concatenate(ys)
        self.fmean = X.mean(axis=0).astype(np.float32)
        self.fstd = np.maximum(X.std(axis=0), 1e-6).astype(np.float32)
        self.yscale = float(np.sqrt(np.mean(Y ** 2))) or 1.0
        perm = np.random.default_rng(seed).permutation(len(X))
        Xn = ((X - self.fmean) / self.fstd)[perm]
        Yn = (Y / self.yscale)[perm]
        self.meta = dict(problems=len(problems), rows=len(X), steps=steps, batch_rows=batch_rows, lr=lr)
        return self.net.train(Xn, Yn, steps, min(batch_rows, len(Xn)), lr)

    def density(self, problem):
        f, A = self.features(problem)
        return self.net.forward((f - self.fmean) / self.fstd).astype(np.float64) * self.yscale, A

    def start_vector(self, problem):
        """(N, 3): the predicted momentum of every time slice (density x
        vertex area)."""
        m, A = self.density(problem)
        return m * A[:, None]

    def start_vectors(self, problems, threads=8):
        """start_vector of every problem: the features of several problems at
        once on host threads (each library call runs on a stream of its
        own), then one network pass over all their rows (the network is
        row-wise, so the result is that of start_vector)."""
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(max(1, int(threads))) as ex:
            fa = list(ex.map(self.features, problems))
        X = np.concatenate([f for f, _ in fa])
        m = self.net.forward((X - self.fmean) / self.fstd).astype(np.float64) * self.yscale
        out, o = [], 0
        for f, A in fa:
            out.append(m[o:o + len(f)] * A[:, None])
            o += len(f)
        return out

    def predict(self, problem, time_num):
        """(T-1, N, 3) initial momentum stack for one problem."""
        return np.repeat(self.start_vector(problem)[None], time_num - 1, axis=0)

    def save(self, path, **extra):
        meta = dict(self.meta, **extra)
        np.savez(path, params=self.net.get_params(), fmean=self.fmean, fstd=self.fstd, yscale=self.yscale,
                 hidden=self.net.hidden, layers=self.net.layers, vf_types=np.array([self.vf[0], self.vf[2]]),
                 vf_widths=np.array([self.vf[1], self.vf[3]]), meta_keys=np.array(list(meta.keys())),
                 meta_vals=np.array([float(v) for v in meta.values()]))

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        ws = cls(int(z["hidden"]), int(z["layers"]), cen_type=str(z["vf_types"][0]), cen_width=float(z["vf_widths"][0]),
                 dir_type=str(z["vf_types"][1]), dir_width=float(z["vf_widths"][1]))
        ws.net.set_params(z["params"])
        ws.fmean, ws.fstd, ws.yscale = z["fmean"], z["fstd"], float(z["yscale"])
        ws.meta = {str(k): float(v) for k, v in zip(z["meta_keys"], z["meta_vals"])}
        return ws

    def close(self):
        self.net.close()


def choose_start(batch, base, scales=(0.25, 0.5, 0.75, 1.0, 1.25, 1.5)):
    """Per problem, the multiple c of the constant-in-time start `base` (one
    (N_p, 3) array per problem; c in scales, or c = 0: zero momentum) with the
    lowest objective, from one energy-only batched evaluation per multiple on
    the device (Batch.objective_scan); a later multiple replaces an earlier
    one only when strictly lower.  Returns (chosen multiple per problem, its
    objective, the zero-start objective)."""
    cs = np.concatenate([[0.0], np.asarray(scales, dtype=np.float64)])
    f = batch.objective_scan(base, cs)
    best = np.zeros(f.shape[1], dtype=np.int64)
    cols = np.arange(f.shape[1])
    for k in range(1, len(cs)):
        best = np.where(f[k] < f[best, cols], k, best)
    return cs[best], f[best, cols], f[0]
