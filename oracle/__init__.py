"""CPU checker for the LDDMM hot path -- TEST INFRASTRUCTURE ONLY.

`liboracle.so` (built from lddmm_oracle.c by this folder's Makefile) is an
FP64, single-threaded restatement of the reference's CPU path
(dnhsieh/LDDMM Cpp/src).  `ref()` loads the reference's own kernel sources
compiled into `_ref/libref{2,3}d.so` when they were built.

Only tests/, the smoke check of the repository's entry module and the
`cpu_baseline` leg of bench.py may import this module; the product package
never does.

All arrays use the reference's host layout: point sets D x N column-major
(numpy shape (N, D), C order), stacks (T-1, N, D), element tables (E, V)
int32 0-based, the adjacency table (N, rows) int32.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_refs = {}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Problem(C.Structure):
    _fields_ = [("dim", C.c_int), ("ndeNum", C.c_int), ("timeNum", C.c_int), ("knlOrder", C.c_int),
                ("knlWidth", C.c_double), ("timeStp", C.c_double), ("disWgt", C.c_double),
                ("iniNdeMat", C.c_void_p), ("elmVtx", C.c_void_p), ("elmIfo", C.c_void_p),
                ("elmNum", C.c_int), ("ifoRows", C.c_int), ("tgtNum", C.c_int),
                ("tgtCen", C.c_void_p), ("tgtDir", C.c_void_p), ("tgtVol", C.c_void_p),
                ("cenType", C.c_char), ("dirType", C.c_char),
                ("cenWidth", C.c_double), ("dirWidth", C.c_double), ("tgtNde", C.c_void_p)]


class LbfgsOpts(C.Structure):
    _fields_ = [("itrMax", C.c_int), ("memNum", C.c_int), ("vbsFlg", C.c_int),
                ("tolVal", C.c_double), ("wolfe1", C.c_double), ("wolfe2", C.c_double)]


class Trace(C.Structure):
    _fields_ = [("f", C.c_void_p), ("gnorm", C.c_void_p), ("step", C.c_void_p), ("evals", C.c_void_p),
                ("iters", C.c_int), ("status", C.c_int), ("totalEvals", C.c_int)]


OBJFUN = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double))


def build():
    """Compile liboracle.so (and the reference kernels when the tree exists)."""
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])
    if os.path.isdir("/root/reference/Cpp/src"):
        subprocess.check_call(["make", "-s", "-C", _HERE, "ref"])


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_bessel_p01.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_compute_kernel.argtypes = [_dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int]
        L.orc_compute_kernel_cross.argtypes = [_dp, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int]
        L.orc_multiply_kernel.argtypes = [_dp, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int]
        L.orc_multiply_kernel_cross.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int]
        L.orc_dq_kernel.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int]
        L.orc_dq_kernel_cross.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int]
        L.orc_element_geometry.argtypes = [_dp, _dp, _dp, _dp, _ip, C.c_int, C.c_int]
        L.orc_varifold_terms.argtypes = [C.POINTER(C.c_double)] * 3 + [_dp, _ip, _dp, _dp, _dp, C.c_char, C.c_double,
                                                                        C.c_char, C.c_double, C.c_int, C.c_int, C.c_int]
        L.orc_varifold.argtypes = [C.POINTER(C.c_double), C.c_void_p, _dp, _ip, _ip, _dp, _dp, _dp, C.c_char,
                                   C.c_double, C.c_char, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_get_elm_ifo.argtypes = [C.c_void_p, _ip, C.c_int, C.c_int, C.c_int]
        L.orc_objgrd.argtypes = [C.POINTER(C.c_double), _dp, _dp, C.POINTER(Problem), C.c_void_p]
        L.orc_l2_dist.argtypes = [C.POINTER(C.c_double), _dp, _dp, _dp, C.c_double, C.c_int, C.c_int]
        L.orc_l2_dist.restype = None
        L.orc_get_direction.argtypes = [_dp, C.c_double, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, _dp, C.c_size_t]
        L.orc_get_direction.restype = None
        L.orc_line_search.argtypes = [OBJFUN, C.c_void_p, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double,
                                      C.c_double, _dp, _dp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                      C.POINTER(C.c_int), C.c_size_t, C.c_int]
        L.orc_lbfgs.argtypes = [OBJFUN, C.c_void_p, C.POINTER(C.c_double), _dp, _dp, C.c_size_t,
                                C.POINTER(LbfgsOpts), C.c_void_p]
        L.orc_lddmm_run.argtypes = [C.POINTER(Problem), _dp, C.POINTER(LbfgsOpts), _dp, C.POINTER(C.c_double),
                                    C.c_void_p]
        _lib = L
    return _lib


def ref(dim):
    """The reference's own kernels (compiled from /root/reference in this
    container into _ref/); None when they were not built."""
    if dim in _refs:
        return _refs[dim]
    path = os.path.join(_HERE, "_ref", "libref%dd.so" % dim)
    if not os.path.exists(path):
        _refs[dim] = None
        return None
    L = C.CDLL(path)
    L.ref_p01.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.ref_compute_kernel.argtypes = [_dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int]
    L.ref_compute_kernel_cross.argtypes = [_dp, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int]
    L.ref_multiply_kernel.argtypes = [_dp, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int]
    L.ref_multiply_kernel_cross.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int]
    L.ref_dq_kernel.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int]
    L.ref_dq_kernel_cross.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int]
    L.ref_varifold.argtypes = [C.POINTER(C.c_double), C.c_void_p, _dp, _ip, _ip, _dp, _dp, _dp, C.c_char,
                               C.c_double, C.c_char, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    _refs[dim] = L
    return L


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _chk(rc):
    if rc < 0:
        raise ValueError("oracle rejected the arguments (code %d)" % rc)
    return rc


# ---------------------------------------------------------------- operators
def bessel_p01(x, L=None):
    p0, p1 = C.c_double(), C.c_double()
    (L or lib()).orc_bessel_p01(float(x), C.byref(p0), C.byref(p1))
    return p0.value, p1.value


def compute_kernel(q, order, width):
    q = _f64(q)
    n, d = q.shape
    K = np.empty((n, n))
    _chk(lib().orc_compute_kernel(K, q, order, width, n, d))
    return K


def compute_kernel_cross(qi, qj, order, width):
    """Returns K with K[j, i] = k(qi_i, qj_j) (the reference's Ni x Nj
    column-major matrix, read as a (Nj, Ni) C-order array)."""
    qi, qj = _f64(qi), _f64(qj)
    K = np.empty((qj.shape[0], qi.shape[0]))
    _chk(lib().orc_compute_kernel_cross(K, qi, qj, order, width, qi.shape[0], qj.shape[0], qi.shape[1]))
    return K


def multiply_kernel(q, a, order, width):
    q, a = _f64(q), _f64(a)
    v = np.empty_like(q)
    _chk(lib().orc_multiply_kernel(v, q, a, order, width, q.shape[0], q.shape[1]))
    return v


def multiply_kernel_cross(qi, qj, a, order, width):
    qi, qj, a = _f64(qi), _f64(qj), _f64(a)
    v = np.empty_like(qi)
    _chk(lib().orc_multiply_kernel_cross(v, qi, qj, a, order, width, qi.shape[0], qj.shape[0], qi.shape[1]))
    return v


def dq_kernel(q, l, r, order, width):
    q, l, r = _f64(q), _f64(l), _f64(r)
    g = np.empty_like(q)
    _chk(lib().orc_dq_kernel(g, q, l, r, order, width, q.shape[0], q.shape[1]))
    return g


def dq_kernel_cross(qi, qj, l, r, order, width):
    qi, qj, l, r = _f64(qi), _f64(qj), _f64(l), _f64(r)
    gi, gj = np.empty_like(qi), np.empty_like(qj)
    _chk(lib().orc_dq_kernel_cross(gi, gj, qi, qj, l, r, order, width, qi.shape[0], qj.shape[0], qi.shape[1]))
    return gi, gj


def element_geometry(q, elm):
    q, elm = _f64(q), _i32(elm)
    ne, d = elm.shape[0], q.shape[1]
    c, u, a = np.empty((ne, d)), np.empty((ne, d)), np.empty(ne)
    lib().orc_element_geometry(c, u, a, q, elm, ne, d)
    return c, u, a


def get_elm_ifo(elm, nq):
    """Adjacency table of a 0-based (E, V) element table, shape (nq, rows)."""
    elm = _i32(elm)
    ne, vpe = elm.shape
    rows = _chk(lib().orc_get_elm_ifo(None, elm, vpe, ne, nq))
    ifo = np.zeros((nq, rows), dtype=np.int32)
    lib().orc_get_elm_ifo(ifo.ctypes.data, elm, vpe, ne, nq)
    return ifo


def varifold_terms(q, elm, tc, tu, ta, cen_type, cen_w, dir_type, dir_w):
    q, elm = _f64(q), _i32(elm)
    dd, dt, tt = C.c_double(), C.c_double(), C.c_double()
    _chk(lib().orc_varifold_terms(C.byref(dd), C.byref(dt), C.byref(tt), q, elm, _f64(tc), _f64(tu), _f64(ta),
                                  cen_type.encode(), cen_w, dir_type.encode(), dir_w, elm.shape[0],
                                  len(ta), q.shape[1]))
    return dd.value, dt.value, tt.value


def l2_dist(q, tgt, w):
    """l2Dist (MATLAB/src/l2Dist.m:1-5): (w/2 |q - tgt|^2, w (q - tgt))."""
    q, tgt = _f64(q), _f64(tgt)
    f = C.c_double()
    g = np.empty_like(q)
    lib().orc_l2_dist(C.byref(f), g, q, tgt, C.c_double(w), q.shape[0], q.shape[1])
    return f.value, g


def varifold(q, elm, ifo, tc, tu, ta, cen_type, cen_w, dir_type, dir_w, grad=True, L=None, prefix="orc"):
    q, elm, ifo = _f64(q), _i32(elm), _i32(ifo)
    val = C.c_double()
    dq = np.zeros_like(q) if grad else None
    fn = getattr(L or lib(), prefix + "_varifold")
    _chk(fn(C.byref(val), dq.ctypes.data if grad else None, q, elm, ifo, _f64(tc), _f64(tu), _f64(ta),
            cen_type.encode(), cen_w, dir_type.encode(), dir_w, q.shape[0], elm.shape[0], ifo.shape[1],
            len(ta), q.shape[1]))
    return (val.value, dq) if grad else val.value


# ---------------------------------------------------------------- objective
class OracleProblem:
    """Keeps the numpy arrays of an orc_problem alive.  With tgt_nde (N, D)
    the matching term is l2Dist against those landmarks (the element and
    target varifold arguments are then ignored and may be empty)."""

    def __init__(self, q0, elm, tgt_cen, tgt_dir, tgt_vol, cen_type, cen_w, dir_type, dir_w, dis_wgt,
                 knl_order, knl_width, time_stp, time_num, ifo=None, tgt_nde=None):
        self.tl = None if tgt_nde is None else _f64(tgt_nde)
        self.q0 = _f64(q0)
        self.elm = _i32(elm)
        self.ifo = _i32(ifo) if ifo is not None else get_elm_ifo(self.elm, self.q0.shape[0])
        self.tc, self.tu, self.ta = _f64(tgt_cen), _f64(tgt_dir), _f64(tgt_vol)
        n, d = self.q0.shape
        self.dim, self.n, self.T = d, n, int(time_num)
        self.varnum = d * n * (self.T - 1)
        P = Problem()
        P.dim, P.ndeNum, P.timeNum, P.knlOrder = d, n, int(time_num), int(knl_order)
        P.knlWidth, P.timeStp, P.disWgt = knl_width, time_stp, dis_wgt
        P.iniNdeMat, P.elmVtx, P.elmIfo = self.q0.ctypes.data, self.elm.ctypes.data, self.ifo.ctypes.data
        P.elmNum, P.ifoRows, P.tgtNum = self.elm.shape[0], self.ifo.shape[1], len(self.ta)
        P.tgtCen, P.tgtDir, P.tgtVol = self.tc.ctypes.data, self.tu.ctypes.data, self.ta.ctypes.data
        P.cenType, P.dirType = cen_type.encode(), dir_type.encode()
        P.cenWidth, P.dirWidth = cen_w, dir_w
        P.tgtNde = self.tl.ctypes.data if self.tl is not None else None
        self.P = P

    def objgrd(self, alp, want_traj=False):
        alp = _f64(alp).reshape(-1)
        g = np.empty(self.varnum)
        f = C.c_double()
        traj = np.empty((self.T, self.n, self.dim)) if want_traj else None
        _chk(lib().orc_objgrd(C.byref(f), g, alp, C.byref(self.P), traj.ctypes.data if want_traj else None))
        return (f.value, g, traj) if want_traj else (f.value, g)

    def run(self, alp0=None, itr_max=200, tol=0.0, wolfe1=0.0, wolfe2=0.5, mem=5, verbose=False):
        x0 = np.zeros(self.varnum) if alp0 is None else _f64(alp0).reshape(-1)
        out = np.empty(self.varnum)
        tr, keep = make_trace(itr_max)
        o = LbfgsOpts(itr_max, mem, int(verbose), tol, wolfe1, wolfe2)
        f = C.c_double()
        st = lib().orc_lddmm_run(C.byref(self.P), x0, C.byref(o), out, C.byref(f), C.byref(tr))
        return out, f.value, trace_dict(tr, keep, st)


def make_trace(itr_max):
    keep = dict(f=np.zeros(itr_max + 1), gnorm=np.zeros(itr_max + 1), step=np.zeros(itr_max + 1),
                evals=np.zeros(itr_max + 1, dtype=np.int32))
    tr = Trace(keep["f"].ctypes.data, keep["gnorm"].ctypes.data, keep["step"].ctypes.data,
               keep["evals"].ctypes.data, 0, 0, 0)
    return tr, keep


def trace_dict(tr, keep, status):
    k = tr.iters
    return dict(f=keep["f"][:k + 1].copy(), gnorm=keep["gnorm"][:k + 1].copy(), step=keep["step"][:k + 1].copy(),
                evals=keep["evals"][:k + 1].copy(), iters=k, status=status, total_evals=tr.totalEvals)


def lbfgs(fun, x0, itr_max=100, tol=1e-10, wolfe1=1e-4, wolfe2=0.9, mem=5):
    """L-BFGS on a Python objective fun(x) -> (f, g); for optimiser tests."""
    x = _f64(x0).copy()
    n = x.size

    @OBJFUN
    def cb(_ctx, fp, gp, xp):
        xv = np.ctypeslib.as_array(xp, shape=(n,))
        f, g = fun(xv.copy())
        fp[0] = f
        np.ctypeslib.as_array(gp, shape=(n,))[:] = g

    g = np.empty(n)
    f = C.c_double()
    tr, keep = make_trace(itr_max)
    o = LbfgsOpts(itr_max, mem, 0, tol, wolfe1, wolfe2)
    st = lib().orc_lbfgs(cb, None, C.byref(f), g, x, n, C.byref(o), C.byref(tr))
    return x, f.value, trace_dict(tr, keep, st)


def line_search(fun, x, g, d, f0, c1, c2, tol=0.0):
    """One weak-Wolfe line search on fun(x) -> (f, g); returns
    (code, x_new, g_new, f_new, step, count)."""
    x, g, d = _f64(x), _f64(g), _f64(d)
    n = x.size

    @OBJFUN
    def cb(_ctx, fp, gp, xp):
        f, gg = fun(np.ctypeslib.as_array(xp, shape=(n,)).copy())
        fp[0] = f
        np.ctypeslib.as_array(gp, shape=(n,))[:] = gg

    xt, gt = np.empty(n), np.empty(n)
    ft, step, cnt = C.c_double(), C.c_double(), C.c_int()
    code = lib().orc_line_search(cb, None, x, g, d, f0, c1, c2, tol, xt, gt, C.byref(ft), C.byref(step),
                                 C.byref(cnt), n, 1)
    return code, xt, gt, ft.value, step.value, cnt.value


def get_direction(H0, g, S, Y, new_idx, his_num, mem):
    g = _f64(g)
    d = np.empty_like(g)
    rec = np.zeros(mem)
    lib().orc_get_direction(d, H0, g, _f64(S), _f64(Y), new_idx, his_num, mem, rec, g.size)
    return d
