"""lddmm_amd -- LDDMM landmark/curve/surface registration with a varifold
matching term on AMD Instinct MI355X (gfx950).

The compute path is liblddmm_amd.so (hand-written HIP kernels behind the
C-ABI in include/lddmm.h).  This module is a thin ctypes mirror of the
reference's MATLAB-facing interface (dnhsieh/LDDMM Cpp/src/LDDMM.m,
multiplyKernel MEX, getElmIfo.m, computeNodePosition.m) so driver code and
tests read like the reference's.  There is no CPU fallback: every call goes
to the GPU library and raises when it is missing or fails.

Array conventions: a MATLAB D x N matrix is a numpy array of shape (N, D)
in C order (the same memory); a D x N x (T-1) stack is shape (T-1, N, D);
element tables are (E, V).  `LDDMM()` and `getElmIfo()` take MATLAB's
1-based element tables, like the reference's .m files; the lower-level
functions take 0-based int32 tables, like the MEX entry points.
"""
import ctypes as C
import os

import numpy as np

__all__ = ["lib", "LDDMM", "getElmIfo", "computeNodePosition", "elementVarifold", "multiplyKernel", "dqKernel", "computeKernel",
           "varifold", "Objective", "run", "l2Dist", "run_l2", "spatial_order", "set_pair_culling", "set_pair_expansion", "LddmmError", "profile_enable", "profile_read", "profile_reset"]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblddmm_amd.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class LddmmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("lddmm error %d: %s" % (code, msg))
        self.code = code


class Options(C.Structure):
    _fields_ = [("precision", C.c_int), ("no_fp64_refine", C.c_int), ("rel_energy_tol", C.c_double),
                ("refine_rel_tol", C.c_double), ("target_energy", C.c_double), ("has_target", C.c_int),
                ("pair_target_blocks", C.c_int), ("progress", C.c_int), ("reserved", C.c_int * 3)]


class Result(C.Structure):
    _fields_ = [("fcn", C.c_double), ("grad_norm", C.c_double), ("seconds", C.c_double),
                ("pairs_per_eval", C.c_double), ("iterations", C.c_int), ("evaluations", C.c_int),
                ("status", C.c_int), ("fp32_iterations", C.c_int), ("reached_target", C.c_int),
                ("reserved", C.c_int * 3)]


PRECISIONS = {"fp64": 0, "fp32": 1}

_RUN_ARGS = [_dp, _dp, _ip, _ip, _dp, _dp, _dp, C.c_char, C.c_double, C.c_char, C.c_double, C.c_double,
             C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
             C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp]


def lib():
    """The HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("liblddmm_amd.so is not built (run `python -m lddmm_amd.build`); "
                               "there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        L.lddmm_last_error.restype = C.c_char_p
        L.lddmm_run_2d.argtypes = _RUN_ARGS
        L.lddmm_run_3d.argtypes = _RUN_ARGS
        L.lddmm_run_ex.argtypes = [C.c_int] + _RUN_ARGS + [C.POINTER(Options), C.POINTER(Result)]
        L.lddmm_run_dev.argtypes = [C.c_int] + [C.c_void_p] * 7 + _RUN_ARGS[7:26] + [
            C.c_void_p, C.POINTER(Options), C.POINTER(Result), C.c_void_p]
        L.lddmm_objective_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, _dp, _ip, _ip, _dp, _dp, _dp,
                                             C.c_char, C.c_double, C.c_char, C.c_double, C.c_double, C.c_int,
                                             C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.lddmm_objgrd.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.lddmm_objective_trajectory.argtypes = [C.c_void_p, _dp]
        L.lddmm_objgrd_dev.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.lddmm_objective_synchronize.argtypes = [C.c_void_p]
        L.lddmm_objective_pairs.argtypes = [C.c_void_p, _dp]
        L.lddmm_objective_target_term.argtypes = [C.c_void_p, _dp]
        L.lddmm_objective_set_precision.argtypes = [C.c_void_p, C.c_int]
        L.lddmm_objective_destroy.argtypes = [C.c_void_p]
        L.lddmm_objective_destroy.restype = None
        L.lddmm_multiply_kernel.argtypes = [_dp, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int]
        L.lddmm_multiply_kernel_cross.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int,
                                                  C.c_int]
        L.lddmm_dq_kernel.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int]
        L.lddmm_dq_kernel_cross.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int,
                                            C.c_int]
        L.lddmm_compute_kernel.argtypes = [_dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int]
        L.lddmm_compute_kernel_cross.argtypes = [_dp, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int]
        L.lddmm_varifold.argtypes = [_dp, _dp, _dp, _ip, _ip, _dp, _dp, _dp, C.c_char, C.c_double, C.c_char,
                                     C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.lddmm_get_elm_ifo.argtypes = [_ip, _ip, C.c_int, C.c_int, C.c_int]
        L.lddmm_spatial_order.argtypes = [_ip, _dp, C.c_int, C.c_int]
        L.lddmm_set_pair_culling.argtypes = [C.c_int]
        L.lddmm_set_pair_expansion.argtypes = [C.c_int]
        L.lddmm_compute_node_position.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_double, C.c_double, C.c_int,
                                                  C.c_int, C.c_int]
        L.lddmm_device_count.argtypes = [C.POINTER(C.c_int)]
        L.lddmm_device_malloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.lddmm_device_free.argtypes = [C.c_void_p]
        L.lddmm_memcpy_htod.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.lddmm_memcpy_dtoh.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.lddmm_profile_enable.argtypes = [C.c_int]
        L.lddmm_profile_read.argtypes = [C.POINTER(C.c_char_p), _dp, C.POINTER(C.c_longlong), C.c_int]
        _lib = L
    return _lib


def _check(rc, allow_positive=False):
    if rc < 0 or (rc > 0 and not allow_positive):
        raise LddmmError(rc, lib().lddmm_last_error().decode(errors="replace"))
    return rc


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp if a.dtype == np.float64 else _ip)


def _ch(c):
    return c.encode() if isinstance(c, str) else bytes([c])


# ---------------------------------------------------------------- data formats
def getElmIfo(elmVtxMat, ndeNum, one_based=True):
    """Adjacency table of getElmIfo.m (Cpp/src/getElmIfo.m:1-59): returns an
    int32 array of shape (ndeNum, rows) (MATLAB's rows x ndeNum)."""
    elm = _i32(np.asarray(elmVtxMat) - (1 if one_based else 0))
    ne, vpe = elm.shape
    rows = _check(lib().lddmm_get_elm_ifo(None, _p(elm), vpe, ne, ndeNum), allow_positive=True)
    ifo = np.zeros((ndeNum, rows), dtype=np.int32)
    _check(lib().lddmm_get_elm_ifo(_p(ifo), _p(elm), vpe, ne, ndeNum), allow_positive=True)
    return ifo


def spatial_order(pts):
    """The internal order of a point set (N x D): perm[s] = the point in slot
    s; every aligned block of 64 or 128 slots is spatially compact.  Every
    entry point takes and returns the caller's order; this shows the order
    the library computes in."""
    q = _f64(pts)
    perm = np.empty(q.shape[0], dtype=np.int32)
    _check(lib().lddmm_spatial_order(_p(perm), _p(q), q.shape[0], q.shape[1]))
    return perm


def set_pair_culling(on):
    """Culling of FP32 Gaussian tile pairs whose kernel values are all below
    2^-48 (default on); off computes every pair, with results that agree to
    a few FP64 ulps."""
    _check(lib().lddmm_set_pair_culling(int(bool(on))))


def set_pair_cull_level(bits):
    """Skip FP32 Gaussian tile pairs whose kernel values are all below 2^-bits
    (default 48; 24..126)."""
    _check(lib().lddmm_set_pair_cull_level(int(bits)))


def set_pair_groups(on):
    """Culling by precomputed 64-source group bounds, surviving groups packed
    four to a tile (default on); off tests each staged quarter tile."""
    _check(lib().lddmm_set_pair_groups(int(bool(on))))


def set_pair_expansion(on):
    """Expanded form of the FP32 Gaussian pair terms (default on): r^2 as
    |x_i|^2 + (|x_j|^2 - 2 x_i.x_j) with the factor exp2(-|x_i|^2) applied
    per target; off uses the direct form (results differ by FP32 rounding)."""
    _check(lib().lddmm_set_pair_expansion(int(bool(on))))


def measure_rates():
    """Measured chip rates: dict(exp2_f32=v_exp_f32/s, fma_f32=FP32 FMA/s, fma_f64=FP64 FMA/s)."""
    r = (C.c_double * 3)()
    _check(lib().lddmm_measure_rates(r))
    return dict(exp2_f32=r[0], fma_f32=r[1], fma_f64=r[2])


def pair_count_begin():
    """Start counting the pairs the culling pair kernels compute (diagnostic)."""
    _check(lib().lddmm_pair_count_begin())


def pair_count_end():
    """Stop counting; returns the executed pairs since pair_count_begin() of
    the culling kernels: dict(fwd_pairs=K.alpha, bwd_pairs=adjoint, vf_pairs=varifold)."""
    v = (C.c_ulonglong * 3)()
    _check(lib().lddmm_pair_count_end(v))
    return dict(fwd_pairs=v[0], bwd_pairs=v[1], vf_pairs=v[2])


# ---------------------------------------------------------------- operators
def multiplyKernel(*args):
    """multiplyKernel(lmk, alp, order, width) or
    multiplyKernel(lmki, lmkj, alp, order, width) (Cpp/src/multiplyKernelITF.cpp:16-61)."""
    if len(args) == 4:
        q, a, order, width = args
        q, a = _f64(q), _f64(a)
        v = np.empty_like(q)
        _check(lib().lddmm_multiply_kernel(_p(v), _p(q), _p(a), order, width, q.shape[0], q.shape[1]))
        return v
    if len(args) == 5:
        qi, qj, a, order, width = args
        qi, qj, a = _f64(qi), _f64(qj), _f64(a)
        v = np.empty_like(qi)
        _check(lib().lddmm_multiply_kernel_cross(_p(v), _p(qi), _p(qj), _p(a), order, width, qi.shape[0],
                                                 qj.shape[0], qi.shape[1]))
        return v
    raise TypeError("The number of inputs must be 4 or 5.")


def dqKernel(*args):
    """dqKernel(lmk, lft, rgt, order, width) -> dq, or
    dqKernel(lmki, lmkj, lft, rgt, order, width) -> (dqi, dqj)
    (MATLAB/src/dqKernelITF.cpp:16-72)."""
    if len(args) == 5:
        q, l, r, order, width = args
        q, l, r = _f64(q), _f64(l), _f64(r)
        g = np.empty_like(q)
        _check(lib().lddmm_dq_kernel(_p(g), _p(q), _p(l), _p(r), order, width, q.shape[0], q.shape[1]))
        return g
    if len(args) == 6:
        qi, qj, l, r, order, width = args
        qi, qj, l, r = _f64(qi), _f64(qj), _f64(l), _f64(r)
        gi, gj = np.empty_like(qi), np.empty_like(qj)
        _check(lib().lddmm_dq_kernel_cross(_p(gi), _p(gj), _p(qi), _p(qj), _p(l), _p(r), order, width,
                                           qi.shape[0], qj.shape[0], qi.shape[1]))
        return gi, gj
    raise TypeError("The number of inputs must be 5 or 6.")


def computeKernel(*args):
    """computeKernel(lmk, order, width) -> (N, N) with K[i, j];
    computeKernel(lmki, lmkj, order, width) -> (Nj, Ni) holding the
    reference's Ni x Nj column-major matrix (MATLAB/src/computeKernelITF.cpp:15-59)."""
    if len(args) == 3:
        q, order, width = args
        q = _f64(q)
        K = np.empty((q.shape[0], q.shape[0]))
        _check(lib().lddmm_compute_kernel(_p(K), _p(q), order, width, q.shape[0], q.shape[1]))
        return K
    if len(args) == 4:
        qi, qj, order, width = args
        qi, qj = _f64(qi), _f64(qj)
        K = np.empty((qj.shape[0], qi.shape[0]))
        _check(lib().lddmm_compute_kernel_cross(_p(K), _p(qi), _p(qj), order, width, qi.shape[0], qj.shape[0],
                                                qi.shape[1]))
        return K
    raise TypeError("The number of inputs must be 3 or 4.")


def varifold(q, elm, ifo, tgt_cen, tgt_dir, tgt_vol, cen_type, cen_width, dir_type, dir_width, grad=True):
    """Varifold distance (and vertex gradient) of the deformed set q with
    0-based (E, V) elements against a target given by centres, unit
    directions and volumes (Cpp/src/varifold3D.cpp:393-435)."""
    q, elm, ifo = _f64(q), _i32(elm), _i32(ifo)
    tc, tu, ta = _f64(tgt_cen), _f64(tgt_dir), _f64(tgt_vol)
    val = C.c_double()
    dq = np.empty_like(q) if grad else None
    _check(lib().lddmm_varifold(C.byref(val), _p(dq), _p(q), _p(elm), _p(ifo), _p(tc), _p(tu), _p(ta),
                                _ch(cen_type), cen_width, _ch(dir_type), dir_width, q.shape[0], elm.shape[0],
                                ifo.shape[1], len(ta), q.shape[1]))
    return (val.value, dq) if grad else val.value


def elementVarifold(ndeMat, elmVtxMat):
    """Element centres, unit directions and volumes of a mesh (the target
    preprocessing of Cpp/examples/circle2square/run_circle2square.m:21-24;
    landmarksToVarifoldKernel, Cpp/src/varifold3D.cpp:8-39).  ndeMat (N, D),
    elmVtxMat (E, D) 0-based -> (cen (E, D), dir (E, D), vol (E,))."""
    L = lib()
    if not hasattr(L.lddmm_element_varifold, "_set"):
        L.lddmm_element_varifold.argtypes = [_dp, _dp, _dp, _dp, _ip, C.c_int, C.c_int, C.c_int]
        L.lddmm_element_varifold._set = True
    q, e = _f64(ndeMat), _i32(elmVtxMat)
    n, d = q.shape
    cen, udir, vol = np.empty((len(e), d)), np.empty((len(e), d)), np.empty(len(e))
    _check(L.lddmm_element_varifold(_p(cen), _p(udir), _p(vol), _p(q), _p(e), n, len(e), d))
    return cen, udir, vol


def computeNodePosition(iniNdeMat, alpStk, knlOrder, knlWidth, timeStp, timeNum):
    """Trajectory and velocities of the Euler flow (Cpp/src/computeNodePosition.m:1-20):
    returns ndeStk (T, N, D) and vlcStk (T-1, N, D)."""
    q0, a = _f64(iniNdeMat), _f64(alpStk)
    n, d = q0.shape
    nde = np.empty((timeNum, n, d))
    vlc = np.empty((max(timeNum - 1, 0), n, d))
    _check(lib().lddmm_compute_node_position(_p(nde), _p(vlc), _p(q0), _p(a), knlOrder, knlWidth, timeStp,
                                             timeNum, n, d))
    return nde, vlc


def l2Dist(ndeMat, tgtMat, wgtVal):
    """[fcnVal, grdMat] = l2Dist(ndeMat, tgtMat, wgtVal) of MATLAB/src/l2Dist.m:1-5
    on the GPU: w/2 |q - tgt|^2 and w (q - tgt); arrays (N, D)."""
    L = lib()
    if not hasattr(L.lddmm_l2_dist, "_set"):
        L.lddmm_l2_dist.argtypes = [_dp, _dp, _dp, _dp, C.c_double, C.c_int, C.c_int]
        L.lddmm_objective_create_l2.argtypes = [C.POINTER(C.c_void_p), C.c_int, _dp, _dp, C.c_double, C.c_int,
                                                C.c_double, C.c_double, C.c_int, C.c_int]
        L.lddmm_run_l2.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_int, C.c_double, C.c_double, C.c_int,
                                   C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, _dp,
                                   C.POINTER(Options), C.POINTER(Result)]
        L.lddmm_l2_dist._set = True
    q, t = _f64(ndeMat), _f64(tgtMat)
    if q.shape != t.shape:
        raise ValueError("ndeMat and tgtMat must have the same shape")
    f = C.c_double()
    g = np.empty_like(q)
    _check(L.lddmm_l2_dist(C.byref(f), _p(g), _p(q), _p(t), float(wgtVal), q.shape[0], q.shape[1]))
    return f.value, g


def run_l2(q0, alp0, tgt, dis_wgt, knl_order, knl_width, time_stp, time_num, itr_max, tol, wolfe1, wolfe2, mem,
           verbose=False, precision="fp64", rel_energy_tol=0.0):
    """Landmark registration with the l2Dist matching term (MATLAB/src/LDDMM.m
    with disHdl = @(q) l2Dist(q, tgt, w)).  Returns (alpha (T-1, N, D), info)."""
    l2Dist(np.zeros((1, 2)), np.zeros((1, 2)), 0.0)   # binds the signatures
    q0, tgt = _f64(q0), _f64(tgt)
    n, d = q0.shape
    T = int(time_num)
    a0 = np.zeros(d * n * (T - 1)) if alp0 is None else _f64(alp0).reshape(-1)
    out = np.empty(d * n * (T - 1))
    opts = Options(PRECISIONS[precision], 0, rel_energy_tol, 0.0, 0.0, 0, 0, 0)
    res = Result()
    st = _check(lib().lddmm_run_l2(d, _p(q0), _p(a0), _p(tgt), dis_wgt, knl_order, knl_width, time_stp, T, itr_max,
                                   tol, wolfe1, wolfe2, mem, int(bool(verbose)), n, _p(out), C.byref(opts),
                                   C.byref(res)), allow_positive=True)
    info = dict(fcn=res.fcn, grad_norm=res.grad_norm, seconds=res.seconds, iterations=res.iterations,
                evaluations=res.evaluations, status=st)
    return out.reshape(T - 1, n, d), info


# ---------------------------------------------------------------- objective
class Objective:
    """objgrd (Cpp/src/objgrd.cpp:26-90) on the GPU for a fixed problem.
    elm is a 0-based (E, V) table; ifo the (N, rows) adjacency table."""

    def __init__(self, q0, elm, ifo, tgt_cen, tgt_vol, tgt_dir, cen_type, cen_width, dir_type, dir_width,
                 dis_wgt, knl_order, knl_width, time_stp, time_num, precision="fp64"):
        self.q0, self.elm, self.ifo = _f64(q0), _i32(elm), _i32(ifo)
        self.tc, self.ta, self.tu = _f64(tgt_cen), _f64(tgt_vol), _f64(tgt_dir)
        self.n, self.dim = self.q0.shape
        self.T = int(time_num)
        self.varnum = self.dim * self.n * (self.T - 1)
        h = C.c_void_p()
        _check(lib().lddmm_objective_create(C.byref(h), self.dim, _p(self.q0), _p(self.elm), _p(self.ifo),
                                            _p(self.tc), _p(self.ta), _p(self.tu), _ch(cen_type), cen_width,
                                            _ch(dir_type), dir_width, dis_wgt, knl_order, knl_width, time_stp,
                                            self.T, self.n, self.elm.shape[0], self.ifo.shape[1], len(self.ta)))
        self._h = h
        if precision != "fp64":
            self.set_precision(precision)

    @classmethod
    def l2(cls, q0, tgt, dis_wgt, knl_order, knl_width, time_stp, time_num):
        """The objective with the l2Dist landmark matching term
        (MATLAB/src/objgrd.m with disHdl = l2Dist) against targets tgt (N, D)."""
        l2Dist(np.zeros((1, 2)), np.zeros((1, 2)), 0.0)   # binds the signatures
        self = cls.__new__(cls)
        self.q0, self.tl = _f64(q0), _f64(tgt)
        self.n, self.dim = self.q0.shape
        self.T = int(time_num)
        self.varnum = self.dim * self.n * (self.T - 1)
        h = C.c_void_p()
        _check(lib().lddmm_objective_create_l2(C.byref(h), self.dim, _p(self.q0), _p(self.tl), dis_wgt, knl_order,
                                               knl_width, time_stp, self.T, self.n))
        self._h = h
        return self

    def set_precision(self, precision):
        """'fp64' (the reference's arithmetic) or 'fp32' (FP32 pair terms,
        FP64 differences, cross-tile sums, energies and reductions)."""
        _check(lib().lddmm_objective_set_precision(self._h, PRECISIONS[precision]))

    def __call__(self, alp):
        a = _f64(alp).reshape(-1)
        if a.size != self.varnum:
            raise ValueError("alpha must hold D*N*(T-1) = %d values" % self.varnum)
        g = np.empty(self.varnum)
        f = C.c_double()
        _check(lib().lddmm_objgrd(self._h, C.byref(f), _p(g), _p(a)))
        return f.value, g

    def eval_device(self, d_fcn, d_grd, d_alp):
        """Asynchronous evaluation on device arrays (device layout)."""
        _check(lib().lddmm_objgrd_dev(self._h, d_fcn, d_grd, d_alp))

    def synchronize(self):
        _check(lib().lddmm_objective_synchronize(self._h))

    def pairs_per_eval(self):
        v = C.c_double()
        _check(lib().lddmm_objective_pairs(self._h, C.byref(v)))
        return v.value

    def trajectory(self):
        out = np.empty((self.T, self.n, self.dim))
        _check(lib().lddmm_objective_trajectory(self._h, _p(out)))
        return out

    def target_term(self):
        v = C.c_double()
        _check(lib().lddmm_objective_target_term(self._h, C.byref(v)))
        return v.value

    def close(self):
        if getattr(self, "_h", None):
            lib().lddmm_objective_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- registration
def run(q0, alp0, elm, ifo, tgt_cen, tgt_vol, tgt_dir, cen_type, cen_width, dir_type, dir_width, dis_wgt,
        knl_order, knl_width, time_stp, time_num, itr_max, tol, wolfe1, wolfe2, mem, verbose=False,
        rel_energy_tol=0.0, pair_target_blocks=0, precision="fp64", fp64_refine=True, refine_rel_tol=0.0,
        target_energy=None):
    """The LDDMM2D/LDDMM3D MEX call (Cpp/src/LDDMM.cpp:11-98) with 0-based
    tables.  Returns (optAlpVec reshaped (T-1, N, D), result dict)."""
    q0, elm, ifo = _f64(q0), _i32(elm), _i32(ifo)
    tc, ta, tu = _f64(tgt_cen), _f64(tgt_vol), _f64(tgt_dir)
    n, d = q0.shape
    T = int(time_num)
    a0 = np.zeros(d * n * (T - 1)) if alp0 is None else _f64(alp0).reshape(-1)
    out = np.empty(d * n * (T - 1))
    opts = Options(PRECISIONS[precision], 0 if fp64_refine else 1, rel_energy_tol, refine_rel_tol,
                   0.0 if target_energy is None else float(target_energy), int(target_energy is not None),
                   pair_target_blocks)
    res = Result()
    st = _check(lib().lddmm_run_ex(d, _p(q0), _p(a0), _p(elm), _p(ifo), _p(tc), _p(ta), _p(tu), _ch(cen_type),
                                   cen_width, _ch(dir_type), dir_width, dis_wgt, knl_order, knl_width, time_stp, T,
                                   itr_max, tol, wolfe1, wolfe2, mem, int(bool(verbose)), n, elm.shape[0],
                                   ifo.shape[1], len(ta), _p(out), C.byref(opts), C.byref(res)),
                allow_positive=True)
    info = dict(fcn=res.fcn, grad_norm=res.grad_norm, seconds=res.seconds, pairs_per_eval=res.pairs_per_eval,
                iterations=res.iterations, evaluations=res.evaluations, status=st,
                fp32_iterations=res.fp32_iterations, reached_target=bool(res.reached_target))
    return out.reshape(T - 1, n, d), info


def LDDMM(iniNdeMat, iniAlpStk, disObj, knlOrder, knlWidth, timeEnd, timeStp, optObj):
    """[ndeStk, vlcStk, alpStk] = LDDMM(...) of Cpp/src/LDDMM.m:1-46.
    disObj / optObj are dicts with the reference's field names; the element
    table disObj['dfmElmVtxMat'] is 1-based (E, V) as in MATLAB."""
    q0 = _f64(iniNdeMat)
    n, d = q0.shape
    T = int(np.floor(timeEnd / timeStp)) + 1
    elm1 = np.asarray(disObj["dfmElmVtxMat"])
    ifo = getElmIfo(elm1, n)
    elm = _i32(elm1 - 1)
    alp0 = None if iniAlpStk is None or np.size(iniAlpStk) == 0 else iniAlpStk
    alp, _ = run(q0, alp0, elm, ifo, disObj["tgtCenPosMat"], np.ravel(disObj["tgtElmVolVec"]),
                 disObj["tgtUniDirMat"], disObj["cenKnlType"], disObj["cenKnlWidth"], disObj["dirKnlType"],
                 disObj["dirKnlWidth"], disObj["disWgt"], knlOrder, knlWidth, timeStp, T,
                 int(optObj["MaxIterations"]), optObj["OptimalityTolerance"], optObj["Wolfe1"], optObj["Wolfe2"],
                 int(optObj["Columns"]), optObj.get("Verbose", False))
    ndeStk, vlcStk = computeNodePosition(q0, alp, knlOrder, knlWidth, timeStp, T)
    return ndeStk, vlcStk, alp


# ---------------------------------------------------------------- profiling
class Batch:
    """Many independent registrations of uneven sizes evaluated together
    (lddmm_batch_*).  problems: list of dicts with q0 (N, D), elm (E, V)
    0-based, tgt_cen, tgt_dir (Et, D), tgt_vol (Et,), knl_width; the kernel
    and varifold types, weight and time grid are shared."""

    def __init__(self, problems, cen_type, cen_width, dir_type, dir_width, dis_wgt, knl_order, time_stp, time_num,
                 precision="fp64", chunk=0):
        L = lib()
        if not hasattr(L.lddmm_batch_create, "_set"):
            L.lddmm_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int, _ip, _ip, _ip, _dp, _ip, _dp,
                                             _dp, _dp, _dp, C.c_char, C.c_double, C.c_char, C.c_double, C.c_double,
                                             C.c_int, C.c_double, C.c_int, C.c_int, C.c_int]
            L.lddmm_batch_objgrd.argtypes = [C.c_void_p, _dp, _dp, _dp]
            L.lddmm_batch_run.argtypes = [C.c_void_p, _dp, _dp, C.c_int, C.c_double, C.c_double, C.c_double,
                                          C.c_int, C.POINTER(Options), _dp, C.POINTER(Result), C.POINTER(C.c_int)]
            L.lddmm_batch_objective_scan.argtypes = [C.c_void_p, _dp, _dp, C.c_int, _dp]
            L.lddmm_batch_run_scaled.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_int, C.c_double, C.c_double,
                                                 C.c_double, C.c_int, C.POINTER(Options), _dp, C.POINTER(Result),
                                                 C.POINTER(C.c_int)]
            L.lddmm_batch_info.argtypes = [C.c_void_p, C.POINTER(C.c_int), _dp]
            L.lddmm_batch_set_graphs.argtypes = [C.c_void_p, C.c_int]
            L.lddmm_batch_tuning.argtypes = [C.c_void_p] + [C.POINTER(C.c_int)] * 3
            L.lddmm_batch_set_tuning.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
            L.lddmm_batch_destroy.argtypes = [C.c_void_p]
            L.lddmm_batch_destroy.restype = None
            L.lddmm_batch_create._set = True
        self.P = len(problems)
        self.dim = np.asarray(problems[0]["q0"]).shape[1]
        self.T = int(time_num)
        self.n = np.array([len(p["q0"]) for p in problems], dtype=np.int32)
        ne = np.array([len(p["elm"]) for p in problems], dtype=np.int32)
        nt = np.array([len(p["tgt_vol"]) for p in problems], dtype=np.int32)
        q0 = _f64(np.concatenate([p["q0"] for p in problems]))
        elm = _i32(np.concatenate([p["elm"] for p in problems]))
        tc = _f64(np.concatenate([p["tgt_cen"] for p in problems]))
        tu = _f64(np.concatenate([p["tgt_dir"] for p in problems]))
        ta = _f64(np.concatenate([p["tgt_vol"] for p in problems]))
        widths = _f64([p["knl_width"] for p in problems])
        h = C.c_void_p()
        _check(L.lddmm_batch_create(C.byref(h), self.dim, self.P, _p(self.n), _p(ne), _p(nt), _p(q0), _p(elm),
                                    _p(tc), _p(ta), _p(tu), _p(widths), _ch(cen_type), cen_width, _ch(dir_type),
                                    dir_width, dis_wgt, knl_order, time_stp, self.T, PRECISIONS[precision],
                                    int(chunk)))
        self._h = h
        self.sizes = self.dim * self.n * (self.T - 1)
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)])

    def split(self, flat):
        return [flat[self.offsets[p]:self.offsets[p + 1]].reshape(self.T - 1, self.n[p], self.dim)
                for p in range(self.P)]

    def objgrd(self, alps):
        """alps: list of (T-1, N_p, D) momenta -> (energies, list of gradients)."""
        a = _f64(np.concatenate([np.ravel(x) for x in alps]))
        if a.size != int(self.offsets[-1]):
            raise ValueError("the momenta do not match the batch's sizes")
        f = np.empty(self.P)
        g = np.empty_like(a)
        _check(lib().lddmm_batch_objgrd(self._h, _p(f), _p(g), _p(a)))
        return f, self.split(g)

    def run(self, itr_max, tol, wolfe1, wolfe2, mem, alp0=None, precision="fp64", rel_energy_tol=0.0,
            fp64_refine=True, refine_rel_tol=0.0, targets=None, progress=0):
        a0 = None if alp0 is None else _f64(np.concatenate([np.ravel(x) for x in alp0]))
        if a0 is not None and a0.size != int(self.offsets[-1]):
            raise ValueError("the initial momenta do not match the batch's sizes")
        return self._run(lambda *r: lib().lddmm_batch_run(self._h, _p(a0), *r), itr_max, tol, wolfe1, wolfe2, mem,
                         precision, rel_energy_tol, fp64_refine, refine_rel_tol, targets, progress)

    def _vecs(self, base):
        base = [np.asarray(u, dtype=np.float64) for u in base]
        if len(base) != self.P or any(u.shape != (self.n[p], self.dim) for p, u in enumerate(base)):
            raise ValueError("one (N_p, D) vector array per problem expected")
        return _f64(np.concatenate([np.ravel(u) for u in base]))

    def objective_scan(self, base, scales):
        """Energies (len(scales), P), in the batch's precision, at the momenta
        alpha_t = c u_p for every time slice and each multiple c; base: one
        (N_p, D) array per problem.  Energies only (lddmm_batch_objective_scan)."""
        u = self._vecs(base)
        sc = _f64(scales)
        f = np.empty((len(sc), self.P))
        _check(lib().lddmm_batch_objective_scan(self._h, _p(u), _p(sc), len(sc), _p(f)))
        return f

    def run_scaled(self, base, scales, itr_max, tol, wolfe1, wolfe2, mem, precision="fp64", rel_energy_tol=0.0,
                   fp64_refine=True, refine_rel_tol=0.0, targets=None, progress=0):
        """run() from the start alpha_t = scales[p] u_p (every time slice),
        built on the device (lddmm_batch_run_scaled)."""
        u = self._vecs(base)
        sc = _f64(scales)
        if sc.shape != (self.P,):
            raise ValueError("one start multiple per problem expected")
        return self._run(lambda *r: lib().lddmm_batch_run_scaled(self._h, _p(u), _p(sc), *r), itr_max, tol, wolfe1,
                         wolfe2, mem, precision, rel_energy_tol, fp64_refine, refine_rel_tol, targets, progress)

    def _run(self, call, itr_max, tol, wolfe1, wolfe2, mem, precision, rel_energy_tol, fp64_refine, refine_rel_tol,
             targets, progress):
        out = np.empty(int(self.offsets[-1]))
        opts = Options(PRECISIONS[precision], 0 if fp64_refine else 1, rel_energy_tol, refine_rel_tol, 0.0, 0, 0,
                       int(progress))
        res = (Result * self.P)()
        rounds = C.c_int()
        tg = None if targets is None else _f64(targets)
        if tg is not None and tg.shape != (self.P,):
            raise ValueError("one target energy per problem expected")
        _check(call(_p(out), itr_max, tol, wolfe1, wolfe2, mem, C.byref(opts), _p(tg), res, C.byref(rounds)))
        infos = [dict(fcn=r.fcn, grad_norm=r.grad_norm, seconds=r.seconds, pairs_per_eval=r.pairs_per_eval,
                      iterations=r.iterations, evaluations=r.evaluations, status=r.status,
                      fp32_iterations=r.fp32_iterations, reached_target=bool(r.reached_target)) for r in res]
        return self.split(out), infos, rounds.value

    def set_graphs(self, on):
        """Replay captured hipGraphs for evaluations (default) or launch directly."""
        _check(lib().lddmm_batch_set_graphs(self._h, int(bool(on))))

    def info(self):
        ch = C.c_int()
        pairs = np.empty(self.P)
        _check(lib().lddmm_batch_info(self._h, C.byref(ch), _p(pairs)))
        return ch.value, pairs

    def tuning(self):
        """(chunk, kv_lanes, bwd_lanes): source points per work item and FP32
        targets per lane of the K.alpha and adjoint pair launches."""
        v = [C.c_int() for _ in range(3)]
        _check(lib().lddmm_batch_tuning(self._h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def set_tuning(self, chunk, kv_lanes, bwd_lanes):
        """Fix the launch shape (see tuning)."""
        _check(lib().lddmm_batch_set_tuning(self._h, int(chunk), int(kv_lanes), int(bwd_lanes)))

    def close(self):
        if getattr(self, "_h", None):
            lib().lddmm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A device allocation made through the library (no torch involved)."""

    def __init__(self, host=None, nbytes=None):
        self.nbytes = int(host.nbytes if host is not None else nbytes)
        self.ptr = C.c_void_p()
        _check(lib().lddmm_device_malloc(C.byref(self.ptr), max(self.nbytes, 8)))
        if host is not None:
            h = np.ascontiguousarray(host)
            _check(lib().lddmm_memcpy_htod(self.ptr, h.ctypes.data, self.nbytes))

    def get(self, dtype=np.float64):
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        _check(lib().lddmm_memcpy_dtoh(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().lddmm_device_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def profile_enable(on=True):
    _check(lib().lddmm_profile_enable(int(bool(on))))


def profile_reset():
    _check(lib().lddmm_profile_reset())


def profile_read():
    """{group: (milliseconds, launches)} of the HIP-event kernel accounting."""
    names = (C.c_char_p * 32)()
    ms = (C.c_double * 32)()
    cnt = (C.c_longlong * 32)()
    k = _check(lib().lddmm_profile_read(names, ms, cnt, 32), allow_positive=True)
    return {names[i].decode(): (ms[i], cnt[i]) for i in range(k)}
