// capi.cpp -- the extern "C" boundary declared in include/lddmm.h.
//
// Host-layout entry points transpose to the device (SoA) layout on upload
// and back on download; the device-layout entry points take the caller's
// device pointers as they are (CUDA/src/LDDMM.m:7-19 is the layout
// contract).  Exceptions never cross the boundary: every entry point
// returns a status and keeps the message for lddmm_last_error().
#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/lddmm.h"
#include "batch.hpp"
#include "engine.hpp"
#include "kernels.hpp"
#include "mlp.hpp"

using namespace lddmm;

namespace {

thread_local std::string g_err;

template <typename F>
int guarded(F &&f)
{
    try {
        g_err.clear();
        return f();
    } catch (const Error &e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_err = "host allocation failed";
        return kErrAlloc;
    } catch (const std::exception &e) {
        g_err = e.what();
        return kErrArg;
    }
}

void require(bool ok, const char *msg)
{
    if (!ok) throw Error(kErrArg, msg);
}

// RAII device buffer
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    void alloc(size_t count)
    {
        n = count;
        if (count == 0) return;
        if (hipMalloc((void **)&p, count * sizeof(T)) != hipSuccess) {
            p = nullptr;
            throw Error(kErrAlloc, "device allocation of " + std::to_string(count * sizeof(T)) + " bytes failed");
        }
    }
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    void upload(const T *src, hipStream_t st)
    {
        if (n) LDDMM_HIP(hipMemcpyAsync(p, src, n * sizeof(T), hipMemcpyHostToDevice, st));
    }
    void download(T *dst, hipStream_t st) const
    {
        if (n) LDDMM_HIP(hipMemcpyAsync(dst, p, n * sizeof(T), hipMemcpyDeviceToHost, st));
    }
};

struct Stream {
    hipStream_t s = nullptr;
    Stream() { LDDMM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
    ~Stream()
    {
        if (s) (void)hipStreamDestroy(s);
    }
    void sync() { LDDMM_HIP(hipStreamSynchronize(s)); }
};

// AoS (D x N column-major) -> SoA (N x D column-major), `slices` stacked sets
template <typename T>
std::vector<T> to_soa(const T *src, int n, int d, int slices = 1)
{
    std::vector<T> out((size_t)n * d * slices);
    for (int t = 0; t < slices; ++t) {
        const T *s = src + (size_t)t * n * d;
        T *o = out.data() + (size_t)t * n * d;
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < d; ++c) o[(size_t)c * n + i] = s[(size_t)i * d + c];
    }
    return out;
}

template <typename T>
void from_soa(T *dst, const T *soa, int n, int d, int slices = 1)
{
    for (int t = 0; t < slices; ++t) {
        const T *s = soa + (size_t)t * n * d;
        T *o = dst + (size_t)t * n * d;
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < d; ++c) o[(size_t)i * d + c] = s[(size_t)c * n + i];
    }
}

// Host-side problem arrays converted to the device layout and uploaded.
struct UploadedProblem {
    DevBuf<double> q0, tc, tu, ta, tl;
    DevBuf<int> elm, ifo;
    DeviceProblem p;
};

// l2 landmark matching problem: source and target landmarks (host layout)
void upload_l2_problem(UploadedProblem &u, int dim, int n, int T, int order, double width, double h, double w,
                       const double *q0, const double *tgt, hipStream_t st)
{
    require(dim == 2 || dim == 3, "dimNum must be 2 or 3");
    require(n > 0, "ndeNum must be positive");
    require(T >= 2, "timeNum must be at least 2");
    require(q0 && tgt, "null landmark array");
    DeviceProblem &p = u.p;
    p.dim = dim;
    p.n = n;
    p.T = T;
    p.order = order;
    p.width = width;
    p.h = h;
    p.w = w;
    p.ne = p.nt = 0;
    auto hq = to_soa(q0, n, dim), ht = to_soa(tgt, n, dim);
    u.q0.alloc(hq.size());
    u.tl.alloc(ht.size());
    u.q0.upload(hq.data(), st);
    u.tl.upload(ht.data(), st);
    LDDMM_HIP(hipStreamSynchronize(st));   // host vectors go out of scope
    p.q0 = u.q0.p;
    p.tl = u.tl.p;
    validate(p);
}

void check_common(int dim, int n, int ne, int ifo_rows, int nt, int T)
{
    require(dim == 2 || dim == 3, "dimNum must be 2 or 3");
    require(n > 0, "ndeNum must be positive");
    require(ne > 0, "dfmElmNum must be positive");
    require(nt >= 0, "tgtElmNum must be non-negative");
    require(ifo_rows >= 1, "dfmElmIfoNum must be at least 1");
    require(T >= 2, "timeNum must be at least 2");
}

void fill_params(DeviceProblem &p, int dim, int n, int T, int order, double width, double h, double w, char ct,
                 double cw, char dt, double dw, int ne, int nt, int rows)
{
    p.dim = dim;
    p.n = n;
    p.T = T;
    p.order = order;
    p.width = width;
    p.h = h;
    p.w = w;
    p.vp = VfParams{ct, dt, cw, dw};
    p.ne = ne;
    p.nt = nt;
    p.ifo_rows = rows;
}

void upload_host_problem(UploadedProblem &u, int dim, const double *q0, const int32_t *elm, const int32_t *ifo,
                         const double *tc, const double *ta, const double *tu, int n, int ne, int rows, int nt,
                         hipStream_t st)
{
    require(q0 && elm && ifo, "null input array");
    require(nt == 0 || (tc && ta && tu), "null target array");
    const int vpe = dim;   // segments in 2D, triangles in 3D (VFDVTXNUM, Cpp/src/include/constants.h)
    auto hq = to_soa(q0, n, dim);
    auto helm = to_soa(elm, ne, vpe);
    auto hifo = to_soa(ifo, n, rows);
    for (int v : helm) require(v >= 0 && v < n, "dfmElmVtxMat holds a vertex index outside [0, ndeNum)");
    for (int k = 0; k < n; ++k) {
        int adj = hifo[k];
        require(adj >= 0 && 1 + (long long)adj * vpe <= rows, "dfmElmIfoMat count exceeds its row count");
        for (int a = 0; a < adj; ++a) {
            int e = hifo[(size_t)(1 + a * vpe) * n + k];
            require(e >= 0 && e < ne, "dfmElmIfoMat holds an element index outside [0, dfmElmNum)");
            if (vpe == 3)
                for (int z = 1; z <= 2; ++z) {
                    int v = hifo[(size_t)(1 + a * vpe + z) * n + k];
                    require(v >= 0 && v < n, "dfmElmIfoMat holds a vertex index outside [0, ndeNum)");
                }
        }
    }
    u.q0.alloc(hq.size());
    u.q0.upload(hq.data(), st);
    u.elm.alloc(helm.size());
    u.elm.upload(helm.data(), st);
    u.ifo.alloc(hifo.size());
    u.ifo.upload(hifo.data(), st);
    if (nt > 0) {
        auto htc = to_soa(tc, nt, dim), htu = to_soa(tu, nt, dim);
        u.tc.alloc(htc.size());
        u.tc.upload(htc.data(), st);
        u.tu.alloc(htu.size());
        u.tu.upload(htu.data(), st);
        u.ta.alloc(nt);
        u.ta.upload(ta, st);
    }
    LDDMM_HIP(hipStreamSynchronize(st));   // host vectors go out of scope
    u.p.q0 = u.q0.p;
    u.p.elm = u.elm.p;
    u.p.ifo = u.ifo.p;
    u.p.tc = u.tc.p;
    u.p.tu = u.tu.p;
    u.p.ta = u.ta.p;
}

LbfgsOptions make_opts(int itrMax, double tol, double w1, double w2, int mem, int vbs, const lddmm_options *o)
{
    LbfgsOptions lo;
    lo.itr_max = itrMax;
    lo.tol = tol;
    lo.wolfe1 = w1;
    lo.wolfe2 = w2;
    lo.mem = mem;
    lo.verbose = vbs != 0;
    if (o) {
        lo.rel_tol = o->rel_energy_tol;
        lo.fp64_refine = o->no_fp64_refine == 0;
        if (o->refine_rel_tol > 0.0) lo.refine_rel_tol = o->refine_rel_tol;
        lo.has_target = o->has_target != 0;
        lo.target = o->target_energy;
    }
    return lo;
}

int precision_of(const lddmm_options *o)
{
    if (!o) return kFP64;
    require(o->precision == LDDMM_PRECISION_FP64 || o->precision == LDDMM_PRECISION_FP32,
            "precision must be LDDMM_PRECISION_FP64 or LDDMM_PRECISION_FP32");
    return o->precision == LDDMM_PRECISION_FP32 ? kFP32 : kFP64;
}

void fill_result(lddmm_result *r, const LbfgsResult &lr, const Objective &obj)
{
    if (!r) return;
    std::memset(r, 0, sizeof(*r));
    r->fcn = lr.f;
    r->grad_norm = lr.gnorm;
    r->seconds = lr.seconds;
    r->pairs_per_eval = obj.pairs_per_eval();
    r->iterations = lr.iters;
    r->evaluations = lr.evals;
    r->status = lr.status;
    r->fp32_iterations = lr.fp32_iters;
    r->reached_target = lr.reached_target;
}

int run_host(int dim, const double *iniNdeMat, const double *iniAlpVec, const int32_t *elm, const int32_t *ifo,
             const double *tc, const double *ta, const double *tu, char ct, double cw, char dt, double dw,
             double w, int order, double width, double h, int T, int itrMax, double tol, double w1, double w2,
             int mem, int vbs, int n, int ne, int rows, int nt, double *optAlpVec, const lddmm_options *opts,
             lddmm_result *res)
{
    check_common(dim, n, ne, rows, nt, T);
    require(iniAlpVec && optAlpVec, "null momentum array");
    LbfgsOptions lo = make_opts(itrMax, tol, w1, w2, mem, vbs, opts);
    Stream st;
    UploadedProblem u;
    fill_params(u.p, dim, n, T, order, width, h, w, ct, cw, dt, dw, ne, nt, rows);
    u.p.prec = precision_of(opts);
    validate(u.p);
    upload_host_problem(u, dim, iniNdeMat, elm, ifo, tc, ta, tu, n, ne, rows, nt, st.s);
    Objective obj(u.p, st.s, opts ? opts->pair_target_blocks : 0);
    auto hx = to_soa(iniAlpVec, n, dim, T - 1);
    DevBuf<double> x(hx.size());
    x.upload(hx.data(), st.s);
    LbfgsResult lr = run_lbfgs(obj, x.p, lo, st.s);
    x.download(hx.data(), st.s);
    st.sync();
    from_soa(optAlpVec, hx.data(), n, dim, T - 1);
    fill_result(res, lr, obj);
    return lr.status;
}

}  // namespace

struct lddmm_objective {
    Stream st;
    UploadedProblem u;
    std::unique_ptr<Objective> obj;
    DevBuf<double> alp, grd, f;
};

extern "C" {

const char *lddmm_last_error(void) { return g_err.c_str(); }

int lddmm_run_ex(int dimNum, const double *iniNdeMat, const double *iniAlpVec, const int32_t *dfmElmVtxMat,
                 const int32_t *dfmElmIfoMat, const double *tgtCenPosMat, const double *tgtElmVolVec,
                 const double *tgtUniDirMat, char cenKnlType, double cenKnlWidth, char dirKnlType,
                 double dirKnlWidth, double disWgt, int knlOrder, double knlWidth, double timeStp, int timeNum,
                 int itrMax, double tolVal, double wolfe1, double wolfe2, int memNum, int vbsFlg, int ndeNum,
                 int dfmElmNum, int dfmElmIfoNum, int tgtElmNum, double *optAlpVec, const lddmm_options *opts,
                 lddmm_result *result)
{
    return guarded([&] {
        return run_host(dimNum, iniNdeMat, iniAlpVec, dfmElmVtxMat, dfmElmIfoMat, tgtCenPosMat, tgtElmVolVec,
                        tgtUniDirMat, cenKnlType, cenKnlWidth, dirKnlType, dirKnlWidth, disWgt, knlOrder, knlWidth,
                        timeStp, timeNum, itrMax, tolVal, wolfe1, wolfe2, memNum, vbsFlg, ndeNum, dfmElmNum,
                        dfmElmIfoNum, tgtElmNum, optAlpVec, opts, result);
    });
}

int lddmm_run_2d(const double *iniNdeMat, const double *iniAlpVec, const int32_t *dfmElmVtxMat,
                 const int32_t *dfmElmIfoMat, const double *tgtCenPosMat, const double *tgtElmVolVec,
                 const double *tgtUniDirMat, char cenKnlType, double cenKnlWidth, char dirKnlType,
                 double dirKnlWidth, double disWgt, int knlOrder, double knlWidth, double timeStp, int timeNum,
                 int itrMax, double tolVal, double wolfe1, double wolfe2, int memNum, int vbsFlg, int ndeNum,
                 int dfmElmNum, int dfmElmIfoNum, int tgtElmNum, double *optAlpVec)
{
    return lddmm_run_ex(2, iniNdeMat, iniAlpVec, dfmElmVtxMat, dfmElmIfoMat, tgtCenPosMat, tgtElmVolVec,
                        tgtUniDirMat, cenKnlType, cenKnlWidth, dirKnlType, dirKnlWidth, disWgt, knlOrder, knlWidth,
                        timeStp, timeNum, itrMax, tolVal, wolfe1, wolfe2, memNum, vbsFlg, ndeNum, dfmElmNum,
                        dfmElmIfoNum, tgtElmNum, optAlpVec, nullptr, nullptr);
}

int lddmm_run_3d(const double *iniNdeMat, const double *iniAlpVec, const int32_t *dfmElmVtxMat,
                 const int32_t *dfmElmIfoMat, const double *tgtCenPosMat, const double *tgtElmVolVec,
                 const double *tgtUniDirMat, char cenKnlType, double cenKnlWidth, char dirKnlType,
                 double dirKnlWidth, double disWgt, int knlOrder, double knlWidth, double timeStp, int timeNum,
                 int itrMax, double tolVal, double wolfe1, double wolfe2, int memNum, int vbsFlg, int ndeNum,
                 int dfmElmNum, int dfmElmIfoNum, int tgtElmNum, double *optAlpVec)
{
    return lddmm_run_ex(3, iniNdeMat, iniAlpVec, dfmElmVtxMat, dfmElmIfoMat, tgtCenPosMat, tgtElmVolVec,
                        tgtUniDirMat, cenKnlType, cenKnlWidth, dirKnlType, dirKnlWidth, disWgt, knlOrder, knlWidth,
                        timeStp, timeNum, itrMax, tolVal, wolfe1, wolfe2, memNum, vbsFlg, ndeNum, dfmElmNum,
                        dfmElmIfoNum, tgtElmNum, optAlpVec, nullptr, nullptr);
}

int lddmm_run_dev(int dimNum, const double *d_iniNdeMat, const double *d_iniAlpVec, const int32_t *d_elm,
                  const int32_t *d_ifo, const double *d_tc, const double *d_ta, const double *d_tu, char cenKnlType,
                  double cenKnlWidth, char dirKnlType, double dirKnlWidth, double disWgt, int knlOrder,
                  double knlWidth, double timeStp, int timeNum, int itrMax, double tolVal, double wolfe1,
                  double wolfe2, int memNum, int vbsFlg, int ndeNum, int dfmElmNum, int dfmElmIfoNum,
                  int tgtElmNum, double *d_optAlpVec, const lddmm_options *opts, lddmm_result *result,
                  void *stream)
{
    return guarded([&] {
        check_common(dimNum, ndeNum, dfmElmNum, dfmElmIfoNum, tgtElmNum, timeNum);
        require(d_iniNdeMat && d_iniAlpVec && d_elm && d_ifo && d_optAlpVec, "null device array");
        require(tgtElmNum == 0 || (d_tc && d_ta && d_tu), "null target array");
        hipStream_t st = (hipStream_t)stream;
        LbfgsOptions lo = make_opts(itrMax, tolVal, wolfe1, wolfe2, memNum, vbsFlg, opts);
        DeviceProblem p;
        fill_params(p, dimNum, ndeNum, timeNum, knlOrder, knlWidth, timeStp, disWgt, cenKnlType, cenKnlWidth,
                    dirKnlType, dirKnlWidth, dfmElmNum, tgtElmNum, dfmElmIfoNum);
        p.prec = precision_of(opts);
        p.q0 = d_iniNdeMat;
        p.elm = d_elm;
        p.ifo = d_ifo;
        p.tc = d_tc;
        p.ta = d_ta;
        p.tu = d_tu;
        Objective obj(p, st, opts ? opts->pair_target_blocks : 0);
        LDDMM_HIP(hipMemcpyAsync(d_optAlpVec, d_iniAlpVec, sizeof(double) * p.varnum(), hipMemcpyDeviceToDevice,
                                 st));
        LbfgsResult lr = run_lbfgs(obj, d_optAlpVec, lo, st);
        fill_result(result, lr, obj);
        return lr.status;
    });
}

// ---------------------------------------------------------------- objective handle
int lddmm_objective_create(lddmm_objective **out, int dimNum, const double *iniNdeMat, const int32_t *dfmElmVtxMat,
                           const int32_t *dfmElmIfoMat, const double *tgtCenPosMat, const double *tgtElmVolVec,
                           const double *tgtUniDirMat, char cenKnlType, double cenKnlWidth, char dirKnlType,
                           double dirKnlWidth, double disWgt, int knlOrder, double knlWidth, double timeStp,
                           int timeNum, int ndeNum, int dfmElmNum, int dfmElmIfoNum, int tgtElmNum)
{
    return guarded([&] {
        require(out != nullptr, "null handle pointer");
        *out = nullptr;
        check_common(dimNum, ndeNum, dfmElmNum, dfmElmIfoNum, tgtElmNum, timeNum);
        auto h = std::make_unique<lddmm_objective>();
        fill_params(h->u.p, dimNum, ndeNum, timeNum, knlOrder, knlWidth, timeStp, disWgt, cenKnlType, cenKnlWidth,
                    dirKnlType, dirKnlWidth, dfmElmNum, tgtElmNum, dfmElmIfoNum);
        validate(h->u.p);
        upload_host_problem(h->u, dimNum, iniNdeMat, dfmElmVtxMat, dfmElmIfoMat, tgtCenPosMat, tgtElmVolVec,
                            tgtUniDirMat, ndeNum, dfmElmNum, dfmElmIfoNum, tgtElmNum, h->st.s);
        h->obj = std::make_unique<Objective>(h->u.p, h->st.s);
        h->alp.alloc(h->u.p.varnum());
        h->grd.alloc(h->u.p.varnum());
        h->f.alloc(1);
        *out = h.release();
        return 0;
    });
}

int lddmm_objgrd(lddmm_objective *obj, double *fcnVal, double *grdStk, const double *alpStk)
{
    return guarded([&] {
        require(obj && fcnVal && grdStk && alpStk, "null argument");
        const DeviceProblem &p = obj->u.p;
        auto ha = to_soa(alpStk, p.n, p.dim, p.T - 1);
        obj->alp.upload(ha.data(), obj->st.s);
        obj->obj->eval(obj->alp.p, obj->grd.p, obj->f.p);
        obj->grd.download(ha.data(), obj->st.s);
        obj->f.download(fcnVal, obj->st.s);
        obj->st.sync();
        from_soa(grdStk, ha.data(), p.n, p.dim, p.T - 1);
        return 0;
    });
}

int lddmm_objgrd_dev(lddmm_objective *obj, double *d_fcnVal, double *d_grdStk, const double *d_alpStk)
{
    return guarded([&] {
        require(obj && d_fcnVal && d_grdStk && d_alpStk, "null argument");
        obj->obj->eval(d_alpStk, d_grdStk, d_fcnVal);
        return 0;
    });
}

int lddmm_objective_synchronize(lddmm_objective *obj)
{
    return guarded([&] {
        require(obj != nullptr, "null argument");
        obj->st.sync();
        return 0;
    });
}

int lddmm_objective_pairs(lddmm_objective *obj, double *pairsPerEval)
{
    return guarded([&] {
        require(obj && pairsPerEval, "null argument");
        *pairsPerEval = obj->obj->pairs_per_eval();
        return 0;
    });
}

int lddmm_objective_trajectory(lddmm_objective *obj, double *ndeStk)
{
    return guarded([&] {
        require(obj && ndeStk, "null argument");
        const DeviceProblem &p = obj->u.p;
        std::vector<double> h((size_t)p.dim * p.n * p.T);
        DevBuf<double> d(h.size());
        obj->obj->trajectory_to(d.p);
        d.download(h.data(), obj->st.s);
        obj->st.sync();
        from_soa(ndeStk, h.data(), p.n, p.dim, p.T);
        return 0;
    });
}

int lddmm_objective_set_precision(lddmm_objective *obj, int precision)
{
    return guarded([&] {
        require(obj != nullptr, "null argument");
        require(precision == LDDMM_PRECISION_FP64 || precision == LDDMM_PRECISION_FP32, "unknown precision");
        obj->st.sync();
        obj->u.p.prec = precision == LDDMM_PRECISION_FP32 ? kFP32 : kFP64;
        obj->obj.reset();
        obj->obj = std::make_unique<Objective>(obj->u.p, obj->st.s);
        return 0;
    });
}

int lddmm_objective_target_term(lddmm_objective *obj, double *value)
{
    return guarded([&] {
        require(obj && value, "null argument");
        *value = obj->obj->target_target();
        return 0;
    });
}

void lddmm_objective_destroy(lddmm_objective *obj) { delete obj; }

// ---------------------------------------------------------------- l2 landmark matching
int lddmm_l2_dist(double *fcnVal, double *grdMat, const double *ndeMat, const double *tgtMat, double wgtVal,
                  int ndeNum, int dimNum)
{
    return guarded([&] {
        require(fcnVal && grdMat && ndeMat && tgtMat, "null argument");
        require(dimNum == 2 || dimNum == 3, "dimNum must be 2 or 3");
        require(ndeNum > 0, "ndeNum must be positive");
        Stream st;
        auto hq = to_soa(ndeMat, ndeNum, dimNum), ht = to_soa(tgtMat, ndeNum, dimNum);
        const int nb = l2_match_blocks(ndeNum);
        DevBuf<double> q(hq.size()), t(ht.size()), p(hq.size()), parts(nb), zero(1), f(1), e(1);
        q.upload(hq.data(), st.s);
        t.upload(ht.data(), st.s);
        const double z = 0.0;
        zero.upload(&z, st.s);
        e.upload(&z, st.s);
        launch_l2_match(dimNum, q.p, t.p, ndeNum, wgtVal, p.p, parts.p, st.s);
        // energy kernel with no flow partials: f = w * (sum parts + 0)
        launch_energy(e.p, 0, parts.p, nb, zero.p, 0.0, wgtVal, f.p, st.s);
        p.download(hq.data(), st.s);
        f.download(fcnVal, st.s);
        st.sync();
        from_soa(grdMat, hq.data(), ndeNum, dimNum);
        for (size_t k = 0; k < hq.size(); ++k) grdMat[k] = -grdMat[k];   // p = -grad
        return 0;
    });
}

int lddmm_objective_create_l2(lddmm_objective **out, int dimNum, const double *iniNdeMat, const double *tgtNdeMat,
                              double disWgt, int knlOrder, double knlWidth, double timeStp, int timeNum, int ndeNum)
{
    return guarded([&] {
        require(out != nullptr, "null handle pointer");
        *out = nullptr;
        auto h = std::make_unique<lddmm_objective>();
        upload_l2_problem(h->u, dimNum, ndeNum, timeNum, knlOrder, knlWidth, timeStp, disWgt, iniNdeMat, tgtNdeMat,
                          h->st.s);
        h->obj = std::make_unique<Objective>(h->u.p, h->st.s);
        h->alp.alloc(h->u.p.varnum());
        h->grd.alloc(h->u.p.varnum());
        h->f.alloc(1);
        *out = h.release();
        return 0;
    });
}

int lddmm_run_l2(int dimNum, const double *iniNdeMat, const double *iniAlpVec, const double *tgtNdeMat,
                 double disWgt, int knlOrder, double knlWidth, double timeStp, int timeNum, int itrMax, double tolVal,
                 double wolfe1, double wolfe2, int memNum, int vbsFlg, int ndeNum, double *optAlpVec,
                 const lddmm_options *opts, lddmm_result *result)
{
    return guarded([&] {
        require(iniAlpVec && optAlpVec, "null momentum array");
        LbfgsOptions lo = make_opts(itrMax, tolVal, wolfe1, wolfe2, memNum, vbsFlg, opts);
        Stream st;
        UploadedProblem u;
        u.p.prec = precision_of(opts);
        upload_l2_problem(u, dimNum, ndeNum, timeNum, knlOrder, knlWidth, timeStp, disWgt, iniNdeMat, tgtNdeMat,
                          st.s);
        Objective obj(u.p, st.s, opts ? opts->pair_target_blocks : 0);
        auto hx = to_soa(iniAlpVec, ndeNum, dimNum, timeNum - 1);
        DevBuf<double> x(hx.size());
        x.upload(hx.data(), st.s);
        LbfgsResult lr = run_lbfgs(obj, x.p, lo, st.s);
        x.download(hx.data(), st.s);
        st.sync();
        from_soa(optAlpVec, hx.data(), ndeNum, dimNum, timeNum - 1);
        fill_result(result, lr, obj);
        return lr.status;
    });
}

// ---------------------------------------------------------------- batches
struct lddmm_batch {
    Stream st;
    std::unique_ptr<Batch> b;
};

int lddmm_batch_create(lddmm_batch **out, int dimNum, int numProblems, const int32_t *ndeNums,
                       const int32_t *dfmElmNums, const int32_t *tgtElmNums, const double *iniNdeMats,
                       const int32_t *dfmElmVtxMats, const double *tgtCenPosMats, const double *tgtElmVolVecs,
                       const double *tgtUniDirMats, const double *knlWidths, char cenKnlType, double cenKnlWidth,
                       char dirKnlType, double dirKnlWidth, double disWgt, int knlOrder, double timeStp,
                       int timeNum, int precision, int chunk)
{
    return guarded([&] {
        require(out != nullptr, "null handle pointer");
        *out = nullptr;
        require(numProblems > 0, "numProblems must be positive");
        require(ndeNums && dfmElmNums && tgtElmNums && iniNdeMats && dfmElmVtxMats && knlWidths, "null argument");
        require(precision == LDDMM_PRECISION_FP64 || precision == LDDMM_PRECISION_FP32, "unknown precision");
        BatchSpec sp;
        sp.dim = dimNum;
        sp.T = timeNum;
        sp.order = knlOrder;
        sp.prec = precision == LDDMM_PRECISION_FP32 ? kFP32 : kFP64;
        sp.h = timeStp;
        sp.w = disWgt;
        sp.vp = VfParams{cenKnlType, dirKnlType, cenKnlWidth, dirKnlWidth};
        sp.chunk = chunk;
        long long tt = 0;
        for (int p = 0; p < numProblems; ++p) {
            sp.n.push_back(ndeNums[p]);
            sp.ne.push_back(dfmElmNums[p]);
            sp.nt.push_back(tgtElmNums[p]);
            sp.width.push_back(knlWidths[p]);
            tt += tgtElmNums[p];
        }
        require(tt == 0 || (tgtCenPosMats && tgtElmVolVecs && tgtUniDirMats), "null target array");
        auto h = std::make_unique<lddmm_batch>();
        h->b = std::make_unique<Batch>(sp, iniNdeMats, dfmElmVtxMats, tgtCenPosMats, tgtElmVolVecs, tgtUniDirMats,
                                       h->st.s);
        *out = h.release();
        return 0;
    });
}

int lddmm_batch_objgrd(lddmm_batch *b, double *fcnVals, double *grdStks, const double *alpStks)
{
    return guarded([&] {
        require(b && fcnVals && grdStks && alpStks, "null argument");
        Batch &B = *b->b;
        auto ha = B.to_device_layout(alpStks);
        DevBuf<double> a(ha.size()), g(ha.size()), f(B.P());
        a.upload(ha.data(), b->st.s);
        B.eval(a.p, g.p, f.p, std::vector<int>(B.P(), B.spec().prec));
        g.download(ha.data(), b->st.s);
        f.download(fcnVals, b->st.s);
        b->st.sync();
        B.to_host_layout(ha, grdStks);
        return 0;
    });
}

// runs the batch from the device-layout start x and returns the results in
// the host layout
static void batch_run_from(lddmm_batch *b, DevBuf<double> &x, double *optAlpStks, int itrMax, double tolVal,
                           double wolfe1, double wolfe2, int memNum, const lddmm_options *opts,
                           const double *targetEnergies, lddmm_result *results, int *rounds)
{
    Batch &B = *b->b;
    BatchLbfgsOptions bo;
    bo.base = make_opts(itrMax, tolVal, wolfe1, wolfe2, memNum, 0, opts);
    if (opts) {
        bo.prec = precision_of(opts);
        bo.progress = opts->progress;
    }
    if (targetEnergies) bo.targets.assign(targetEnergies, targetEnergies + B.P());
    BatchRunResult r = B.run(x.p, bo);
    std::vector<double> hx((size_t)B.varnum_total());
    x.download(hx.data(), b->st.s);
    b->st.sync();
    B.to_host_layout(hx, optAlpStks);
    if (rounds) *rounds = r.rounds;
    if (results) {
        for (int p = 0; p < B.P(); ++p) {
            lddmm_result &q = results[p];
            std::memset(&q, 0, sizeof(q));
            q.fcn = r.prob[p].f;
            q.grad_norm = r.prob[p].gnorm;
            q.seconds = r.seconds;
            q.pairs_per_eval = B.pairs_per_eval(p);
            q.iterations = r.prob[p].iters;
            q.evaluations = r.prob[p].evals;
            q.status = r.prob[p].status;
            q.fp32_iterations = r.prob[p].fp32_iters;
            q.reached_target = r.prob[p].reached_target;
        }
    }
}

int lddmm_batch_run(lddmm_batch *b, const double *iniAlpStks, double *optAlpStks, int itrMax, double tolVal,
                    double wolfe1, double wolfe2, int memNum, const lddmm_options *opts,
                    const double *targetEnergies, lddmm_result *results, int *rounds)
{
    return guarded([&] {
        require(b && optAlpStks, "null argument");
        Batch &B = *b->b;
        DevBuf<double> x((size_t)B.varnum_total());
        if (iniAlpStks) {
            const std::vector<double> hx = B.to_device_layout(iniAlpStks);
            x.upload(hx.data(), b->st.s);
            b->st.sync();
        } else if (x.n) {
            LDDMM_HIP(hipMemsetAsync(x.p, 0, x.n * sizeof(double), b->st.s));
        }
        batch_run_from(b, x, optAlpStks, itrMax, tolVal, wolfe1, wolfe2, memNum, opts, targetEnergies, results,
                       rounds);
        return 0;
    });
}

int lddmm_batch_objective_scan(lddmm_batch *b, const double *baseVecs, const double *scales, int nScales,
                               double *fcnVals)
{
    return guarded([&] {
        require(b && baseVecs && scales && fcnVals, "null argument");
        require(nScales > 0, "nScales must be positive");
        Batch &B = *b->b;
        const int P = B.P();
        const std::vector<double> hu = B.vec_to_device_layout(baseVecs);
        std::vector<double> hs((size_t)nScales * P);
        for (int k = 0; k < nScales; ++k) std::fill(hs.begin() + (size_t)k * P, hs.begin() + (size_t)(k + 1) * P, scales[k]);
        DevBuf<double> u(hu.size()), s(hs.size()), x((size_t)B.varnum_total()), f((size_t)P),
            fall((size_t)nScales * P);
        u.upload(hu.data(), b->st.s);
        s.upload(hs.data(), b->st.s);
        const std::vector<int> prec(P, B.spec().prec);
        // one set of buffers for every multiple, so that the evaluation's
        // captured graph is replayed rather than captured per multiple
        for (int k = 0; k < nScales; ++k) {
            B.const_momentum(u.p, s.p + (size_t)k * P, x.p);
            B.eval(x.p, nullptr, f.p, prec);
            LDDMM_HIP(hipMemcpyAsync(fall.p + (size_t)k * P, f.p, (size_t)P * sizeof(double), hipMemcpyDeviceToDevice,
                                     b->st.s));
        }
        fall.download(fcnVals, b->st.s);
        b->st.sync();
        return 0;
    });
}

int lddmm_batch_run_scaled(lddmm_batch *b, const double *baseVecs, const double *startScales, double *optAlpStks,
                           int itrMax, double tolVal, double wolfe1, double wolfe2, int memNum,
                           const lddmm_options *opts, const double *targetEnergies, lddmm_result *results,
                           int *rounds)
{
    return guarded([&] {
        require(b && baseVecs && startScales && optAlpStks, "null argument");
        Batch &B = *b->b;
        const std::vector<double> hu = B.vec_to_device_layout(baseVecs);
        DevBuf<double> u(hu.size()), s((size_t)B.P()), x((size_t)B.varnum_total());
        u.upload(hu.data(), b->st.s);
        s.upload(startScales, b->st.s);
        B.const_momentum(u.p, s.p, x.p);
        b->st.sync();
        batch_run_from(b, x, optAlpStks, itrMax, tolVal, wolfe1, wolfe2, memNum, opts, targetEnergies, results,
                       rounds);
        return 0;
    });
}

int lddmm_batch_info(lddmm_batch *b, int *chunk, double *pairsPerEval)
{
    return guarded([&] {
        require(b != nullptr, "null argument");
        if (chunk) *chunk = b->b->chunk();
        if (pairsPerEval)
            for (int p = 0; p < b->b->P(); ++p) pairsPerEval[p] = b->b->pairs_per_eval(p);
        return 0;
    });
}

int lddmm_batch_tuning(lddmm_batch *b, int *chunk, int *kvLanes, int *bwdLanes)
{
    return guarded([&] {
        require(b != nullptr, "null argument");
        if (chunk) *chunk = b->b->chunk();
        if (kvLanes) *kvLanes = b->b->kv_lanes();
        if (bwdLanes) *bwdLanes = b->b->bwd_lanes();
        return 0;
    });
}

int lddmm_batch_set_tuning(lddmm_batch *b, int chunk, int kvLanes, int bwdLanes)
{
    return guarded([&] {
        require(b != nullptr, "null argument");
        b->st.sync();
        b->b->set_tuning(chunk, kvLanes, bwdLanes);
        return 0;
    });
}

int lddmm_batch_set_graphs(lddmm_batch *b, int on)
{
    return guarded([&] {
        require(b != nullptr, "null argument");
        b->st.sync();
        b->b->set_graphs(on != 0);
        return 0;
    });
}

void lddmm_batch_destroy(lddmm_batch *b) { delete b; }

// ---------------------------------------------------------------- learned warm start
int lddmm_point_features(float *featMat, int dimNum, const double *ndeMat, int ndeNum, const int32_t *elmVtxMat,
                         int elmNum, const double *tgtCenPosMat, const double *tgtUniDirMat,
                         const double *tgtElmVolVec, int tgtElmNum, double knlWidth)
{
    return guarded([&] {
        require(featMat && ndeMat && elmVtxMat, "null argument");
        require(tgtElmNum >= 0 && (tgtElmNum == 0 || (tgtCenPosMat && tgtUniDirMat && tgtElmVolVec)),
                "null target array");
        Stream st;
        point_features(dimNum, ndeMat, ndeNum, elmVtxMat, elmNum, tgtCenPosMat, tgtUniDirMat, tgtElmVolVec,
                       tgtElmNum, knlWidth, featMat, st.s);
        return 0;
    });
}

struct lddmm_mlp {
    Stream st;
    std::unique_ptr<Mlp> m;
};

int lddmm_mlp_create(lddmm_mlp **out, int hidden, int layers, uint64_t seed)
{
    return guarded([&] {
        require(out != nullptr, "null handle pointer");
        *out = nullptr;
        auto h = std::make_unique<lddmm_mlp>();
        h->m = std::make_unique<Mlp>(hidden, layers, (unsigned long long)seed, h->st.s);
        *out = h.release();
        return 0;
    });
}

int lddmm_mlp_param_count(lddmm_mlp *m, int *count)
{
    return guarded([&] {
        require(m && count, "null argument");
        *count = m->m->param_count();
        return 0;
    });
}

int lddmm_mlp_get_params(lddmm_mlp *m, float *params)
{
    return guarded([&] {
        require(m && params, "null argument");
        m->m->get_params(params);
        return 0;
    });
}

int lddmm_mlp_set_params(lddmm_mlp *m, const float *params)
{
    return guarded([&] {
        require(m && params, "null argument");
        m->m->set_params(params);
        return 0;
    });
}

int lddmm_mlp_forward(lddmm_mlp *m, const float *featMat, int rows, float *outMat)
{
    return guarded([&] {
        require(m && featMat && outMat, "null argument");
        require(rows > 0, "rows must be positive");
        DevBuf<float> x((size_t)rows * kFeat), y((size_t)rows * 3);
        x.upload(featMat, m->st.s);
        m->m->forward(x.p, rows, y.p);
        y.download(outMat, m->st.s);
        m->st.sync();
        return 0;
    });
}

int lddmm_mlp_train(lddmm_mlp *m, const float *featMat, const float *tgtMat, int rows, int steps, int batchRows,
                    float lr, double *losses, float *grad)
{
    return guarded([&] {
        require(m && featMat && tgtMat, "null argument");
        require(rows > 0 && steps >= 0 && batchRows > 0, "rows and batchRows must be positive");
        require(!grad || (steps == 1 && batchRows >= rows), "grad needs one full-batch step");
        DevBuf<float> x((size_t)rows * kFeat), y((size_t)rows * 3);
        x.upload(featMat, m->st.s);
        y.upload(tgtMat, m->st.s);
        int start = 0;
        for (int s = 0; s < steps; ++s) {
            const int cnt = std::min(batchRows, rows - start);
            const double l = m->m->train_step(x.p + (size_t)start * kFeat, y.p + (size_t)start * 3, cnt, lr, grad);
            if (losses) losses[s] = l;
            start += cnt;
            if (start >= rows) start = 0;
        }
        m->st.sync();
        return 0;
    });
}

void lddmm_mlp_destroy(lddmm_mlp *m) { delete m; }

int lddmm_gemm_bf16(float *C, const uint16_t *A, const uint16_t *B, int M, int N, int K, int transA, int transB,
                    int splitK)
{
    return guarded([&] {
        require(C && A && B, "null argument");
        require(M > 0 && N > 0 && K > 0 && M % 32 == 0 && N % 32 == 0 && K % 16 == 0,
                "M, N must be positive multiples of 32 and K of 16");
        require(splitK >= 1, "splitK must be at least 1");
        Stream st;
        DevBuf<uint16_t> a((size_t)M * K), b((size_t)K * N);
        DevBuf<float> c((size_t)M * N), part((size_t)M * N * splitK);
        a.upload(A, st.s);
        b.upload(B, st.s);
        gemm_bf16(transA != 0, transB != 0, (const __bf16 *)a.p, transA ? M : K, (const __bf16 *)b.p,
                  transB ? K : N, c.p, N, M, N, K, splitK, part.p, st.s);
        c.download(C, st.s);
        st.sync();
        return 0;
    });
}

// ---------------------------------------------------------------- operators
int lddmm_multiply_kernel_cross(double *vlcMat, const double *lmkiMat, const double *lmkjMat, const double *alpMat,
                                int knlOrder, double knlWidth, int lmkiNum, int lmkjNum, int dimNum)
{
    return guarded([&] {
        require(vlcMat && lmkiMat && lmkjMat && alpMat, "null argument");
        require(dimNum == 2 || dimNum == 3, "dimNum must be 2 or 3");
        require(lmkiNum >= 0 && lmkjNum >= 0, "negative size");
        require(knlWidth > 0.0, "knlWidth must be positive");
        if (lmkiNum == 0) return 0;
        if (lmkjNum == 0) {
            std::memset(vlcMat, 0, sizeof(double) * dimNum * lmkiNum);
            return 0;
        }
        Stream st;
        auto hqi = to_soa(lmkiMat, lmkiNum, dimNum), hqj = to_soa(lmkjMat, lmkjNum, dimNum);
        auto ha = to_soa(alpMat, lmkjNum, dimNum);
        DevBuf<double> qi(hqi.size()), qj(hqj.size()), a(ha.size()), v(hqi.size());
        qi.upload(hqi.data(), st.s);
        qj.upload(hqj.data(), st.s);
        a.upload(ha.data(), st.s);
        PairSplit sp = plan_kv(lmkiNum, lmkjNum, kFP64, 2048);
        DevBuf<double> part((size_t)sp.splits * dimNum * lmkiNum);
        {
            ProfScope ps(kProfOperator, st.s);
            launch_kv(dimNum, knlOrder, kFP64, qi.p, lmkiNum, qj.p, a.p, lmkjNum, knlWidth, sp, part.p, st.s);
            launch_sum_splits(dimNum, sp.splits, lmkiNum, part.p, 1.0, v.p, st.s);
        }
        v.download(hqi.data(), st.s);
        st.sync();
        from_soa(vlcMat, hqi.data(), lmkiNum, dimNum);
        return 0;
    });
}

int lddmm_multiply_kernel_cross_noalloc(double *d_vlcMat, const double *d_lmkiMat, const double *d_lmkjMat,
                                        const double *d_alpMat, int knlOrder, double knlWidth, int lmkiNum,
                                        int lmkjNum, int dimNum, void *stream)
{
    return guarded([&] {
        require(d_vlcMat && d_lmkiMat && d_lmkjMat && d_alpMat, "null argument");
        require(dimNum == 2 || dimNum == 3, "dimNum must be 2 or 3");
        require(lmkiNum >= 0 && lmkjNum >= 0, "negative size");
        require(knlWidth > 0.0, "knlWidth must be positive");
        if (lmkiNum == 0) return 0;
        hipStream_t st = static_cast<hipStream_t>(stream);
        double *scr = noalloc_scratch();
        if (lmkjNum == 0) {   // no sources: zero splits sum to zeros
            launch_sum_splits(dimNum, 0, lmkiNum, scr, 1.0, d_vlcMat, st);
            return 0;
        }
        // the plan of lddmm_multiply_kernel_cross: the same kernels, the same sums
        const PairSplit sp = plan_kv(lmkiNum, lmkjNum, kFP64, 2048);
        const size_t need = (size_t)sp.splits * dimNum * lmkiNum;
        if (need > kNoAllocScratchDoubles)
            throw Error(kErrArg, "lddmm_multiply_kernel_cross_noalloc: these sizes need " + std::to_string(need) +
                                     " doubles of scratch, more than its " +
                                     std::to_string(kNoAllocScratchDoubles));
        launch_kv(dimNum, knlOrder, kFP64, d_lmkiMat, lmkiNum, d_lmkjMat, d_alpMat, lmkjNum, knlWidth, sp, scr, st);
        launch_sum_splits(dimNum, sp.splits, lmkiNum, scr, 1.0, d_vlcMat, st);
        return 0;
    });
}

int lddmm_multiply_kernel(double *vlcMat, const double *lmkMat, const double *alpMat, int knlOrder, double knlWidth,
                          int lmkNum, int dimNum)
{
    return lddmm_multiply_kernel_cross(vlcMat, lmkMat, lmkMat, alpMat, knlOrder, knlWidth, lmkNum, lmkNum, dimNum);
}

int lddmm_dq_kernel(double *dqKMat, const double *lmkMat, const double *lftMat, const double *rgtMat, int knlOrder,
                    double knlWidth, int lmkNum, int dimNum)
{
    return guarded([&] {
        require(dqKMat && lmkMat && lftMat && rgtMat, "null argument");
        require(dimNum == 2 || dimNum == 3, "dimNum must be 2 or 3");
        require(lmkNum >= 0, "negative size");
        require(knlWidth > 0.0, "knlWidth must be positive");
        if (lmkNum == 0) return 0;
        Stream st;
        auto hq = to_soa(lmkMat, lmkNum, dimNum), hl = to_soa(lftMat, lmkNum, dimNum),
             hr = to_soa(rgtMat, lmkNum, dimNum);
        DevBuf<double> q(hq.size()), l(hl.size()), r(hr.size()), out(hq.size());
        q.upload(hq.data(), st.s);
        l.upload(hl.data(), st.s);
        r.upload(hr.data(), st.s);
        PairSplit sp = plan_single(lmkNum, lmkNum, 2048);
        DevBuf<double> part((size_t)sp.splits * dimNum * lmkNum);
        {
            ProfScope ps(kProfOperator, st.s);
            launch_dq_self(dimNum, knlOrder, kFP64, q.p, l.p, r.p, lmkNum, knlWidth, sp, part.p, st.s);
            launch_sum_splits(dimNum, sp.splits, lmkNum, part.p, 1.0, out.p, st.s);
        }
        out.download(hq.data(), st.s);
        st.sync();
        from_soa(dqKMat, hq.data(), lmkNum, dimNum);
        return 0;
    });
}

int lddmm_dq_kernel_cross(double *dqiKMat, double *dqjKMat, const double *lmkiMat, const double *lmkjMat,
                          const double *lftMat, const double *rgtMat, int knlOrder, double knlWidth, int lmkiNum,
                          int lmkjNum, int dimNum)
{
    return guarded([&] {
        require(dqiKMat && dqjKMat && lmkiMat && lmkjMat && lftMat && rgtMat, "null argument");
        require(dimNum == 2 || dimNum == 3, "dimNum must be 2 or 3");
        require(lmkiNum >= 0 && lmkjNum >= 0, "negative size");
        require(knlWidth > 0.0, "knlWidth must be positive");
        if (lmkiNum == 0 || lmkjNum == 0) {
            std::memset(dqiKMat, 0, sizeof(double) * dimNum * lmkiNum);
            std::memset(dqjKMat, 0, sizeof(double) * dimNum * lmkjNum);
            return 0;
        }
        Stream st;
        auto hqi = to_soa(lmkiMat, lmkiNum, dimNum), hqj = to_soa(lmkjMat, lmkjNum, dimNum);
        auto hl = to_soa(lftMat, lmkiNum, dimNum), hr = to_soa(rgtMat, lmkjNum, dimNum);
        DevBuf<double> qi(hqi.size()), qj(hqj.size()), l(hl.size()), r(hr.size());
        DevBuf<double> oi(hqi.size()), oj(hqj.size());
        qi.upload(hqi.data(), st.s);
        qj.upload(hqj.data(), st.s);
        l.upload(hl.data(), st.s);
        r.upload(hr.data(), st.s);
        PairSplit spi = plan_single(lmkiNum, lmkjNum, 2048), spj = plan_single(lmkjNum, lmkiNum, 2048);
        DevBuf<double> part(std::max((size_t)spi.splits * dimNum * lmkiNum, (size_t)spj.splits * dimNum * lmkjNum));
        {
            ProfScope ps(kProfOperator, st.s);
            launch_dq_cross(dimNum, knlOrder, kFP64, qi.p, l.p, lmkiNum, qj.p, r.p, lmkjNum, knlWidth, spi, part.p, st.s);
            launch_sum_splits(dimNum, spi.splits, lmkiNum, part.p, 1.0, oi.p, st.s);
            launch_dq_cross(dimNum, knlOrder, kFP64, qj.p, r.p, lmkjNum, qi.p, l.p, lmkiNum, knlWidth, spj, part.p, st.s);
            launch_sum_splits(dimNum, spj.splits, lmkjNum, part.p, 1.0, oj.p, st.s);
        }
        oi.download(hqi.data(), st.s);
        oj.download(hqj.data(), st.s);
        st.sync();
        from_soa(dqiKMat, hqi.data(), lmkiNum, dimNum);
        from_soa(dqjKMat, hqj.data(), lmkjNum, dimNum);
        return 0;
    });
}

int lddmm_compute_kernel_cross(double *knlMat, const double *lmkiMat, const double *lmkjMat, int knlOrder,
                               double knlWidth, int lmkiNum, int lmkjNum, int dimNum)
{
    return guarded([&] {
        require(knlMat && lmkiMat && lmkjMat, "null argument");
        require(dimNum == 2 || dimNum == 3, "dimNum must be 2 or 3");
        require(lmkiNum >= 0 && lmkjNum >= 0, "negative size");
        require(knlWidth > 0.0, "knlWidth must be positive");
        if (lmkiNum == 0 || lmkjNum == 0) return 0;
        Stream st;
        auto hqi = to_soa(lmkiMat, lmkiNum, dimNum), hqj = to_soa(lmkjMat, lmkjNum, dimNum);
        DevBuf<double> qi(hqi.size()), qj(hqj.size()), K((size_t)lmkiNum * lmkjNum);
        qi.upload(hqi.data(), st.s);
        qj.upload(hqj.data(), st.s);
        {
            ProfScope ps(kProfOperator, st.s);
            launch_dense_kernel(dimNum, knlOrder, qi.p, lmkiNum, qj.p, lmkjNum, knlWidth, K.p, st.s);
        }
        K.download(knlMat, st.s);
        st.sync();
        return 0;
    });
}

int lddmm_compute_kernel(double *knlMat, const double *lmkMat, int knlOrder, double knlWidth, int lmkNum,
                         int dimNum)
{
    // symmetric: the cross layout K[j*N + i] equals the self layout K[i*N + j]
    return lddmm_compute_kernel_cross(knlMat, lmkMat, lmkMat, knlOrder, knlWidth, lmkNum, lmkNum, dimNum);
}

int lddmm_varifold(double *vfdVal, double *dqVfdMat, const double *dfmLmkPosMat, const int32_t *dfmElmVtxMat,
                   const int32_t *dfmElmIfoMat, const double *tgtCenPosMat, const double *tgtUniDirMat,
                   const double *tgtElmVolVec, char cenKnlType, double cenKnlWidth, char dirKnlType,
                   double dirKnlWidth, int dfmLmkNum, int dfmElmNum, int dfmElmIfoNum, int tgtElmNum, int dimNum)
{
    return guarded([&] {
        require(vfdVal != nullptr, "null argument");
        check_common(dimNum, dfmLmkNum, dfmElmNum, dfmElmIfoNum, tgtElmNum, 2);
        Stream st;
        UploadedProblem u;
        fill_params(u.p, dimNum, dfmLmkNum, 2, -1, 1.0, 0.0, 1.0, cenKnlType, cenKnlWidth, dirKnlType, dirKnlWidth,
                    dfmElmNum, tgtElmNum, dfmElmIfoNum);
        validate(u.p);
        upload_host_problem(u, dimNum, dfmLmkPosMat, dfmElmVtxMat, dfmElmIfoMat, tgtCenPosMat, tgtElmVolVec,
                            tgtUniDirMat, dfmLmkNum, dfmElmNum, dfmElmIfoNum, tgtElmNum, st.s);
        const int D = dimNum, ne = dfmElmNum, nt = tgtElmNum, nq = dfmLmkNum;
        PairSplit spd = plan_vf(ne, ne, 1024), spt = plan_vf(ne, std::max(nt, 1), 1024);
        PairSplit sptt = plan_vf(std::max(nt, 1), std::max(nt, 1), 2048);
        DevBuf<double> cen((size_t)D * ne), dir((size_t)D * ne), vol(ne), dc((size_t)D * ne), dd((size_t)D * ne);
        DevBuf<double> part(std::max((size_t)(spd.splits + spt.splits) * vf_comps(D) * ne,
                                     (size_t)sptt.splits * std::max(nt, 1)));
        DevBuf<double> ep(std::max(vf_finalize_blocks(ne), vf_value_finalize_blocks(std::max(nt, 1))));
        DevBuf<double> scratch(kDotBlocks), vals(2), dq((size_t)D * nq);
        {
            ProfScope ps(kProfOperator, st.s);
            launch_elm_geometry(D, u.p.q0, nq, u.p.elm, ne, cen.p, dir.p, vol.p, st.s);
            launch_vf_pairs(D, kFP64, u.p.vp, cen.p, dir.p, vol.p, ne, u.p.tc, u.p.tu, u.p.ta, nt, spd, spt, part.p, st.s);
            launch_vf_finalize(D, u.p.vp, spd.splits, spd.splits + (nt > 0 ? spt.splits : 0), ne, part.p, dir.p, vol.p, dc.p,
                               dd.p, ep.p, st.s);
            launch_sum(ep.p, vf_finalize_blocks(ne), scratch.p, vals.p, st.s);
            if (dqVfdMat) launch_vf_gather(D, u.p.q0, nq, u.p.ifo, dfmElmIfoNum, dc.p, dd.p, ne, 1.0, dq.p, st.s);
            if (nt > 0) {
                launch_vf_value_pairs(D, u.p.vp, u.p.tc, u.p.tu, u.p.ta, nt, sptt, part.p, st.s);
                launch_vf_value_finalize(sptt.splits, nt, part.p, u.p.ta, ep.p, st.s);
                launch_sum(ep.p, vf_value_finalize_blocks(nt), scratch.p, vals.p + 1, st.s);
            } else {
                LDDMM_HIP(hipMemsetAsync(vals.p + 1, 0, sizeof(double), st.s));
            }
        }
        double hv[2];
        vals.download(hv, st.s);
        std::vector<double> hdq;
        if (dqVfdMat) {
            hdq.resize((size_t)D * nq);
            dq.download(hdq.data(), st.s);
        }
        st.sync();
        *vfdVal = hv[0] + hv[1];
        if (dqVfdMat) from_soa(dqVfdMat, hdq.data(), nq, D);
        return 0;
    });
}

// ---------------------------------------------------------------- data formats
int lddmm_spatial_order(int32_t *perm, const double *ptsMat, int ptsNum, int dimNum)
{
    return guarded([&] {
        require(ptsNum >= 0, "negative size");
        require(ptsNum == 0 || (perm && ptsMat), "null argument");
        require(dimNum == 2 || dimNum == 3, "dimNum must be 2 or 3");
        const std::vector<int> v =
            spatial_order(ptsNum, dimNum, [&](int i, int c) { return ptsMat[(size_t)i * dimNum + c]; });
        std::copy(v.begin(), v.end(), perm);
        return 0;
    });
}

int lddmm_set_pair_culling(int on)
{
    set_pair_culling(on != 0);
    return 0;
}

int lddmm_set_pair_cull_level(int bits)
{
    return guarded([&] {
        set_pair_cull_bits(bits);
        return 0;
    });
}

int lddmm_set_pair_groups(int on)
{
    set_pair_groups(on != 0);
    return 0;
}

int lddmm_set_pair_expansion(int on)
{
    set_pair_expansion(on != 0);
    return 0;
}

int lddmm_measure_rates(double *rates)
{
    return guarded([&] {
        require(rates != nullptr, "null argument");
        Stream st;
        measure_rates(rates, st.s);
        return 0;
    });
}

int lddmm_pair_count_begin(void)
{
    return guarded([&] {
        require(pair_counter() == nullptr, "a pair count is already running");
        unsigned long long *c = nullptr;
        LDDMM_HIP(hipMalloc((void **)&c, 3 * sizeof(unsigned long long)));
        const hipError_t e = hipMemset(c, 0, 3 * sizeof(unsigned long long));
        if (e != hipSuccess) {
            (void)hipFree(c);
            LDDMM_HIP(e);
        }
        set_pair_counter(c);
        return 0;
    });
}

int lddmm_pair_count_end(unsigned long long *executedPairs)
{
    return guarded([&] {
        unsigned long long *c = pair_counter();
        require(c != nullptr, "no pair count is running");
        set_pair_counter(nullptr);
        unsigned long long v[3] = {0, 0, 0};
        hipError_t e = hipDeviceSynchronize();
        if (e == hipSuccess) e = hipMemcpy(v, c, sizeof(v), hipMemcpyDeviceToHost);
        (void)hipFree(c);
        LDDMM_HIP(e);
        if (executedPairs)
            for (int k = 0; k < 3; ++k) executedPairs[k] = v[k];
        return 0;
    });
}

int lddmm_get_elm_ifo(int32_t *elmIfoMat, const int32_t *elmVtxMat, int vtxPerElm, int elmNum, int ndeNum)
{
    return guarded([&] {
        require(elmVtxMat != nullptr, "null element table");
        require(vtxPerElm == 2 || vtxPerElm == 3, "vtxPerElm must be 2 or 3");
        require(elmNum >= 0 && ndeNum >= 0, "negative size");
        std::vector<int> cnt(ndeNum, 0);
        for (long long k = 0; k < (long long)elmNum * vtxPerElm; ++k) {
            int v = elmVtxMat[k];
            require(v >= 0 && v < ndeNum, "element table holds a vertex outside [0, ndeNum)");
            ++cnt[v];
        }
        int max_adj = 0;
        for (int c : cnt) max_adj = std::max(max_adj, c);
        const int rows = 1 + max_adj * vtxPerElm;
        if (!elmIfoMat) return rows;
        std::memset(elmIfoMat, 0, sizeof(int32_t) * (size_t)rows * ndeNum);
        // elements in ascending order, as MATLAB's find() returns them
        for (int e = 0; e < elmNum; ++e) {
            const int32_t *ev = elmVtxMat + (size_t)e * vtxPerElm;
            for (int k = 0; k < vtxPerElm; ++k) {
                int32_t *col = elmIfoMat + (size_t)ev[k] * rows;
                int32_t *slot = col + 1 + (size_t)col[0] * vtxPerElm;
                ++col[0];
                slot[0] = e;
                if (vtxPerElm == 2) {
                    slot[1] = k == 0 ? -1 : 1;
                } else {
                    // (qa, qb) so that dd x (qa - qb) is the normal's derivative
                    static const int pa[3] = {2, 0, 1}, pb[3] = {1, 2, 0};
                    slot[1] = ev[pa[k]];
                    slot[2] = ev[pb[k]];
                }
            }
        }
        return rows;
    });
}

int lddmm_element_varifold(double *cenPosMat, double *uniDirMat, double *elmVolVec, const double *ndeMat,
                           const int32_t *elmVtxMat, int ndeNum, int elmNum, int dimNum)
{
    return guarded([&] {
        require(cenPosMat && uniDirMat && elmVolVec && ndeMat && elmVtxMat, "null argument");
        require(dimNum == 2 || dimNum == 3, "dimNum must be 2 or 3");
        require(ndeNum > 0 && elmNum > 0, "ndeNum and elmNum must be positive");
        for (long long k = 0; k < (long long)elmNum * dimNum; ++k)
            require(elmVtxMat[k] >= 0 && elmVtxMat[k] < ndeNum, "element table holds a vertex outside [0, ndeNum)");
        Stream st;
        auto hq = to_soa(ndeMat, ndeNum, dimNum);
        auto he = to_soa(elmVtxMat, elmNum, dimNum);
        DevBuf<double> q(hq.size()), cen((size_t)dimNum * elmNum), dir((size_t)dimNum * elmNum), vol(elmNum);
        DevBuf<int> e(he.size());
        q.upload(hq.data(), st.s);
        e.upload(he.data(), st.s);
        launch_elm_geometry(dimNum, q.p, ndeNum, e.p, elmNum, cen.p, dir.p, vol.p, st.s);
        std::vector<double> hc((size_t)dimNum * elmNum), hd((size_t)dimNum * elmNum);
        cen.download(hc.data(), st.s);
        dir.download(hd.data(), st.s);
        vol.download(elmVolVec, st.s);
        st.sync();
        from_soa(cenPosMat, hc.data(), elmNum, dimNum);
        from_soa(uniDirMat, hd.data(), elmNum, dimNum);
        return 0;
    });
}

int lddmm_compute_node_position(double *ndeStk, double *vlcStk, const double *iniNdeMat, const double *alpStk,
                                int knlOrder, double knlWidth, double timeStp, int timeNum, int ndeNum, int dimNum)
{
    return guarded([&] {
        require(iniNdeMat && alpStk, "null argument");
        require(dimNum == 2 || dimNum == 3, "dimNum must be 2 or 3");
        require(ndeNum > 0 && timeNum >= 1, "bad sizes");
        require(knlWidth > 0.0, "knlWidth must be positive");
        Stream st;
        const int D = dimNum, n = ndeNum, T = timeNum;
        const size_t dn = (size_t)D * n;
        auto hq = to_soa(iniNdeMat, n, D);
        auto ha = to_soa(alpStk, n, D, std::max(T - 1, 0));
        DevBuf<double> q(dn * T), a(std::max(ha.size(), (size_t)1)), v(std::max(dn * (T - 1), (size_t)1));
        LDDMM_HIP(hipMemcpyAsync(q.p, hq.data(), dn * sizeof(double), hipMemcpyHostToDevice, st.s));
        if (!ha.empty()) LDDMM_HIP(hipMemcpyAsync(a.p, ha.data(), ha.size() * sizeof(double), hipMemcpyHostToDevice, st.s));
        PairSplit sp = plan_kv(n, n, kFP64, 2048);
        DevBuf<double> part((size_t)sp.splits * D * n), ep(fwd_reduce_blocks(n));
        for (int t = 0; t < T - 1; ++t) {
            ProfScope ps(kProfOperator, st.s);
            const double *qt = q.p + (size_t)t * dn, *at = a.p + (size_t)t * dn;
            launch_kv(D, knlOrder, kFP64, qt, n, qt, at, n, knlWidth, sp, part.p, st.s);
            launch_fwd_reduce(D, sp.splits, n, part.p, qt, at, timeStp, q.p + (size_t)(t + 1) * dn,
                              v.p + (size_t)t * dn, ep.p, st.s);
        }
        std::vector<double> hn(dn * T), hv(dn * std::max(T - 1, 0));
        q.download(hn.data(), st.s);
        if (!hv.empty()) LDDMM_HIP(hipMemcpyAsync(hv.data(), v.p, hv.size() * sizeof(double), hipMemcpyDeviceToHost, st.s));
        st.sync();
        if (ndeStk) from_soa(ndeStk, hn.data(), n, D, T);
        if (vlcStk && T > 1) from_soa(vlcStk, hv.data(), n, D, T - 1);
        return 0;
    });
}

// ---------------------------------------------------------------- utilities
int lddmm_device_count(int *count)
{
    return guarded([&] {
        require(count != nullptr, "null argument");
        int c = 0;
        hipError_t e = hipGetDeviceCount(&c);
        *count = e == hipSuccess ? c : 0;
        return 0;
    });
}

int lddmm_device_malloc(void **ptr, size_t bytes)
{
    return guarded([&] {
        require(ptr != nullptr, "null argument");
        if (hipMalloc(ptr, bytes) != hipSuccess) throw Error(kErrAlloc, "hipMalloc failed");
        return 0;
    });
}

int lddmm_device_free(void *ptr)
{
    return guarded([&] {
        LDDMM_HIP(hipFree(ptr));
        return 0;
    });
}

int lddmm_memcpy_htod(void *dst, const void *src, size_t bytes)
{
    return guarded([&] {
        LDDMM_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
        return 0;
    });
}

int lddmm_memcpy_dtoh(void *dst, const void *src, size_t bytes)
{
    return guarded([&] {
        LDDMM_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        return 0;
    });
}

int lddmm_device_synchronize(void)
{
    return guarded([&] {
        LDDMM_HIP(hipDeviceSynchronize());
        return 0;
    });
}

int lddmm_profile_enable(int on)
{
    return guarded([&] {
        Profiler::get().enable(on != 0);
        return 0;
    });
}

int lddmm_profile_reset(void)
{
    return guarded([&] {
        Profiler::get().reset();
        return 0;
    });
}

int lddmm_profile_read(const char **names, double *ms, long long *counts, int maxEntries)
{
    return guarded([&] {
        double m[kProfCount];
        long long c[kProfCount];
        Profiler::get().read(m, c);
        const int k = std::min(maxEntries, (int)kProfCount);
        for (int i = 0; i < k; ++i) {
            if (names) names[i] = prof_name(i);
            if (ms) ms[i] = m[i];
            if (counts) counts[i] = c[i];
        }
        return k;
    });
}

}  // extern "C"
