// batch.hip -- batched evaluation and L-BFGS for many uneven registrations.
//
// Layout: every per-problem array is a contiguous block in a concatenated
// device array, SoA inside the block: points [D][N_p] at D*vo_p, the flow
// [T][D][N_p] at T*D*vo_p, momenta and gradients [T-1][D][N_p] at
// (T-1)*D*vo_p, elements [V][E_p] at V*eo_p (vertex indices local to the
// problem), adjacency [rows][N_p] at rows*vo_p, targets [D][Et_p] at
// D*to_p.  Work tables list (problem, tile, chunk) items of equal size
// (64 targets x `chunk` sources), so a problem of 20,000 points contributes
// many items and one of 500 points a few, and one launch per Euler step
// keeps the whole chip busy.  All sums are in a fixed order: results do not
// depend on which other problems share the launch.
#include <algorithm>
#include <array>
#include <tuple>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <atomic>
#include <exception>
#include <thread>
#include <tuple>

#include <unistd.h>

#include "batch.hpp"
#include "pair_ops.hpp"
#include "varifold_ops.hpp"

namespace lddmm {

namespace {

// fn(p) for every problem p < P on up to 16 host threads (the set-up loops
// write disjoint per-problem ranges); the first exception is rethrown
template <class Fn>
void for_each_problem(int P, Fn &&fn)
{
    const int hw = (int)std::thread::hardware_concurrency();
    const int nt = std::max(1, std::min({P, 16, hw > 0 ? hw : 1}));
    if (nt == 1) {
        for (int p = 0; p < P; ++p) fn(p);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::exception_ptr> err(nt);
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t) {
        try {
            pool.emplace_back([&, t] {
                try {
                    for (int p = next.fetch_add(1); p < P; p = next.fetch_add(1)) fn(p);
                } catch (...) {
                    err[t] = std::current_exception();
                }
            });
        } catch (...) {
            break;   // no more threads: the started ones take every problem
        }
    }
    if (pool.empty())
        for (int p = next.fetch_add(1); p < P; p = next.fetch_add(1)) fn(p);
    for (std::thread &th : pool) th.join();
    for (std::exception_ptr &e : err)
        if (e) std::rethrow_exception(e);
}

using namespace ops;

constexpr int kTile = 64;          // targets per work item (one per lane)
constexpr int kVecBlk = 4096;      // elements per block of the batched vector kernels
constexpr int kMaxBDots = 16;   // dot products per batched launch

struct BatchView {
    const BatchDesc *desc;
    double *nde, *padj, *part, *vpart;
    const double *alp;
    double *grd;
    const double *q0, *tc, *tu, *ta;
    const int *elm, *ifo;
    int rows, T;
    double h, w;
    double *ecen, *edir, *evol, *dc, *dd, *efwd, *evf;
    int cull;     // culling level of FP32 Gaussian tile pairs (values < 2^-cull skipped; 0: off)
    int expand;   // expanded form of the FP32 Gaussian pair terms
    unsigned long long *exec;   // executed-pair counter (pair_counter), or null
    // group bounds of q_t, of the deformed element centres and of the targets
    // (BatchDesc gvo, geo, gto) for the culling operators; used when gpts /
    // gvf (point kernels / varifold) are set
    double4 *grpv, *grpe, *grpt;
    int gpts, gvf;
    // per group of q_t (one time slice, BatchDesc gvo / T): the largest
    // |alpha_t| and |alpha_t - p| components, the column scales of the
    // matrix-core adjoint (kb_group_maxima)
    float2 *grpm = nullptr;
};

// x[i] = scale * u_i with u_i a fixed pseudo-random value in [-1/2, 1/2):
// stand-in momenta and adjoints for timing launches
__global__ __launch_bounds__(kTile) void kb_fill_pattern(double *__restrict__ x, long long n, double scale)
{
    for (long long i = (long long)blockIdx.x * kTile + threadIdx.x; i < n; i += (long long)gridDim.x * kTile) {
        const unsigned h = (unsigned)(i * 2654435761ULL) ^ (unsigned)(i >> 16);
        x[i] = scale * ((double)(h >> 8) * (1.0 / 16777216.0) - 0.5);
    }
}

// x = the momentum stack of every problem p with each time slice s[p] u_p
// (u: D x N_p blocks in the device layout); blockIdx.y is the problem
__global__ __launch_bounds__(kTile) void kb_const_momentum(const BatchDesc *__restrict__ desc, int D, int T1,
                                                           const double *__restrict__ u, const double *__restrict__ s,
                                                           double *__restrict__ x)
{
    const BatchDesc d = desc[blockIdx.y];
    const long long m = (long long)D * d.n;
    const double c = s[blockIdx.y];
    const double *up = u + (long long)D * d.vo;
    double *xp = x + (long long)D * T1 * d.vo;
    for (long long i = (long long)blockIdx.x * kTile + threadIdx.x; i < m * T1; i += (long long)gridDim.x * kTile)
        xp[i] = c * up[i % m];
}

// bounds of the aligned 64-point groups: q_t (which = 0), the deformed
// element centres (1) or the target centres (2), one group per item
template <int D>
__global__ __launch_bounds__(kTile) void kb_group_bounds(BatchView v, int t, int which, const int4 *__restrict__ items)
{
    const int4 it = items[blockIdx.x];
    const BatchDesc d = v.desc[it.x];
    if (which == 0) {
        const double *q = v.nde + (size_t)v.T * D * d.vo + (size_t)t * D * d.n;
        group_bounds_wave<D>(q, d.n, it.y, threadIdx.x, v.grpv + d.gvo + (size_t)t * ((d.n + 63) / 64));
    } else if (which == 1) {
        group_bounds_wave<D>(v.ecen + (size_t)D * d.eo, d.ne, it.y, threadIdx.x, v.grpe + d.geo);
    } else {
        group_bounds_wave<D>(v.tc + (size_t)D * d.to, d.nt, it.y, threadIdx.x, v.grpt + d.gto);
    }
}

// per aligned 64-point group of alpha_t and the current adjoint p: the
// largest |alpha| and |alpha - p| components (FP32, rounded as the
// matrix-core adjoint stages them), one group per item
template <int D>
__global__ __launch_bounds__(kTile) void kb_group_maxima(BatchView v, int t, const int4 *__restrict__ items)
{
    const int4 it = items[blockIdx.x];
    const BatchDesc d = v.desc[it.x];
    const int i = it.y * 64 + threadIdx.x;
    const double *a = v.alp + (size_t)(v.T - 1) * D * d.vo + (size_t)t * D * d.n;
    const double *p = v.padj + (size_t)D * d.vo;
    float am = 0.0f, wm = 0.0f;
    if (i < d.n) {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            const double ac = a[(size_t)c * d.n + i], pc = p[(size_t)c * d.n + i];
            am = fmaxf(am, fabsf((float)ac));
            wm = fmaxf(wm, fabsf((float)(ac - pc)));
        }
    }
    am = wave_max(am);
    wm = wave_max(wm);
    if (threadIdx.x == 0) v.grpm[d.gvo / v.T + it.y] = make_float2(am, wm);
}

template <int NC>
__device__ __forceinline__ void sum_splits(const double *__restrict__ part, int splits, int n, int i,
                                           double (&v)[NC])
{
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] = 0.0;
    int k = 0;
    for (; k + 4 <= splits; k += 4) {
        double t[4][NC];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int c = 0; c < NC; ++c) t[u][c] = part[((size_t)(k + u) * NC + c) * n + i];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int c = 0; c < NC; ++c) v[c] += t[u][c];
    }
    for (; k < splits; ++k)
#pragma unroll
        for (int c = 0; c < NC; ++c) v[c] += part[((size_t)k * NC + c) * n + i];
}

// ------------------------------------------------------------- flow kernels
template <int D, int ORDER, class P, int R>
__global__ __launch_bounds__(kBlock) void kb_kv(BatchView v, int t, const int4 *__restrict__ items)
{
    using Op = OpKV<D, ORDER, P, R>;
    using real = typename Op::real;
    __shared__ __align__(16) char smem[PairSmem<Op>::bytes];
    const int4 it = items[xcd_remap(blockIdx.x, gridDim.x)];
    const BatchDesc d = v.desc[it.x];
    const double *q = v.nde + (size_t)v.T * D * d.vo + (size_t)t * D * d.n;
    const double *a = v.alp + (size_t)(v.T - 1) * D * d.vo + (size_t)t * D * d.n;
    typename Op::Args A{q,
                        q,
                        a,
                        d.n,
                        d.n,
                        (real)d.iw,
                        (real)d.iw2,
                        cull_r2(v.cull),
                        expand_r2(v.expand != 0),
                        Op::KF::kappa(d.iw, d.iw2),
                        v.exec,
                        v.gpts ? v.grpv + d.gvo + (size_t)t * ((d.n + 63) / 64) : nullptr};
    const int j0 = it.z * d.chunk, j1 = min(d.n, j0 + d.chunk);
    pairs_block<Op>(A, d.n, it.y, j0, j1, v.part + d.po + (size_t)it.z * D * d.n, smem);
}

template <int D, int ORDER, class P, int R>
__global__ __launch_bounds__(kBlock) void kb_bwd(BatchView v, int t, const int4 *__restrict__ items)
{
    using Op = OpBwd<D, ORDER, P, R>;
    using real = typename Op::real;
    __shared__ __align__(16) char smem[PairSmem<Op>::bytes];
    const int4 it = items[xcd_remap(blockIdx.x, gridDim.x)];
    const BatchDesc d = v.desc[it.x];
    const double *q = v.nde + (size_t)v.T * D * d.vo + (size_t)t * D * d.n;
    const double *a = v.alp + (size_t)(v.T - 1) * D * d.vo + (size_t)t * D * d.n;
    const double kap = Op::KF::kappa(d.iw, d.iw2);
    typename Op::Args A{q,
                        a,
                        v.padj + (size_t)D * d.vo,
                        d.n,
                        (real)d.iw,
                        (real)d.iw2,
                        cull_r2(v.cull),
                        expand_r2(v.expand != 0),
                        kap,
                        Op::KF::cmul(d.iw, d.iw2) / kap,
                        v.exec,
                        v.gpts ? v.grpv + d.gvo + (size_t)t * ((d.n + 63) / 64) : nullptr};
    const int j0 = it.z * d.chunk, j1 = min(d.n, j0 + d.chunk);
    pairs_block<Op>(A, d.n, it.y, j0, j1, v.part + d.po + (size_t)it.z * 2 * D * d.n, smem);
}

#include "mx_bwd.hpp"

template <int D>
__global__ __launch_bounds__(kTile) void kb_fwd_reduce(BatchView v, int t, const int4 *__restrict__ items)
{
    const int4 it = items[blockIdx.x];
    const BatchDesc d = v.desc[it.x];
    const int i = it.y * kTile + threadIdx.x;
    if (i >= d.n) return;
    double vv[D];
    sum_splits<D>(v.part + d.po, d.splits, d.n, i, vv);
    const double *q = v.nde + (size_t)v.T * D * d.vo + (size_t)t * D * d.n;
    double *qn = v.nde + (size_t)v.T * D * d.vo + (size_t)(t + 1) * D * d.n;
    const double *a = v.alp + (size_t)(v.T - 1) * D * d.vo + (size_t)t * D * d.n;
    double e = 0.0;
#pragma unroll
    for (int c = 0; c < D; ++c) {
        const size_t o = (size_t)c * d.n + i;
        e += a[o] * vv[c];
        qn[o] = q[o] + v.h * vv[c];
    }
    double *ef = v.efwd + d.vo + i;
    *ef = (t == 0 ? 0.0 : *ef) + e;
}

template <int D>
__global__ __launch_bounds__(kTile) void kb_bwd_reduce(BatchView v, int t, const int4 *__restrict__ items)
{
    const int4 it = items[blockIdx.x];
    const BatchDesc d = v.desc[it.x];
    const int i = it.y * kTile + threadIdx.x;
    if (i >= d.n) return;
    double vv[2 * D];
    sum_splits<2 * D>(v.part + d.po, d.splits, d.n, i, vv);
    double *g = v.grd + (size_t)(v.T - 1) * D * d.vo + (size_t)t * D * d.n;
    double *p = v.padj + (size_t)D * d.vo;
#pragma unroll
    for (int c = 0; c < D; ++c) {
        const size_t o = (size_t)c * d.n + i;
        g[o] = v.h * vv[c];
        p[o] = p[o] - v.h * vv[D + c];
    }
}

// ---------------------------------------------------------------- varifold
template <int D>
__global__ __launch_bounds__(kTile) void kb_geometry(BatchView v, const int4 *__restrict__ items)
{
    const int4 it = items[blockIdx.x];
    const BatchDesc d = v.desc[it.x];
    const int e = it.y * kTile + threadIdx.x;
    if (e >= d.ne) return;
    const double *q = v.nde + (size_t)v.T * D * d.vo + (size_t)(v.T - 1) * D * d.n;
    const int *elm = v.elm + (size_t)D * d.eo;   // D vertices per element (segment / triangle)
    double *cen = v.ecen + (size_t)D * d.eo, *dir = v.edir + (size_t)D * d.eo, *vol = v.evol + d.eo;
    const int ne = d.ne, n = d.n;
    if constexpr (D == 2) {
        const int i0 = elm[e], i1 = elm[ne + e];
        double x0 = q[i0], y0 = q[n + i0], x1 = q[i1], y1 = q[n + i1];
        double ux = x1 - x0, uy = y1 - y0, a = sqrt(ux * ux + uy * uy);
        cen[e] = 0.5 * (x0 + x1);
        cen[ne + e] = 0.5 * (y0 + y1);
        dir[e] = ux / a;
        dir[ne + e] = uy / a;
        vol[e] = a;
    } else {
        const int i0 = elm[e], i1 = elm[ne + e], i2 = elm[2 * ne + e];
        double e1[3], e2[3], nn[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double p0 = q[(size_t)c * n + i0], p1 = q[(size_t)c * n + i1], p2 = q[(size_t)c * n + i2];
            e1[c] = p1 - p0;
            e2[c] = p2 - p0;
            cen[(size_t)c * ne + e] = (p0 + p1 + p2) / 3.0;
        }
        nn[0] = e1[1] * e2[2] - e1[2] * e2[1];
        nn[1] = -e1[0] * e2[2] + e1[2] * e2[0];
        nn[2] = e1[0] * e2[1] - e1[1] * e2[0];
        double l = sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) dir[(size_t)c * ne + e] = nn[c] / l;
        vol[e] = 0.5 * l;
    }
}

template <int D, int CEN, int DIR, class P>
__global__ __launch_bounds__(kBlock) void kb_vf(BatchView v, double ics2, double ids2,
                                                const int4 *__restrict__ items)
{
    using Op = OpVf<D, CEN, DIR, P>;
    using real = typename Op::real;
    __shared__ __align__(16) char smem[PairSmem<Op>::bytes];
    const int4 it = items[xcd_remap(blockIdx.x, gridDim.x)];
    const BatchDesc d = v.desc[it.x];
    const double *ci = v.ecen + (size_t)D * d.eo, *ui = v.edir + (size_t)D * d.eo;
    const bool tgt = it.z >= d.vsd;
    typename Op::Args A;
    A.ci = ci;
    A.ui = ui;
    A.ni = d.ne;
    A.cj = tgt ? v.tc + (size_t)D * d.to : ci;
    A.uj = tgt ? v.tu + (size_t)D * d.to : ui;
    A.aj = tgt ? v.ta + d.to : v.evol + d.eo;
    A.nj = tgt ? d.nt : d.ne;
    A.ics2 = (real)ics2;
    A.ids2 = (real)ids2;
    A.cut2 = cull_r2(v.cull);
    A.xr2 = expand_r2(v.expand != 0);
    A.kap = cen_kappa<CEN>(ics2);
    A.exec = v.exec;
    A.grp = v.gvf ? (tgt ? v.grpt + d.gto : v.grpe + d.geo) : nullptr;
    const int j0 = (tgt ? it.z - d.vsd : it.z) * d.vchunk, j1 = min(A.nj, j0 + d.vchunk);
    pairs_block<Op>(A, d.ne, it.y, j0, j1, v.vpart + d.vpo + (size_t)it.z * (2 * D + 2) * d.ne, smem);
}

template <int D>
__global__ __launch_bounds__(kTile) void kb_vf_finalize(BatchView v, double gscale, const int4 *__restrict__ items)
{
    constexpr int NC = 2 * D + 2;
    const int4 it = items[blockIdx.x];
    const BatchDesc d = v.desc[it.x];
    const int i = it.y * kTile + threadIdx.x;
    if (i >= d.ne) return;
    const int ne = d.ne;
    double acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = 0.0;
    const int splits = d.vsd + d.vst;
    for (int s = 0; s < splits; ++s) {
        const double *o = v.vpart + d.vpo + (size_t)s * NC * ne + i;
        const double we = s < d.vsd ? 1.0 : -2.0, wg = s < d.vsd ? 1.0 : -1.0;
        acc[0] += we * o[0];
#pragma unroll
        for (int c = 1; c < NC; ++c) acc[c] += wg * o[(size_t)c * ne];
    }
    const double a = v.evol[d.eo + i];
    const double dv = (D == 3) ? 0.5 : 1.0;
    const double *dir = v.edir + (size_t)D * d.eo;
    double *dc = v.dc + (size_t)D * d.eo, *dd = v.dd + (size_t)D * d.eo;
#pragma unroll
    for (int c = 0; c < D; ++c) {
        dc[(size_t)c * ne + i] = gscale * a * acc[2 + c];
        dd[(size_t)c * ne + i] = 2.0 * dv * (acc[2 + D + c] + acc[1] * dir[(size_t)c * ne + i]);
    }
    v.evf[d.eo + i] = a * acc[0];
}

template <int D>
__global__ __launch_bounds__(kTile) void kb_vf_gather(BatchView v, const int4 *__restrict__ items)
{
    const int4 it = items[blockIdx.x];
    const BatchDesc d = v.desc[it.x];
    const int n = it.y * kTile + threadIdx.x;
    if (n >= d.n) return;
    const int nq = d.n, ne = d.ne;
    const int *ifo = v.ifo + (size_t)v.rows * d.vo;
    const double *q = v.nde + (size_t)v.T * D * d.vo + (size_t)(v.T - 1) * D * d.n;
    const double *dc = v.dc + (size_t)D * d.eo, *dd = v.dd + (size_t)D * d.eo;
    double g[D];
#pragma unroll
    for (int c = 0; c < D; ++c) g[c] = 0.0;
    const int adj = ifo[n];
    for (int k = 0; k < adj; ++k) {
        if constexpr (D == 2) {
            const int e = ifo[(size_t)(1 + 2 * k) * nq + n];
            const int sgn = ifo[(size_t)(2 + 2 * k) * nq + n];
            g[0] += 0.5 * dc[e] + sgn * dd[e];
            g[1] += 0.5 * dc[ne + e] + sgn * dd[ne + e];
        } else {
            const int e = ifo[(size_t)(1 + 3 * k) * nq + n];
            const int ia = ifo[(size_t)(2 + 3 * k) * nq + n];
            const int ib = ifo[(size_t)(3 + 3 * k) * nq + n];
            double ab[3], dv[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                ab[c] = q[(size_t)c * nq + ia] - q[(size_t)c * nq + ib];
                dv[c] = dd[(size_t)c * ne + e];
            }
            g[0] += dc[e] / 3.0 + (dv[1] * ab[2] - dv[2] * ab[1]);
            g[1] += dc[ne + e] / 3.0 + (-dv[0] * ab[2] + dv[2] * ab[0]);
            g[2] += dc[2 * (size_t)ne + e] / 3.0 + (dv[0] * ab[1] - dv[1] * ab[0]);
        }
    }
    double *p = v.padj + (size_t)D * d.vo;
#pragma unroll
    for (int c = 0; c < D; ++c) p[(size_t)c * nq + n] = -v.w * g[c];
}

template <int D>
__global__ __launch_bounds__(kTile) void kb_init_flow(BatchView v, const int4 *__restrict__ items)
{
    const int4 it = items[blockIdx.x];
    const BatchDesc d = v.desc[it.x];
    const int i = it.y * kTile + threadIdx.x;
    if (i >= d.n) return;
    double *q = v.nde + (size_t)v.T * D * d.vo;
    const double *q0 = v.q0 + (size_t)D * d.vo;
#pragma unroll
    for (int c = 0; c < D; ++c) q[(size_t)c * d.n + i] = q0[(size_t)c * d.n + i];
}

__device__ __forceinline__ double block_sum256(double x, double *red)
{
    red[threadIdx.x] = x;
    __syncthreads();
#pragma unroll
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    return red[0];
}

// one block per listed problem: f = h/2 sum efwd + w (sum evf + tt)
__global__ __launch_bounds__(kBlock) void kb_energy(BatchView v, const double *__restrict__ tt,
                                                    const int *__restrict__ probs, double *__restrict__ f)
{
    __shared__ double red[kBlock];
    const int p = probs[blockIdx.x];
    const BatchDesc d = v.desc[p];
    double s = 0.0;
    for (int i = threadIdx.x; i < d.n; i += kBlock) s += v.efwd[d.vo + i];
    const double ef = block_sum256(s, red);
    __syncthreads();
    s = 0.0;
    for (int i = threadIdx.x; i < d.ne; i += kBlock) s += v.evf[d.eo + i];
    const double ev = block_sum256(s, red);
    if (threadIdx.x == 0) f[p] = v.h * 0.5 * ef + v.w * (ev + tt[p]);
}

// target-target term, once per problem
template <int D, int CEN, int DIR>
__global__ __launch_bounds__(kBlock) void kb_tt(BatchView v, double ics2, double ids2, const int4 *__restrict__ items)
{
    using Op = OpVfValue<D, CEN, DIR>;
    __shared__ __align__(16) char smem[PairSmem<Op>::bytes];
    const int4 it = items[xcd_remap(blockIdx.x, gridDim.x)];
    const BatchDesc d = v.desc[it.x];
    typename Op::Args A{v.tc + (size_t)D * d.to, v.tu + (size_t)D * d.to, v.ta + d.to, d.nt, ics2, ids2};
    const int j0 = it.z * d.vchunk, j1 = min(d.nt, j0 + d.vchunk);
    pairs_block<Op>(A, d.nt, it.y, j0, j1, v.vpart + d.tpo + (size_t)it.z * d.nt, smem);
}

__global__ __launch_bounds__(kBlock) void kb_tt_sum(BatchView v, double *__restrict__ tt)
{
    __shared__ double red[kBlock];
    const int p = blockIdx.x;
    const BatchDesc d = v.desc[p];
    double s = 0.0;
    for (int i = threadIdx.x; i < d.nt; i += kBlock) {
        double a = 0.0;
        for (int k = 0; k < d.tsplits; ++k) a += v.vpart[d.tpo + (size_t)k * d.nt + i];
        s += v.ta[d.to + i] * a;
    }
    const double r = block_sum256(s, red);
    if (threadIdx.x == 0) tt[p] = r;
}

// ------------------------------------------------------------- vectors
// A per-problem vector: base + slot[p]*stride + voff[p] (slot == nullptr:
// no slot; slot[p] < 0: problem skipped).
struct VRef {
    const double *base;
    const int *slot;
    long long stride;
};

struct VecTab {
    const int2 *blk;          // (problem, block within problem)
    const long long *voff;    // per-problem variable offset
    const long long *vn;      // per-problem variable count
};

__device__ __forceinline__ bool vref_ptr(const VRef &r, int p, long long voff, const double *&out)
{
    long long o = voff;
    if (r.slot) {
        const int s = r.slot[p];
        if (s < 0) return false;
        o += (long long)s * r.stride;
    }
    out = r.base + o;
    return true;
}

struct BDotArgs {
    VRef x[kMaxBDots], y[kMaxBDots];
};

// First stage of nk batched dot products: one block per chunk of kVecBlk
// elements of one problem (VecTab) computes all nk partial sums of that
// chunk, so that a vector shared by several products (g against every ring
// slot, d in g.d and d.d) streams from HBM once per chunk and its repeated
// reads hit the cache.  Thread t sums elements lo + t, lo + t + kBlock, ... in
// order, and the block sum of each product pairs t with t + w for w = 128, ...,
// 1 (block_sum256's tree): a fixed order, whatever the launch.  Dynamic LDS:
// nk x kBlock doubles.
__global__ __launch_bounds__(kBlock) void kb_dots1_all(BDotArgs a, int nk, VecTab vt, const int *__restrict__ mask,
                                                       double *__restrict__ part, int nent)
{
    extern __shared__ double redk[];
    const int2 b = vt.blk[blockIdx.x];
    const int p = b.x;
    const bool live = !mask || mask[p];
    const long long lo = (long long)b.y * kVecBlk, hi = min(vt.vn[p], lo + kVecBlk);
    const double *x[kMaxBDots], *y[kMaxBDots];
    bool ok[kMaxBDots];
    double s[kMaxBDots];
#pragma unroll
    for (int k = 0; k < kMaxBDots; ++k) {
        s[k] = 0.0;
        x[k] = y[k] = nullptr;
        ok[k] = live && k < nk && vref_ptr(a.x[k], p, vt.voff[p], x[k]) && vref_ptr(a.y[k], p, vt.voff[p], y[k]);
    }
    for (long long i = lo + threadIdx.x; i < hi; i += kBlock) {
#pragma unroll
        for (int k = 0; k < kMaxBDots; ++k)
            if (ok[k]) s[k] += x[k][i] * y[k][i];
    }
#pragma unroll
    for (int k = 0; k < kMaxBDots; ++k)
        if (k < nk) redk[k * kBlock + threadIdx.x] = s[k];
    __syncthreads();
#pragma unroll
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < nk; ++k) redk[k * kBlock + threadIdx.x] += redk[k * kBlock + threadIdx.x + w];
        __syncthreads();
    }
    if ((int)threadIdx.x < nk) part[(size_t)threadIdx.x * nent + blockIdx.x] = redk[threadIdx.x * kBlock];
}

// one block per problem: its entries are first[p] .. first[p]+count[p]-1
__global__ __launch_bounds__(kBlock) void kb_dots2(const double *__restrict__ part, int nent,
                                                   const int *__restrict__ first, const int *__restrict__ count,
                                                   const int *__restrict__ mask, double *__restrict__ out, int P)
{
    __shared__ double red[kBlock];
    const int p = blockIdx.x, k = blockIdx.y;
    if (mask && !mask[p]) return;
    double s = 0.0;
    for (int i = threadIdx.x; i < count[p]; i += kBlock) s += part[(size_t)k * nent + first[p] + i];
    const double r = block_sum256(s, red);
    if (threadIdx.x == 0) out[(size_t)k * P + p] = r;
}

// generic per-problem update over the block table
enum VecOp : int { kTry = 0, kAccept };

struct VecArgs {
    int op;
    double *d;                 // xt (kTry)
    const double *g, *x;       // direction / iterate (kTry)
    double *Sw, *Yw;           // ring buffers [m][vtot] (accept)
    double *xw, *gw;           // iterate / gradient (accept)
    const double *xt, *gt;     // trial point (accept)
    long long vtot;
    const int *mask;           // per problem, 0 = skip
    const int *slot;           // per problem ring slot
    const double *rel;         // per problem step (kTry)
};

__global__ __launch_bounds__(kBlock) void kb_vec(VecArgs a, VecTab vt)
{
    const int2 b = vt.blk[blockIdx.x];
    const int p = b.x;
    if (a.mask && !a.mask[p]) return;
    const int sl = a.slot ? a.slot[p] : 0;
    if (sl < 0) return;
    const long long vo = vt.voff[p];
    const long long lo = (long long)b.y * kVecBlk, hi = min(vt.vn[p], lo + kVecBlk);
    const long long so = (long long)sl * a.vtot + vo;
    const double coef = a.op == kTry ? a.rel[p] : 0.0;
    for (long long i = lo + threadIdx.x; i < hi; i += kBlock) {
        const long long k = vo + i;
        switch (a.op) {
        case kTry: a.d[k] = a.x[k] + coef * a.g[k]; break;   // d := xt, g := direction
        case kAccept: {
            double xn = a.xt[k], gn = a.gt[k];
            a.Sw[so + i] = xn - a.xw[k];
            a.Yw[so + i] = gn - a.gw[k];
            a.xw[k] = xn;
            a.gw[k] = gn;
            break;
        }
        default: break;
        }
    }
}

// L-BFGS direction in the compact form of the two-loop recursion
// (Cpp/src/getDirection.cpp).  With a_i = s_i.g, b_i = y_i.g and per problem
// the tables SY[i][j] = s_i.y_j and YY[i][j] = y_i.y_j over the physical ring
// slots, the recursion's coefficients follow from scalars alone:
//   first loop (newest to oldest)
//     alpha_i = (a_i - sum_{j newer} alpha_j SY[i][j]) / SY[i][i],
//   H0 = gamma = SY[n][n] / YY[n][n] (n the newest pair),
//   second loop (oldest to newest)
//     beta_i = (gamma (b_i - sum_j alpha_j YY[i][j])
//               + sum_{j older} (alpha_j - beta_j) SY[j][i]) / SY[i][i],
// the same alpha and beta as the two loops (s_i.q and y_i.r re-associated),
// and the direction is one pass,
//   d = -gamma g + sum_j gamma alpha_j y_j - sum_j (alpha_j - beta_j) s_j.
// One thread per problem; coef[p] = [g, y_0..y_{m-1}, s_0..s_{m-1}]
// coefficients by slot.
__global__ void kb_lbfgs_coef(const double *__restrict__ dots, const double *__restrict__ SY,
                              const double *__restrict__ YY, const int *__restrict__ dir,
                              const int *__restrict__ his, const int *__restrict__ newest, double *__restrict__ ab,
                              double *__restrict__ coef, int m, int P)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P || !dir[p]) return;
    double *c = coef + (size_t)p * (1 + 2 * m), *al = ab + (size_t)p * 2 * m, *be = al + m;
    for (int j = 0; j < 1 + 2 * m; ++j) c[j] = 0.0;
    const int h = his[p];
    if (h == 0) {   // first iteration: d = -g
        c[0] = -1.0;
        return;
    }
    const double *sy = SY + (size_t)p * m * m, *yy = YY + (size_t)p * m * m;
    const int n = newest[p];
    auto slot = [&](int k) {   // k = 0: newest pair
        const int i = n - k;
        return i < 0 ? i + m : i;
    };
    for (int k = 0; k < h; ++k) {
        const int i = slot(k);
        double t = dots[(size_t)i * P + p];
        for (int q = 0; q < k; ++q) t -= al[slot(q)] * sy[i * m + slot(q)];
        al[i] = t / sy[i * m + i];
    }
    const double gam = sy[n * m + n] / yy[n * m + n];
    for (int k = h - 1; k >= 0; --k) {
        const int i = slot(k);
        double t = dots[(size_t)(m + i) * P + p];
        for (int q = 0; q < h; ++q) t -= al[slot(q)] * yy[i * m + slot(q)];
        t *= gam;
        for (int q = h - 1; q > k; --q) t += (al[slot(q)] - be[slot(q)]) * sy[slot(q) * m + i];
        be[i] = t / sy[i * m + i];
    }
    c[0] = -gam;
    for (int k = 0; k < h; ++k) {
        const int j = slot(k);
        c[1 + j] = gam * al[j];
        c[1 + m + j] = -(al[j] - be[j]);
    }
}

// d = coef_g g + sum over the filled slots (0 .. his-1) of coef_y Y + coef_s S
__global__ __launch_bounds__(kBlock) void kb_dir(double *__restrict__ d, const double *__restrict__ g,
                                                 const double *__restrict__ S, const double *__restrict__ Y,
                                                 long long vtot, const double *__restrict__ coef,
                                                 const int *__restrict__ dir, const int *__restrict__ his, int m,
                                                 VecTab vt)
{
    const int2 b = vt.blk[blockIdx.x];
    const int p = b.x;
    if (!dir[p]) return;
    const long long vo = vt.voff[p];
    const long long lo = (long long)b.y * kVecBlk, hi = min(vt.vn[p], lo + kVecBlk);
    const double *c = coef + (size_t)p * (1 + 2 * m);
    const int h = his[p];
    for (long long i = lo + threadIdx.x; i < hi; i += kBlock) {
        const long long k = vo + i;
        double v = c[0] * g[k];
        for (int j = 0; j < h; ++j) v += c[1 + j] * Y[(size_t)j * vtot + k] + c[1 + m + j] * S[(size_t)j * vtot + k];
        d[k] = v;
    }
}

// the newest pair's row and column of SY and YY: dots[3j..3j+2][p] =
// s_n.y_j, s_j.y_n, y_n.y_j for the filled slots j (n = newest)
__global__ void kb_store_ring(const double *__restrict__ dots, const int *__restrict__ mask,
                              const int *__restrict__ newest, const int *__restrict__ his, double *__restrict__ SY,
                              double *__restrict__ YY, int m, int P)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P || !mask[p]) return;
    const int n = newest[p];
    double *sy = SY + (size_t)p * m * m, *yy = YY + (size_t)p * m * m;
    for (int j = 0; j < his[p]; ++j) {
        sy[n * m + j] = dots[(size_t)(3 * j) * P + p];
        sy[j * m + n] = dots[(size_t)(3 * j + 1) * P + p];
        yy[n * m + j] = yy[j * m + n] = dots[(size_t)(3 * j + 2) * P + p];
    }
}

// a runtime number of targets per lane (1, 2 or 4) onto a template lambda
template <typename F>
void with_lanes(int r, F &&f)
{
    switch (r) {
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    default: throw Error(kErrArg, "targets per lane must be 1, 2 or 4");
    }
}

int cen_code(char t) { return t == 'G' ? 0 : 1; }
int dir_code(char t) { return t == 'B' ? 0 : t == 'L' ? 1 : t == 'O' ? 2 : 3; }

template <typename F>
void dispatch_vf3(int dim, const VfParams &vp, F &&f)
{
    const int cc = cen_code(vp.cen_type), dc = dir_code(vp.dir_type);
    auto with_dim = [&](auto C, auto U) {
        if (dim == 2) f(std::integral_constant<int, 2>{}, C, U);
        else f(std::integral_constant<int, 3>{}, C, U);
    };
    auto with_dir = [&](auto C) {
        switch (dc) {
        case 0: with_dim(C, std::integral_constant<int, 0>{}); break;
        case 1: with_dim(C, std::integral_constant<int, 1>{}); break;
        case 2: with_dim(C, std::integral_constant<int, 2>{}); break;
        default: with_dim(C, std::integral_constant<int, 3>{}); break;
        }
    };
    if (cc == 0) with_dir(std::integral_constant<int, 0>{});
    else with_dir(std::integral_constant<int, 1>{});
}

double inv_sq(double w) { return w != 0.0 ? 1.0 / (w * w) : 0.0; }

// Device buffer owned by a unique_ptr-like holder
template <typename T>
struct DArr {
    T *p = nullptr;
    size_t n = 0;
    void ensure(size_t count)
    {
        if (count <= n) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        if (hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess)
            throw Error(kErrAlloc, "device allocation failed");
        n = count;
    }
    ~DArr()
    {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

// work tables of one active set: per precision group the point-kernel and
// varifold items; vertex / element tiles and the list of problems
struct Batch::Tables {
    DArr<int4> kv[2], bwd[2], vf[2], vtile, etile;
    int nkv[2] = {0, 0}, nbwd[2] = {0, 0}, nvf[2] = {0, 0}, nvtile = 0, netile = 0;
    DArr<int> probs;
    int nprobs = 0;
};

// ======================================================================
Batch::Batch(const BatchSpec &spec, const double *q0, const int *elm, const double *tc, const double *ta,
             const double *tu, hipStream_t st)
    : spec_(spec), st_(st), tab_(new Tables)
{
    // hipGraph replay by default; off when rocprofv3 drives the process
    // (its kernel tracing has crashed the host process during stream capture
    // on this ROCm release) or when LDDMM_BATCH_GRAPHS=0
    for (char **e = environ; e && *e; ++e)
        if (std::strncmp(*e, "ROCPROF", 7) == 0) use_graphs_ = false;
    if (const char *e = std::getenv("LDDMM_BATCH_GRAPHS")) use_graphs_ = std::atoi(e) != 0;
    const int P = spec_.P(), D = spec_.dim;
    if (P <= 0) throw Error(kErrArg, "a batch needs at least one problem");
    if (D != 2 && D != 3) throw Error(kErrArg, "dimension must be 2 or 3");
    if (spec_.T < 2) throw Error(kErrArg, "timeNum must be at least 2");
    if (spec_.order < -1 || spec_.order > 4) throw Error(kErrArg, "knlOrder must be -1 or 0..4");
    if ((int)spec_.ne.size() != P || (int)spec_.nt.size() != P || (int)spec_.width.size() != P)
        throw Error(kErrArg, "per-problem arrays must have one entry per problem");
    {
        DeviceProblem chk;
        chk.dim = D;
        chk.n = 1;
        chk.T = spec_.T;
        chk.order = spec_.order;
        chk.vp = spec_.vp;
        chk.ne = 1;
        validate(chk);
    }
    vo_.resize(P);
    eo_.resize(P);
    to_.resize(P);
    for (int p = 0; p < P; ++p) {
        if (spec_.n[p] <= 0 || spec_.ne[p] <= 0 || spec_.nt[p] < 0) throw Error(kErrArg, "bad problem size");
        if (!(spec_.width[p] > 0.0)) throw Error(kErrArg, "kernel widths must be positive");
        vo_[p] = ntot_;
        eo_[p] = etot_;
        to_[p] = ttot_;
        ntot_ += spec_.n[p];
        etot_ += spec_.ne[p];
        ttot_ += spec_.nt[p];
    }
    vtot_ = (long long)D * (spec_.T - 1) * ntot_;

    // Each problem is held in a spatial order (spatial_order of its initial
    // vertices, element centres and target centres; the element table
    // renumbered), as in Objective; momenta and gradients cross in the
    // host-layout conversions.  Element tables in the internal numbering,
    // host layout [E_p][V]:
    vperm_.resize(ntot_);
    std::vector<int> selm((size_t)D * etot_), tperm(std::max<long long>(ttot_, 1));
    for_each_problem(P, [&](int p) {
        const int n = spec_.n[p], ne = spec_.ne[p], nt = spec_.nt[p];
        const double *qp = q0 + (size_t)D * vo_[p];
        const int *ev = elm + (size_t)D * eo_[p];
        for (long long k = 0; k < (long long)ne * D; ++k)
            if (ev[k] < 0 || ev[k] >= n) throw Error(kErrArg, "element table holds a vertex outside its problem");
        const std::vector<int> vp = spatial_order(n, D, [&](int i, int c) { return qp[(size_t)i * D + c]; });
        std::vector<int> vinv(n);
        for (int s = 0; s < n; ++s) {
            vperm_[vo_[p] + s] = vp[s];
            vinv[vp[s]] = s;
        }
        std::vector<double> cen((size_t)D * ne, 0.0);
        for (int e = 0; e < ne; ++e)
            for (int k = 0; k < D; ++k)
                for (int c = 0; c < D; ++c) cen[(size_t)e * D + c] += qp[(size_t)ev[(size_t)e * D + k] * D + c] / D;
        const std::vector<int> ep = spatial_order(ne, D, [&](int i, int c) { return cen[(size_t)i * D + c]; });
        for (int s = 0; s < ne; ++s)
            for (int k = 0; k < D; ++k) selm[(size_t)D * (eo_[p] + s) + k] = vinv[ev[(size_t)ep[s] * D + k]];
        const double *tp = nt > 0 ? tc + (size_t)D * to_[p] : nullptr;
        const std::vector<int> tq = spatial_order(nt, D, [&](int i, int c) { return tp[(size_t)i * D + c]; });
        for (int s = 0; s < nt; ++s) tperm[to_[p] + s] = tq[s];
    });
    elm = selm.data();

    // adjacency tables per problem (getElmIfo, Cpp/src/getElmIfo.m), padded to one row count
    std::vector<std::vector<int>> ifos(P);
    for (int p = 0; p < P; ++p) {
        const int n = spec_.n[p], ne = spec_.ne[p];
        const int *ev = elm + (size_t)D * eo_[p];
        std::vector<int> cnt(n, 0);
        for (long long k = 0; k < (long long)ne * D; ++k) ++cnt[ev[k]];
        int mx = 0;
        for (int c : cnt) mx = std::max(mx, c);
        rows_ = std::max(rows_, 1 + mx * D);
    }
    std::vector<int> hifo((size_t)rows_ * ntot_, 0);
    for_each_problem(P, [&](int p) {
        const int n = spec_.n[p], ne = spec_.ne[p];
        const int *ev = elm + (size_t)D * eo_[p];
        int *blk = hifo.data() + (size_t)rows_ * vo_[p];   // [rows][n]
        for (int e = 0; e < ne; ++e)
            for (int k = 0; k < D; ++k) {
                const int v = ev[(size_t)e * D + k];
                int &count = blk[v];
                const int base = 1 + count * D;
                blk[(size_t)base * n + v] = e;
                if (D == 2) {
                    blk[(size_t)(base + 1) * n + v] = k == 0 ? -1 : 1;
                } else {
                    static const int pa[3] = {2, 0, 1}, pb[3] = {1, 2, 0};
                    blk[(size_t)(base + 1) * n + v] = ev[(size_t)e * D + pa[k]];
                    blk[(size_t)(base + 2) * n + v] = ev[(size_t)e * D + pb[k]];
                }
                ++count;
            }
    });
    // host layout -> per-problem SoA blocks
    std::vector<double> hq((size_t)D * ntot_), htc((size_t)D * std::max<long long>(ttot_, 1)),
        htu((size_t)D * std::max<long long>(ttot_, 1));
    std::vector<int> helm((size_t)D * etot_);
    std::vector<double> hta(std::max<long long>(ttot_, 1));
    for_each_problem(P, [&](int p) {
        const int n = spec_.n[p], ne = spec_.ne[p], nt = spec_.nt[p];
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < D; ++c)
                hq[(size_t)D * vo_[p] + (size_t)c * n + i] = q0[(size_t)D * (vo_[p] + vperm_[vo_[p] + i]) + c];
        for (int e = 0; e < ne; ++e)
            for (int c = 0; c < D; ++c)
                helm[(size_t)D * eo_[p] + (size_t)c * ne + e] = elm[(size_t)D * (eo_[p] + e) + c];
        for (int e = 0; e < nt; ++e) {
            const long long src = to_[p] + tperm[to_[p] + e];
            for (int c = 0; c < D; ++c) {
                htc[(size_t)D * to_[p] + (size_t)c * nt + e] = tc[(size_t)D * src + c];
                htu[(size_t)D * to_[p] + (size_t)c * nt + e] = tu[(size_t)D * src + c];
            }
            hta[to_[p] + e] = ta[src];
        }
    });

    chunk_ = spec_.chunk > 0 ? ((spec_.chunk + 63) / 64) * 64 : 1024;
    vchunk_ = 1024;
    rkv_ = kKvR32;
    rbwd_ = bwd_r32(spec_.order);
    build_descs();
    size_t part = 0, vpart = 0;
    for (int p = 0; p < P; ++p) {
        part = std::max<size_t>(part, desc_h_[p].po + (size_t)desc_h_[p].splits * 2 * D * spec_.n[p]);
        vpart = std::max<size_t>(vpart, desc_h_[p].vpo + (size_t)(desc_h_[p].vsd + desc_h_[p].vst) * (2 * D + 2) *
                                                             spec_.ne[p]);
        vpart = std::max<size_t>(vpart, desc_h_[p].tpo + (size_t)desc_h_[p].tsplits * std::max(spec_.nt[p], 1));
    }
    // the point-kernel partial buffer is sized for the smallest tunable chunk
    size_t part_max = 0;
    for (int p = 0; p < P; ++p) part_max += (size_t)ceil_div(spec_.n[p], 256) * 2 * D * spec_.n[p];
    part = std::max(part, part_max);
    part_cap_ = part;

    slab_.reserve(&desc_, P);
    slab_.reserve(&q0_, hq.size());
    slab_.reserve(&elm_, helm.size());
    slab_.reserve(&ifo_, hifo.size());
    slab_.reserve(&tc_, htc.size());
    slab_.reserve(&tu_, htu.size());
    slab_.reserve(&ta_, std::max<long long>(ttot_, 1));
    slab_.reserve(&nde_, (size_t)spec_.T * D * ntot_);
    slab_.reserve(&padj_, (size_t)D * ntot_);
    slab_.reserve(&part_, part);
    slab_.reserve(&vpart_, vpart);
    slab_.reserve(&ecen_, (size_t)D * etot_);
    slab_.reserve(&edir_, (size_t)D * etot_);
    slab_.reserve(&evol_, etot_);
    slab_.reserve(&dc_, (size_t)D * etot_);
    slab_.reserve(&dd_, (size_t)D * etot_);
    slab_.reserve(&efwd_, ntot_);
    slab_.reserve(&evf_, etot_);
    slab_.reserve(&tt_, P);
    {
        long long gv = 0, ge = 0, gt = 0;
        for (int p = 0; p < P; ++p) {
            gv += (long long)spec_.T * ((spec_.n[p] + 63) / 64);
            ge += (spec_.ne[p] + 63) / 64;
            gt += (spec_.nt[p] + 63) / 64;
        }
        slab_.reserve(&grpv_, (size_t)std::max(gv, 1LL));
        slab_.reserve(&grpm_, (size_t)std::max(gv / spec_.T, 1LL));
        slab_.reserve(&grpe_, (size_t)std::max(ge, 1LL));
        slab_.reserve(&grpt_, (size_t)std::max(gt, 1LL));
    }
    slab_.allocate();
    LDDMM_HIP(hipMemcpyAsync(q0_, hq.data(), hq.size() * sizeof(double), hipMemcpyHostToDevice, st_));
    LDDMM_HIP(hipMemcpyAsync(elm_, helm.data(), helm.size() * sizeof(int), hipMemcpyHostToDevice, st_));
    LDDMM_HIP(hipMemcpyAsync(ifo_, hifo.data(), hifo.size() * sizeof(int), hipMemcpyHostToDevice, st_));
    if (ttot_ > 0) {
        LDDMM_HIP(hipMemcpyAsync(tc_, htc.data(), (size_t)D * ttot_ * sizeof(double), hipMemcpyHostToDevice, st_));
        LDDMM_HIP(hipMemcpyAsync(tu_, htu.data(), (size_t)D * ttot_ * sizeof(double), hipMemcpyHostToDevice, st_));
        LDDMM_HIP(hipMemcpyAsync(ta_, hta.data(), (size_t)ttot_ * sizeof(double), hipMemcpyHostToDevice, st_));
    }
    LDDMM_HIP(hipMemcpyAsync(desc_, desc_h_.data(), P * sizeof(BatchDesc), hipMemcpyHostToDevice, st_));
    LDDMM_HIP(hipStreamSynchronize(st_));
    compute_target_terms();
    if (spec_.chunk <= 0) autotune();
}

Batch::~Batch() { clear_graphs(); }

void Batch::build_descs()
{
    const int P = spec_.P(), D = spec_.dim;
    desc_h_.assign(P, BatchDesc{});
    long long po = 0, vpo = 0, gvo = 0, geo = 0, gto = 0;
    for (int p = 0; p < P; ++p) {
        BatchDesc &d = desc_h_[p];
        d.vo = vo_[p];
        d.eo = eo_[p];
        d.to = to_[p];
        d.n = spec_.n[p];
        d.ne = spec_.ne[p];
        d.nt = spec_.nt[p];
        d.chunk = chunk_;
        d.splits = ceil_div(d.n, chunk_);
        d.vchunk = vchunk_;
        d.vsd = ceil_div(d.ne, vchunk_);
        d.vst = d.nt > 0 ? ceil_div(d.nt, vchunk_) : 0;
        d.tsplits = d.nt > 0 ? ceil_div(d.nt, vchunk_) : 0;
        d.iw = 1.0 / spec_.width[p];
        d.iw2 = 1.0 / (spec_.width[p] * spec_.width[p]);
        d.po = po;
        po += (long long)d.splits * 2 * D * d.n;
        d.vpo = vpo;
        d.tpo = vpo;   // the target-target pass runs once, before any evaluation
        d.gvo = gvo;
        d.geo = geo;
        d.gto = gto;
        gvo += (long long)spec_.T * ((d.n + 63) / 64);
        geo += (d.ne + 63) / 64;
        gto += (d.nt + 63) / 64;
        vpo += std::max((long long)(d.vsd + d.vst) * (2 * D + 2) * d.ne, (long long)d.tsplits * std::max(d.nt, 1));
    }
}

double Batch::pairs_per_eval(int p) const
{
    const double n = spec_.n[p], T1 = spec_.T - 1;
    return 2.0 * T1 * n * n + (double)spec_.ne[p] * spec_.ne[p] + (double)spec_.ne[p] * spec_.nt[p];
}

void Batch::build_tables(const std::vector<int> &prec_of)
{
    if (prec_of == tab_key_) return;
    clear_graphs();   // captured launches refer to the old tables
    const int P = spec_.P();
    Tables &T = *tab_;
    std::vector<int4> kv[2], bwd[2], vf[2], vt, et;
    std::vector<int> probs;
    for (int p = 0; p < P; ++p) {
        const int pr = prec_of[p];
        if (pr < 0) continue;
        const int g = pr == kFP32 ? 1 : 0;
        const BatchDesc &d = desc_h_[p];
        probs.push_back(p);
        // FP32 pair kernels keep rkv_ / rbwd_ targets per lane (autotuned)
        const int tkv = g ? 64 * rkv_ : 64, tbwd = g ? 64 * rbwd_ : 64;
        for (int s = 0; s < d.splits; ++s) {
            for (int t = 0; t < ceil_div(d.n, tkv); ++t) kv[g].push_back(make_int4(p, t, s, 0));
            for (int t = 0; t < ceil_div(d.n, tbwd); ++t) bwd[g].push_back(make_int4(p, t, s, 0));
        }
        for (int s = 0; s < d.vsd + d.vst; ++s)
            for (int t = 0; t < ceil_div(d.ne, kTile); ++t) vf[g].push_back(make_int4(p, t, s, 0));
        for (int t = 0; t < ceil_div(d.n, kTile); ++t) vt.push_back(make_int4(p, t, 0, 0));
        for (int t = 0; t < ceil_div(d.ne, kTile); ++t) et.push_back(make_int4(p, t, 0, 0));
    }
    auto up = [&](DArr<int4> &a, int &n, const std::vector<int4> &h) {
        n = (int)h.size();
        a.ensure(h.size());
        if (!h.empty())
            LDDMM_HIP(hipMemcpyAsync(a.p, h.data(), h.size() * sizeof(int4), hipMemcpyHostToDevice, st_));
    };
    for (int g = 0; g < 2; ++g) {
        up(T.kv[g], T.nkv[g], kv[g]);
        up(T.bwd[g], T.nbwd[g], bwd[g]);
        up(T.vf[g], T.nvf[g], vf[g]);
    }
    up(T.vtile, T.nvtile, vt);
    up(T.etile, T.netile, et);
    T.nprobs = (int)probs.size();
    T.probs.ensure(probs.size());
    if (!probs.empty())
        LDDMM_HIP(hipMemcpyAsync(T.probs.p, probs.data(), probs.size() * sizeof(int), hipMemcpyHostToDevice, st_));
    LDDMM_HIP(hipStreamSynchronize(st_));   // host vectors go out of scope
    tab_key_ = prec_of;
}

void Batch::compute_target_terms()
{
    const int P = spec_.P(), D = spec_.dim;
    std::vector<int4> items;
    for (int p = 0; p < P; ++p) {
        const BatchDesc &d = desc_h_[p];
        for (int s = 0; s < d.tsplits; ++s)
            for (int t = 0; t < ceil_div(d.nt, kTile); ++t) items.push_back(make_int4(p, t, s, 0));
    }
    BatchView v{};
    v.desc = desc_;
    v.vpart = vpart_;
    v.tc = tc_;
    v.tu = tu_;
    v.ta = ta_;
    DArr<int4> it;
    if (!items.empty()) {
        it.ensure(items.size());
        LDDMM_HIP(hipMemcpyAsync(it.p, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice, st_));
        dispatch_vf3(D, spec_.vp, [&](auto Dc, auto Cc, auto Uc) {
            kb_tt<Dc(), Cc(), Uc()><<<(int)items.size(), kBlock, 0, st_>>>(v, inv_sq(spec_.vp.cen_width),
                                                                           inv_sq(spec_.vp.dir_width), it.p);
        });
        LDDMM_LAUNCH_CHECK();
    }
    kb_tt_sum<<<P, kBlock, 0, st_>>>(v, tt_);
    LDDMM_LAUNCH_CHECK();
    // group bounds of the fixed target centres (FP32 culling of the varifold)
    std::vector<int4> gitems;
    for (int p = 0; p < P; ++p)
        for (int g = 0; g < ceil_div(desc_h_[p].nt, kTile); ++g) gitems.push_back(make_int4(p, g, 0, 0));
    DArr<int4> git;
    if (!gitems.empty()) {
        git.ensure(gitems.size());
        LDDMM_HIP(hipMemcpyAsync(git.p, gitems.data(), gitems.size() * sizeof(int4), hipMemcpyHostToDevice, st_));
        v.grpt = grpt_;
        if (D == 2) kb_group_bounds<2><<<(int)gitems.size(), kTile, 0, st_>>>(v, 0, 2, git.p);
        else kb_group_bounds<3><<<(int)gitems.size(), kTile, 0, st_>>>(v, 0, 2, git.p);
        LDDMM_LAUNCH_CHECK();
    }
    LDDMM_HIP(hipStreamSynchronize(st_));
}

// One batched evaluation replays a captured hipGraph of its ~4T launches
// when the same work tables and buffers were evaluated before: the graph
// is captured on first use per (work tables, alpha, gradient, energy)
// combination and dropped whenever the work tables are rebuilt (a problem
// finished or changed precision).  The captured sequence holds kernel
// launches only (no memset or copy nodes).  With the kernel-time profiler
// on, or graphs switched off, the launches are issued directly.
void Batch::eval(const double *alp, double *grd, double *f_dev, const std::vector<int> &prec_of)
{
    build_tables(prec_of);
    if (tab_->nprobs == 0) return;
    if (!use_graphs_ || Profiler::get().enabled() || pair_counter()) {
        launch_eval(alp, grd, f_dev);
        return;
    }
    const int cull = 2 * pair_cull_bits() + (pair_groups() ? 1 : 0), expand = pair_expansion() ? 1 : 0;
    for (const GraphEntry &e : graphs_)
        if (e.alp == alp && e.grd == grd && e.f == f_dev && e.cull == cull && e.expand == expand) {
            LDDMM_HIP(hipGraphLaunch(e.exec, st_));
            return;
        }
    if (graphs_.size() >= 8) clear_graphs();
    hipGraph_t graph = nullptr;
    LDDMM_HIP(hipStreamBeginCapture(st_, hipStreamCaptureModeRelaxed));
    try {
        launch_eval(alp, grd, f_dev);
    } catch (...) {
        if (hipStreamEndCapture(st_, &graph) == hipSuccess && graph) (void)hipGraphDestroy(graph);
        throw;
    }
    LDDMM_HIP(hipStreamEndCapture(st_, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    LDDMM_HIP(ie);
    graphs_.push_back(GraphEntry{alp, grd, f_dev, cull, expand, exec});
    LDDMM_HIP(hipGraphLaunch(exec, st_));
}

void Batch::clear_graphs()
{
    for (GraphEntry &e : graphs_)
        if (e.exec) (void)hipGraphExecDestroy(e.exec);
    graphs_.clear();
}

void Batch::set_graphs(bool on)
{
    use_graphs_ = on;
    if (!on) clear_graphs();
}

void Batch::launch_eval(const double *alp, double *grd, double *f_dev)
{
    const int D = spec_.dim, T = spec_.T;
    const Tables &tb = *tab_;
    BatchView v{desc_, nde_, padj_, part_, vpart_, alp, grd, q0_, tc_, tu_, ta_, elm_, ifo_, rows_, T,
                spec_.h, spec_.w, ecen_, edir_, evol_, dc_, dd_, efwd_, evf_, pair_cull_bits(),
                pair_expansion() ? 1 : 0, pair_counter(), grpv_, grpe_, grpt_,
                pair_groups() && spec_.order == -1 ? 1 : 0, pair_groups() && spec_.vp.cen_type == 'G' ? 1 : 0};
    v.grpm = grpm_;

    // q_0 of the listed problems
    if (D == 2) kb_init_flow<2><<<tb.nvtile, kTile, 0, st_>>>(v, tb.vtile.p);
    else kb_init_flow<3><<<tb.nvtile, kTile, 0, st_>>>(v, tb.vtile.p);
    LDDMM_LAUNCH_CHECK();
    auto each_prec = [&](auto &&fn) {
        for (int g = 0; g < 2; ++g) {
            if (g == 0) fn(g, PrecF64{});
            else fn(g, PrecF32{});
        }
    };
    auto dispatch = [&](auto &&fn) {
        dispatch_order(spec_.order, [&](auto O) {
            if (D == 2) fn(std::integral_constant<int, 2>{}, O);
            else fn(std::integral_constant<int, 3>{}, O);
        });
    };
    for (int t = 0; t < T - 1; ++t) {
        {
            ProfScope ps(kProfFwdPairs, st_);
            if (v.gpts && tb.nkv[1]) {   // group bounds of q_t for the FP32 culling
                if (D == 2) kb_group_bounds<2><<<tb.nvtile, kTile, 0, st_>>>(v, t, 0, tb.vtile.p);
                else kb_group_bounds<3><<<tb.nvtile, kTile, 0, st_>>>(v, t, 0, tb.vtile.p);
                LDDMM_LAUNCH_CHECK();
            }
            each_prec([&](int g, auto Pc) {
                if (!tb.nkv[g]) return;
                dispatch([&](auto Dc, auto Oc) {
                    with_lanes(decltype(Pc)::fp32 ? rkv_ : 1, [&](auto Rc) {
                        if constexpr (decltype(Pc)::fp32 || Rc() == 1)
                            kb_kv<Dc(), Oc(), decltype(Pc), Rc()><<<tb.nkv[g], kBlock, 0, st_>>>(v, t, tb.kv[g].p);
                    });
                });
                LDDMM_LAUNCH_CHECK();
            });
        }
        ProfScope ps(kProfFwdReduce, st_);
        if (D == 2) kb_fwd_reduce<2><<<tb.nvtile, kTile, 0, st_>>>(v, t, tb.vtile.p);
        else kb_fwd_reduce<3><<<tb.nvtile, kTile, 0, st_>>>(v, t, tb.vtile.p);
        LDDMM_LAUNCH_CHECK();
    }
    {
        ProfScope ps(kProfVfGeometry, st_);
        if (D == 2) kb_geometry<2><<<tb.netile, kTile, 0, st_>>>(v, tb.etile.p);
        else kb_geometry<3><<<tb.netile, kTile, 0, st_>>>(v, tb.etile.p);
        LDDMM_LAUNCH_CHECK();
    }
    {
        ProfScope ps(kProfVfPairs, st_);
        if (v.gvf && tb.nvf[1]) {
            if (D == 2) kb_group_bounds<2><<<tb.netile, kTile, 0, st_>>>(v, 0, 1, tb.etile.p);
            else kb_group_bounds<3><<<tb.netile, kTile, 0, st_>>>(v, 0, 1, tb.etile.p);
            LDDMM_LAUNCH_CHECK();
        }
        each_prec([&](int g, auto Pc) {
            if (!tb.nvf[g]) return;
            dispatch_vf3(D, spec_.vp, [&](auto Dc, auto Cc, auto Uc) {
                kb_vf<Dc(), Cc(), Uc(), decltype(Pc)><<<tb.nvf[g], kBlock, 0, st_>>>(
                    v, inv_sq(spec_.vp.cen_width), inv_sq(spec_.vp.dir_width), tb.vf[g].p);
            });
            LDDMM_LAUNCH_CHECK();
        });
    }
    {
        ProfScope ps(kProfVfFinalize, st_);
        const double gscale = -4.0 * inv_sq(spec_.vp.cen_width);
        if (D == 2) kb_vf_finalize<2><<<tb.netile, kTile, 0, st_>>>(v, gscale, tb.etile.p);
        else kb_vf_finalize<3><<<tb.netile, kTile, 0, st_>>>(v, gscale, tb.etile.p);
        LDDMM_LAUNCH_CHECK();
    }
    {
        ProfScope ps(kProfVfGather, st_);
        if (D == 2) kb_vf_gather<2><<<tb.nvtile, kTile, 0, st_>>>(v, tb.vtile.p);
        else kb_vf_gather<3><<<tb.nvtile, kTile, 0, st_>>>(v, tb.vtile.p);
        LDDMM_LAUNCH_CHECK();
    }
    {
        ProfScope ps(kProfEnergy, st_);
        kb_energy<<<tb.nprobs, kBlock, 0, st_>>>(v, tt_, tb.probs.p, f_dev);
        LDDMM_LAUNCH_CHECK();
    }
    if (grd == nullptr) return;   // energies only
    for (int t = T - 2; t >= 0; --t) {
        {
            ProfScope ps(kProfBwdPairs, st_);
            if (rbwd_ == 4 && mx_ok() && v.gpts && tb.nbwd[1]) {   // column scales of the matrix-core adjoint
                kb_group_maxima<3><<<tb.nvtile, kTile, 0, st_>>>(v, t, tb.vtile.p);
                LDDMM_LAUNCH_CHECK();
            }
            each_prec([&](int g, auto Pc) {
                if (!tb.nbwd[g]) return;
                dispatch([&](auto Dc, auto Oc) {
                    with_lanes(decltype(Pc)::fp32 ? rbwd_ : 1, [&](auto Rc) {
                        if constexpr ((decltype(Pc)::fp32 && Rc() <= 2) || Rc() == 1)
                            kb_bwd<Dc(), Oc(), decltype(Pc), Rc()><<<tb.nbwd[g], kBlock, 0, st_>>>(v, t, tb.bwd[g].p);
                        else if constexpr (decltype(Pc)::fp32 && Rc() == 4 && Dc() == 3 && Oc() == -1)
                            kb_bwd_mx<3><<<tb.nbwd[g], kBlock, 0, st_>>>(v, t, tb.bwd[g].p);
                    });
                });
                LDDMM_LAUNCH_CHECK();
            });
        }
        ProfScope ps(kProfBwdReduce, st_);
        if (D == 2) kb_bwd_reduce<2><<<tb.nvtile, kTile, 0, st_>>>(v, t, tb.vtile.p);
        else kb_bwd_reduce<3><<<tb.nvtile, kTile, 0, st_>>>(v, t, tb.vtile.p);
        LDDMM_LAUNCH_CHECK();
    }
}

void Batch::upload_descs()
{
    LDDMM_HIP(hipMemcpyAsync(desc_, desc_h_.data(), desc_h_.size() * sizeof(BatchDesc), hipMemcpyHostToDevice, st_));
    LDDMM_HIP(hipStreamSynchronize(st_));
    tab_key_.clear();   // the work tables depend on the chunk and the targets per lane
}

void Batch::set_tuning(int chunk, int kv_lanes, int bwd_lanes)
{
    if (chunk <= 0) throw Error(kErrArg, "source chunk must be positive");
    if (kv_lanes != 1 && kv_lanes != 2 && kv_lanes != 4)
        throw Error(kErrArg, "K.alpha targets per lane must be 1, 2 or 4");
    if (bwd_lanes != 1 && bwd_lanes != 2 && !(bwd_lanes == 4 && mx_ok()))
        throw Error(kErrArg, "adjoint targets per lane must be 1 or 2 (4: the matrix-core adjoint of 3D Gaussian "
                             "batches)");
    const int c = ((chunk + 63) / 64) * 64;
    size_t need = 0;
    for (int p = 0; p < spec_.P(); ++p) need += (size_t)ceil_div(spec_.n[p], c) * 2 * spec_.dim * spec_.n[p];
    if (need > part_cap_) throw Error(kErrArg, "source chunk too small for the partial-sum buffer");
    chunk_ = c;
    rkv_ = kv_lanes;
    rbwd_ = bwd_lanes;
    build_descs();
    upload_descs();
}

// Autotuner: for each candidate source chunk (the work-item size) and, in
// FP32, each number of targets per lane (K.alpha 1, 2, 4; adjoint 1, 2),
// times the batched K.alpha and adjoint pair launches of the first time
// slice on this batch, with the culling state of the run, and keeps the
// chunk whose best K.alpha plus best adjoint time is smallest, with those
// two lane counts.  Results are cached per batch signature and culling
// state for the life of the process.
void Batch::autotune()
{
    using Key = std::tuple<int, int, int, int, long long, long long, int>;
    static std::mutex mu;
    static std::map<Key, std::array<int, 3>> cache;
    long long sumsq = 0;
    for (int n : spec_.n) sumsq += (long long)n * n;
    const int cull_key = 2 * pair_cull_bits() + (pair_groups() ? 1 : 0);
    const Key key = std::make_tuple(spec_.dim, spec_.order, spec_.prec, spec_.P(), ntot_, sumsq, cull_key);
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = cache.find(key);
        if (it != cache.end()) {
            chunk_ = it->second[0];
            rkv_ = it->second[1];
            rbwd_ = it->second[2];
            build_descs();
            upload_descs();
            return;
        }
    }
    const int cands[] = {256, 512, 1024, 2048, 4096};
    const int P = spec_.P(), D = spec_.dim;
    const bool f32 = spec_.prec == kFP32;
    const int g = f32 ? 1 : 0;
    std::vector<int> all(P, spec_.prec);
    DArr<double> alp;
    alp.ensure((size_t)vtot_);
    // nonzero stand-in momenta and adjoints: launch times under load depend
    // on the operands (the clock follows the power drawn), so zeros mislead
    kb_fill_pattern<<<1024, kTile, 0, st_>>>(alp.p, vtot_, 1e-2);
    kb_fill_pattern<<<1024, kTile, 0, st_>>>(padj_, (long long)D * ntot_, 1.0);
    LDDMM_LAUNCH_CHECK();
    for (int p = 0; p < P; ++p)
        LDDMM_HIP(hipMemcpyAsync(nde_ + (size_t)spec_.T * D * vo_[p], q0_ + (size_t)D * vo_[p],
                                 (size_t)D * spec_.n[p] * sizeof(double), hipMemcpyDeviceToDevice, st_));
    BatchView v{desc_, nde_, padj_, part_, vpart_, alp.p, nullptr, q0_, tc_, tu_, ta_, elm_, ifo_, rows_,
                spec_.T, spec_.h, spec_.w, ecen_, edir_, evol_, dc_, dd_, efwd_, evf_, pair_cull_bits(),
                pair_expansion() ? 1 : 0, nullptr, grpv_, grpe_, grpt_,
                pair_groups() && spec_.order == -1 ? 1 : 0, 0};
    v.grpm = grpm_;
    hipEvent_t e0, e1;
    LDDMM_HIP(hipEventCreate(&e0));
    LDDMM_HIP(hipEventCreate(&e1));
    // the shortest of three single-launch times, after one untimed launch
    auto timed = [&](auto &&launch) {
        launch();
        float best_ms = 1e30f;
        for (int rep = 0; rep < 3; ++rep) {
            LDDMM_HIP(hipEventRecord(e0, st_));
            launch();
            LDDMM_HIP(hipEventRecord(e1, st_));
            LDDMM_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            LDDMM_HIP(hipEventElapsedTime(&ms, e0, e1));
            best_ms = std::min(best_ms, ms);
        }
        return best_ms;
    };
    const bool log = std::getenv("LDDMM_BATCH_TUNE_LOG") && std::atoi(std::getenv("LDDMM_BATCH_TUNE_LOG")) != 0;
    auto dispatch = [&](auto &&fn) {
        dispatch_order(spec_.order, [&](auto O) {
            if (D == 2) fn(std::integral_constant<int, 2>{}, O);
            else fn(std::integral_constant<int, 3>{}, O);
        });
    };
    auto each = [&](auto Pc) {
        float best = 1e30f;
        int best_c = chunk_, best_kv = rkv_, best_bwd = rbwd_;
        bool bounds = false;
        for (int c : cands) {
            chunk_ = c;
            build_descs();
            upload_descs();
            float t_kv = 1e30f, t_bwd = 1e30f;
            int r_kv = 1, r_bwd = 1;
            for (int r : {1, 2, 4}) {
                if (!f32 && r > 1) break;
                rkv_ = r;
                rbwd_ = r > 2 && !mx_ok() ? 2 : r;
                tab_key_.clear();
                build_tables(all);
                const Tables &tb = *tab_;
                if (v.gpts && f32 && !bounds) {   // group bounds of q_0 for the culling
                    if (D == 2) kb_group_bounds<2><<<tb.nvtile, kTile, 0, st_>>>(v, 0, 0, tb.vtile.p);
                    else kb_group_bounds<3><<<tb.nvtile, kTile, 0, st_>>>(v, 0, 0, tb.vtile.p);
                    LDDMM_LAUNCH_CHECK();
                    bounds = true;
                }
                if (r == 4 && mx_ok() && v.gpts && f32) {   // column scales of the matrix-core adjoint
                    kb_group_maxima<3><<<tb.nvtile, kTile, 0, st_>>>(v, 0, tb.vtile.p);
                    LDDMM_LAUNCH_CHECK();
                }
                const float kv = timed([&] {
                    dispatch([&](auto Dc, auto Oc) {
                        with_lanes(r, [&](auto Rc) {
                            if constexpr (decltype(Pc)::fp32 || Rc() == 1)
                                kb_kv<Dc(), Oc(), decltype(Pc), Rc()><<<tb.nkv[g], kBlock, 0, st_>>>(v, 0,
                                                                                                    tb.kv[g].p);
                        });
                    });
                    LDDMM_LAUNCH_CHECK();
                });
                if (kv < t_kv) {
                    t_kv = kv;
                    r_kv = r;
                }
                if (log) std::fprintf(stderr, "batch tune: chunk %d, K.alpha lanes %d: %.3f ms\n", c, r, kv);
                if (r > 2 && !mx_ok()) continue;
                const float bw = timed([&] {
                    dispatch([&](auto Dc, auto Oc) {
                        with_lanes(r, [&](auto Rc) {
                            if constexpr ((decltype(Pc)::fp32 && Rc() <= 2) || Rc() == 1)
                                kb_bwd<Dc(), Oc(), decltype(Pc), Rc()><<<tb.nbwd[g], kBlock, 0, st_>>>(v, 0,
                                                                                                      tb.bwd[g].p);
                            else if constexpr (decltype(Pc)::fp32 && Rc() == 4 && Dc() == 3 && Oc() == -1)
                                kb_bwd_mx<3><<<tb.nbwd[g], kBlock, 0, st_>>>(v, 0, tb.bwd[g].p);
                        });
                    });
                    LDDMM_LAUNCH_CHECK();
                });
                if (bw < t_bwd) {
                    t_bwd = bw;
                    r_bwd = r;
                }
                if (log) std::fprintf(stderr, "batch tune: chunk %d, adjoint lanes %d: %.3f ms\n", c, r, bw);
            }
            if (t_kv + t_bwd < best) {
                best = t_kv + t_bwd;
                best_c = c;
                best_kv = f32 ? r_kv : kKvR32;
                best_bwd = f32 ? r_bwd : bwd_r32(spec_.order);
            }
        }
        chunk_ = best_c;
        rkv_ = best_kv;
        rbwd_ = best_bwd;
    };
    if (f32) each(PrecF32{});
    else each(PrecF64{});
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    build_descs();
    upload_descs();
    std::lock_guard<std::mutex> gl(mu);
    cache[key] = {chunk_, rkv_, rbwd_};
}

std::vector<double> Batch::to_device_layout(const double *h) const
{
    const int D = spec_.dim, T1 = spec_.T - 1;
    std::vector<double> out((size_t)vtot_);
    for (int p = 0; p < spec_.P(); ++p) {
        const int n = spec_.n[p];
        const size_t off = (size_t)D * T1 * vo_[p];
        for (int t = 0; t < T1; ++t)
            for (int i = 0; i < n; ++i)
                for (int c = 0; c < D; ++c)
                    out[off + (size_t)t * D * n + (size_t)c * n + i] =
                        h[off + (size_t)t * D * n + (size_t)vperm_[vo_[p] + i] * D + c];
    }
    return out;
}

std::vector<double> Batch::vec_to_device_layout(const double *h) const
{
    const int D = spec_.dim;
    std::vector<double> out((size_t)D * ntot_);
    for (int p = 0; p < spec_.P(); ++p) {
        const int n = spec_.n[p];
        const size_t off = (size_t)D * vo_[p];
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < D; ++c) out[off + (size_t)c * n + i] = h[off + (size_t)vperm_[vo_[p] + i] * D + c];
    }
    return out;
}

void Batch::const_momentum(const double *u, const double *s, double *x)
{
    if (spec_.P() > 65535) throw Error(kErrArg, "a constant-in-time start takes at most 65535 problems");
    kb_const_momentum<<<dim3(32, spec_.P()), kTile, 0, st_>>>(desc_, spec_.dim, spec_.T - 1, u, s, x);
    LDDMM_LAUNCH_CHECK();
}

void Batch::to_host_layout(const std::vector<double> &dv, double *h) const
{
    const int D = spec_.dim, T1 = spec_.T - 1;
    for (int p = 0; p < spec_.P(); ++p) {
        const int n = spec_.n[p];
        const size_t off = (size_t)D * T1 * vo_[p];
        for (int t = 0; t < T1; ++t)
            for (int i = 0; i < n; ++i)
                for (int c = 0; c < D; ++c)
                    h[off + (size_t)t * D * n + (size_t)vperm_[vo_[p] + i] * D + c] =
                        dv[off + (size_t)t * D * n + (size_t)c * n + i];
    }
}

// ======================================================================
// Batched L-BFGS: per-problem state machines advanced in rounds; every
// round evaluates the current trial point of every active problem in one
// batched evaluation.  The per-problem arithmetic is that of run_lbfgs
// (Cpp/src/LBFGS.cpp, getDirection.cpp, lineSearch.cpp).
// ======================================================================
BatchRunResult Batch::run(double *x, const BatchLbfgsOptions &bo)
{
    auto t0 = std::chrono::steady_clock::now();
    const LbfgsOptions &o = bo.base;
    const int P = spec_.P(), m = o.mem;
    if (m < 1) throw Error(kErrArg, "memNum must be at least 1");
    const long long V = vtot_;

    // vector block table (fixed: per problem blocks of kVecBlk)
    std::vector<int2> blk;
    std::vector<int> first(P), count(P);
    std::vector<long long> voff(P), vn(P);
    for (int p = 0; p < P; ++p) {
        voff[p] = var_offset(p);
        vn[p] = varnum(p);
        first[p] = (int)blk.size();
        count[p] = ceil_div(vn[p], kVecBlk);
        for (int b = 0; b < count[p]; ++b) blk.push_back(make_int2(p, b));
    }
    const int nent = (int)blk.size();
    Slab s;
    double *g, *d, *xt, *gt, *S, *Y, *sc, *SYt, *YYt, *ab, *coef, *rdots, *part, *rel_d;
    int2 *blk_d;
    int *first_d, *count_d, *imask;
    long long *voff_d, *vn_d;
    // scalar block: f, gg, gd, dd  (P each)
    enum { F = 0, GG, GD, DDs, NS };
    s.reserve(&g, V);
    s.reserve(&d, V);
    s.reserve(&xt, V);
    s.reserve(&gt, V);
    s.reserve(&S, (size_t)V * m);
    s.reserve(&Y, (size_t)V * m);
    s.reserve(&sc, (size_t)NS * P);
    // L-BFGS tables: SY[p][i][j] = s_i.y_j, YY[p][i][j] = y_i.y_j (ring slots),
    // the recursion's alpha/beta, the direction's coefficients, and the
    // batched dots of the direction (2m) and of an accepted pair (3m)
    s.reserve(&SYt, (size_t)P * m * m);
    s.reserve(&YYt, (size_t)P * m * m);
    s.reserve(&ab, (size_t)P * 2 * m);
    s.reserve(&coef, (size_t)P * (1 + 2 * m));
    s.reserve(&rdots, (size_t)P * 3 * m);
    s.reserve(&part, (size_t)kMaxBDots * nent);
    s.reserve(&rel_d, P);
    s.reserve(&blk_d, nent);
    s.reserve(&first_d, P);
    s.reserve(&count_d, P);
    s.reserve(&voff_d, P);
    s.reserve(&vn_d, P);
    // int mask arrays: dir, trial, accept, newest slot, filled slots, then
    // per ring slot j an array holding j where slot j is filled, else -1
    const int NM = 5 + m;
    s.reserve(&imask, (size_t)NM * P);
    s.allocate();
    LDDMM_HIP(hipMemcpyAsync(blk_d, blk.data(), blk.size() * sizeof(int2), hipMemcpyHostToDevice, st_));
    LDDMM_HIP(hipMemcpyAsync(first_d, first.data(), P * sizeof(int), hipMemcpyHostToDevice, st_));
    LDDMM_HIP(hipMemcpyAsync(count_d, count.data(), P * sizeof(int), hipMemcpyHostToDevice, st_));
    LDDMM_HIP(hipMemcpyAsync(voff_d, voff.data(), P * sizeof(long long), hipMemcpyHostToDevice, st_));
    LDDMM_HIP(hipMemcpyAsync(vn_d, vn.data(), P * sizeof(long long), hipMemcpyHostToDevice, st_));
    double *hs = nullptr;
    int *hm = nullptr;
    double *hrel = nullptr;
    LDDMM_HIP(hipHostMalloc((void **)&hs, sizeof(double) * NS * P, hipHostMallocDefault));
    LDDMM_HIP(hipHostMalloc((void **)&hm, sizeof(int) * NM * P, hipHostMallocDefault));
    LDDMM_HIP(hipHostMalloc((void **)&hrel, sizeof(double) * P, hipHostMallocDefault));
    struct HostFree {
        void *a, *b, *c;
        ~HostFree()
        {
            if (a) (void)hipHostFree(a);
            if (b) (void)hipHostFree(b);
            if (c) (void)hipHostFree(c);
        }
    } hf{hs, hm, hrel};
    int *m_dir = imask, *m_trial = imask + P, *m_acc = imask + 2 * P, *m_new = imask + 3 * P,
        *m_his = imask + 4 * P, *m_ring = imask + 5 * P;   // m arrays of P
    int *h_dir = hm, *h_trial = hm + P, *h_acc = hm + 2 * P, *h_new = hm + 3 * P, *h_his = hm + 4 * P,
        *h_ring = hm + 5 * P;
    auto set_ring = [&](int p, int filled) {
        h_his[p] = filled;
        for (int j = 0; j < m; ++j) h_ring[j * P + p] = j < filled ? j : -1;
    };
    VecTab vt{blk_d, voff_d, vn_d};

    // synchronous: the pinned host masks are rewritten right after
    auto upload_masks = [&] {
        LDDMM_HIP(hipMemcpyAsync(imask, hm, sizeof(int) * NM * P, hipMemcpyHostToDevice, st_));
        LDDMM_HIP(hipStreamSynchronize(st_));
    };
    auto vec = [&](VecArgs a) {
        a.vtot = V;
        ProfScope ps(kProfLbfgsVector, st_);
        kb_vec<<<nent, kBlock, 0, st_>>>(a, vt);
        LDDMM_LAUNCH_CHECK();
    };
    // nd dot products per problem into out[k*P + p], kMaxBDots per launch pair
    auto bdots = [&](int nd, const VRef *xs, const VRef *ys, const int *mask, double *out) {
        for (int k0 = 0; k0 < nd; k0 += kMaxBDots) {
            const int nk = std::min(kMaxBDots, nd - k0);
            BDotArgs a{};
            for (int k = 0; k < nk; ++k) {
                a.x[k] = xs[k0 + k];
                a.y[k] = ys[k0 + k];
            }
            ProfScope ps(kProfLbfgsDots, st_);
            kb_dots1_all<<<nent, kBlock, (size_t)nk * kBlock * sizeof(double), st_>>>(a, nk, vt, mask, part, nent);
            LDDMM_LAUNCH_CHECK();
            kb_dots2<<<dim3(P, nk), kBlock, 0, st_>>>(part, nent, first_d, count_d, mask, out + (size_t)k0 * P, P);
            LDDMM_LAUNCH_CHECK();
        }
    };
    auto fetch = [&](int slot0, int nslots) {
        LDDMM_HIP(hipMemcpyAsync(hs + (size_t)slot0 * P, sc + (size_t)slot0 * P, sizeof(double) * nslots * P,
                                 hipMemcpyDeviceToHost, st_));
        LDDMM_HIP(hipStreamSynchronize(st_));
    };
    auto plain = [](const double *p) { return VRef{p, nullptr, 0}; };

    // per-problem state
    struct St {
        bool active = true, need_dir = true;
        int it = 0, new_idx = -1, status = 0, evals = 0, bis = 0, epd = 0, bis_max = 0, epd_max = 0, cnt = 0;
        int prec = kFP64, fp32_iters = 0, reached = 0;
        double f = 0, fprev = 0, glen = 0, slope = 0, dlen = 0, lo = 0, hi = DBL_MAX, rel = 1;
    };
    std::vector<St> st(P);
    const int run_prec = bo.prec >= 0 ? bo.prec : spec_.prec;
    if (run_prec != kFP64 && run_prec != kFP32) throw Error(kErrArg, "unknown precision policy");
    const bool can_refine = o.fp64_refine && run_prec == kFP32;
    std::vector<int> prec_of(P, run_prec);

    // f, g at the start
    eval(x, g, sc + (size_t)F * P, prec_of);
    {
        VRef xs[1] = {plain(g)}, ys[1] = {plain(g)};
        bdots(1, xs, ys, nullptr, sc + (size_t)GG * P);
    }
    fetch(F, 2);
    for (int p = 0; p < P; ++p) {
        st[p].f = hs[F * P + p];
        st[p].glen = std::sqrt(hs[GG * P + p]);
        st[p].evals = 1;
        st[p].prec = run_prec;
        // Cpp/src/LBFGS.cpp:66 (a NaN gradient norm goes on to the direction, where it stops with code 1)
        if (st[p].glen < o.tol || o.itr_max == 0) st[p].active = false;
    }

    auto has_target = [&](int p) { return p < (int)bo.targets.size() || o.has_target; };
    auto target_of = [&](int p) { return p < (int)bo.targets.size() ? bo.targets[p] : o.target; };

    // per-problem refine to FP64 (re-evaluate the current iterate)
    auto refine = [&](const std::vector<int> &which) {
        std::vector<int> pr(P, -1);
        bool any = false;
        for (int p : which) {
            st[p].prec = kFP64;
            pr[p] = kFP64;
            any = true;
        }
        if (!any) return;
        eval(x, g, sc + (size_t)F * P, pr);
        for (int p = 0; p < P; ++p) h_acc[p] = pr[p] >= 0;
        upload_masks();
        VRef xs[1] = {plain(g)}, ys[1] = {plain(g)};
        bdots(1, xs, ys, m_acc, sc + (size_t)GG * P);
        fetch(F, 2);
        for (int p : which) {
            st[p].f = hs[F * P + p];
            st[p].glen = std::sqrt(hs[GG * P + p]);
            st[p].evals += 1;
            st[p].need_dir = true;
            st[p].lo = 0;
            st[p].hi = DBL_MAX;
            st[p].rel = 1;
            st[p].cnt = 0;
        }
    };

    int rounds = 0;
    while (true) {
        int nactive = 0;
        for (int p = 0; p < P; ++p) nactive += st[p].active ? 1 : 0;
        if (!nactive) break;
        ++rounds;
        if (bo.progress > 0 && rounds % bo.progress == 0) {
            const double el =
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            std::fprintf(stderr, "batch round %d: %d of %d problems active, %.1f s\n", rounds, nactive, P, el);
            std::fflush(stderr);
        }

        // ---- directions for the problems starting an iteration
        std::vector<int> dirp;
        for (int p = 0; p < P; ++p)
            if (st[p].active && st[p].need_dir) dirp.push_back(p);
        if (!dirp.empty()) {
            std::memset(hm, 0, sizeof(int) * NM * P);
            for (int j = 0; j < m; ++j)
                for (int p = 0; p < P; ++p) h_ring[j * P + p] = -1;
            for (int p : dirp) {
                h_dir[p] = 1;
                h_new[p] = st[p].new_idx;
                set_ring(p, std::min(st[p].it, m));   // pairs stored so far
            }
            upload_masks();
            {
                // a_j = s_j.g and b_j = y_j.g for the filled slots, one batched pass
                std::vector<VRef> xs(2 * m), ys(2 * m, plain(g));
                for (int j = 0; j < m; ++j) {
                    xs[j] = VRef{S, m_ring + (size_t)j * P, V};
                    xs[m + j] = VRef{Y, m_ring + (size_t)j * P, V};
                }
                bdots(2 * m, xs.data(), ys.data(), m_dir, rdots);
                ProfScope ps(kProfLbfgsVector, st_);
                kb_lbfgs_coef<<<ceil_div(P, 64), 64, 0, st_>>>(rdots, SYt, YYt, m_dir, m_his, m_new, ab, coef, m, P);
                LDDMM_LAUNCH_CHECK();
                kb_dir<<<nent, kBlock, 0, st_>>>(d, g, S, Y, V, coef, m_dir, m_his, m, vt);
                LDDMM_LAUNCH_CHECK();
            }
            {
                VRef xs[2] = {plain(g), plain(d)}, ys[2] = {plain(d), plain(d)};
                bdots(2, xs, ys, m_dir, sc + (size_t)GD * P);
            }
            fetch(GD, 2);
            std::vector<int> stalled;
            for (int p : dirp) {
                St &q = st[p];
                q.slope = hs[GD * P + p];
                q.need_dir = false;
                if (!(q.slope <= -DBL_MIN)) {   // a NaN slope is not a descent direction either
                    if (can_refine && q.prec == kFP32) {
                        stalled.push_back(p);
                        continue;
                    }
                    q.status = 1;
                    q.active = false;
                    continue;
                }
                q.dlen = std::sqrt(hs[DDs * P + p]);
                q.bis_max = (int)std::fmax(50.0, std::log2(1e5 * q.dlen));
                q.epd_max = (int)std::fmax(10.0, std::log2(1e5 / q.dlen));
                q.lo = 0.0;
                q.hi = DBL_MAX;
                q.rel = 1.0;
                q.bis = q.epd = q.cnt = 0;
            }
            refine(stalled);
            for (int p : stalled)
                if (has_target(p) && st[p].f <= target_of(p)) {
                    st[p].reached = 1;
                    st[p].active = false;
                }
            if (!stalled.empty()) continue;
        }

        // ---- one trial point per active problem, one batched evaluation
        std::vector<int> pr(P, -1);
        bool anyt = false;
        for (int p = 0; p < P; ++p) {
            h_trial[p] = st[p].active && !st[p].need_dir;
            hrel[p] = st[p].rel;
            if (h_trial[p]) {
                pr[p] = st[p].prec;
                anyt = true;
            }
        }
        if (!anyt) continue;
        upload_masks();
        LDDMM_HIP(hipMemcpyAsync(rel_d, hrel, sizeof(double) * P, hipMemcpyHostToDevice, st_));
        {
            VecArgs a{};
            a.op = kTry;
            a.d = xt;
            a.x = x;
            a.g = d;
            a.rel = rel_d;
            a.mask = m_trial;
            vec(a);
        }
        eval(xt, gt, sc + (size_t)F * P, pr);
        {
            VRef xs[2] = {plain(gt), plain(gt)}, ys[2] = {plain(gt), plain(d)};
            bdots(2, xs, ys, m_trial, sc + (size_t)GG * P);
        }
        fetch(F, 3);
        std::vector<int> acc, stalled;
        for (int p = 0; p < P; ++p) {
            if (!h_trial[p]) continue;
            St &q = st[p];
            const double fn = hs[F * P + p], gn = std::sqrt(hs[GG * P + p]), sl = hs[GD * P + p];
            ++q.cnt;
            ++q.evals;
            bool ok = false;
            if (gn < o.tol) ok = true;
            else if (fn >= q.f + o.wolfe1 * q.rel * q.slope || std::isnan(fn) || std::isnan(sl))
                q.hi = q.rel;   // a NaN energy or slope shrinks the step like a failed Armijo test
            else if (sl <= o.wolfe2 * q.slope) q.lo = q.rel;
            else ok = true;
            if (ok) {
                acc.push_back(p);
                q.it += 1;
                const double fprev = q.f;
                q.f = fn;
                q.glen = gn;
                if (q.prec == kFP32) ++q.fp32_iters;
                q.need_dir = true;
                // stop rules (checked after the accept below)
                q.reached = 0;
                if (has_target(p) && q.f <= target_of(p)) q.reached = 1;
                q.fprev = fprev;
                continue;
            }
            if (q.hi == DBL_MAX) {
                q.rel *= 2.0;
                ++q.epd;
            } else {
                q.rel = 0.5 * (q.lo + q.hi);
                ++q.bis;
            }
            if (!(q.bis <= q.bis_max && q.epd <= q.epd_max)) {
                if (can_refine && q.prec == kFP32) {
                    stalled.push_back(p);
                } else {
                    q.status = q.hi == DBL_MAX ? 2 : 3;
                    q.active = false;
                }
            }
        }
        if (!acc.empty()) {
            for (int p = 0; p < P; ++p) h_acc[p] = 0;
            for (int p : acc) {
                St &q = st[p];
                q.new_idx = (q.new_idx == m - 1) ? 0 : q.new_idx + 1;
                h_acc[p] = 1;
                h_new[p] = q.new_idx;
                set_ring(p, std::min(q.it, m));   // the accepted pair included
            }
            upload_masks();
            VecArgs a{};
            a.op = kAccept;
            a.xw = x;
            a.gw = g;
            a.xt = xt;
            a.gt = gt;
            a.Sw = S;
            a.Yw = Y;
            a.mask = m_acc;
            a.slot = m_new;
            vec(a);
            {
                // the new pair against every filled slot: s_n.y_j, s_j.y_n, y_n.y_j
                std::vector<VRef> xs(3 * m), ys(3 * m);
                for (int j = 0; j < m; ++j) {
                    const VRef sj{S, m_ring + (size_t)j * P, V}, yj{Y, m_ring + (size_t)j * P, V};
                    xs[3 * j] = VRef{S, m_new, V};
                    ys[3 * j] = yj;
                    xs[3 * j + 1] = sj;
                    ys[3 * j + 1] = VRef{Y, m_new, V};
                    xs[3 * j + 2] = VRef{Y, m_new, V};
                    ys[3 * j + 2] = yj;
                }
                bdots(3 * m, xs.data(), ys.data(), m_acc, rdots);
                kb_store_ring<<<ceil_div(P, 64), 64, 0, st_>>>(rdots, m_acc, m_new, m_his, SYt, YYt, m, P);
                LDDMM_LAUNCH_CHECK();
            }
            for (int p : acc) {
                St &q = st[p];
                const double fprev = q.fprev;
                q.rel = 1.0;
                if (q.reached) {
                    // a target met in FP32 arithmetic is confirmed in FP64 before the run stops
                    if (can_refine && q.prec == kFP32) {
                        q.reached = 0;
                        stalled.push_back(p);
                    } else {
                        q.active = false;
                    }
                    continue;
                }
                if (q.it >= o.itr_max || q.glen < o.tol) {
                    q.active = false;
                    continue;
                }
                const bool small = (fprev - q.f) < o.refine_rel_tol * std::fabs(q.f);
                if (can_refine && q.prec == kFP32 && small) {
                    stalled.push_back(p);
                    continue;
                }
                if (o.rel_tol > 0.0 && (fprev - q.f) < o.rel_tol * std::fabs(q.f)) q.active = false;
            }
        }
        refine(stalled);
        for (int p : stalled) {
            if (has_target(p) && st[p].f <= target_of(p)) {
                st[p].reached = 1;
                st[p].active = false;
            } else if (st[p].it >= o.itr_max) {
                // the FP64 re-evaluation does not grant an iteration beyond the cap
                st[p].active = false;
            }
        }
    }
    LDDMM_HIP(hipStreamSynchronize(st_));
    BatchRunResult r;
    r.rounds = rounds;
    r.prob.resize(P);
    for (int p = 0; p < P; ++p) {
        r.prob[p].f = st[p].f;
        r.prob[p].gnorm = st[p].glen;
        r.prob[p].iters = st[p].it;
        r.prob[p].evals = st[p].evals;
        r.prob[p].status = st[p].status;
        r.prob[p].fp32_iters = st[p].fp32_iters;
        r.prob[p].reached_target = st[p].reached;
    }
    r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return r;
}

}  // namespace lddmm
