// vector_kernels.hip -- deterministic FP64 reductions and the HBM-bound
// vector updates of the device-resident L-BFGS (Cpp/src/LBFGS.cpp:25-112,
// Cpp/src/getDirection.cpp:9-66, Cpp/src/lineSearch.cpp:18-24).
//
// Every inner product is a two-stage reduction over a fixed grid of
// kDotBlocks blocks with a fixed element-to-thread map and fixed tree
// shapes, so results are bitwise repeatable run to run (the reference's
// cublasDdot / ddot order is not specified).  Scalars stay in device
// memory; the two-loop coefficients are formed on the device from them, so
// a direction costs no host round trip.
#include <algorithm>

#include "kernels.hpp"
#include "vector_ops.hpp"

namespace lddmm {

namespace {

__device__ __forceinline__ double block_sum(double v, double *red)
{
    red[threadIdx.x] = v;
    __syncthreads();
#pragma unroll
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    return red[0];
}

struct DotArgs {
    const double *x[kMaxDots];
    const double *y[kMaxDots];
};

__global__ __launch_bounds__(kBlock) void k_dots_stage1(DotArgs args, long long n, double *__restrict__ scratch)
{
    __shared__ double red[kBlock];
    const int k = blockIdx.y;
    const double *__restrict__ x = args.x[k];
    const double *__restrict__ y = args.y[k];
    double s = 0.0;
    const long long stride = (long long)kDotBlocks * kBlock;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) s += x[i] * y[i];
    double t = block_sum(s, red);
    if (threadIdx.x == 0) scratch[(size_t)k * kDotBlocks + blockIdx.x] = t;
}

__global__ __launch_bounds__(kBlock) void k_sum_stage1(const double *__restrict__ x, long long n,
                                                       double *__restrict__ scratch)
{
    __shared__ double red[kBlock];
    double s = 0.0;
    const long long stride = (long long)kDotBlocks * kBlock;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) s += x[i];
    double t = block_sum(s, red);
    if (threadIdx.x == 0) scratch[blockIdx.x] = t;
}

// one block per result: sums kDotBlocks partials in a fixed order
__global__ __launch_bounds__(kBlock) void k_stage2(const double *__restrict__ scratch, double *__restrict__ out)
{
    __shared__ double red[kBlock];
    const double *p = scratch + (size_t)blockIdx.x * kDotBlocks;
    double s = 0.0;
    for (int i = threadIdx.x; i < kDotBlocks; i += kBlock) s += p[i];
    double t = block_sum(s, red);
    if (threadIdx.x == 0) out[blockIdx.x] = t;
}

__global__ __launch_bounds__(kBlock) void k_energy(const double *__restrict__ fp, int nf,
                                                   const double *__restrict__ vp, int nv,
                                                   const double *__restrict__ tt, double h, double w,
                                                   double *__restrict__ out)
{
    __shared__ double red[kBlock];
    double s = 0.0;
    for (int i = threadIdx.x; i < nf; i += kBlock) s += fp[i];
    double ef = block_sum(s, red);
    __syncthreads();
    s = 0.0;
    for (int i = threadIdx.x; i < nv; i += kBlock) s += vp[i];
    double ev = block_sum(s, red);
    if (threadIdx.x == 0) out[0] = h * 0.5 * ef + w * (ev + tt[0]);
}

// l2Dist end point (MATLAB/src/l2Dist.m:1-5): p = -w (q - tl) per vertex,
// per-block partials of |q - tl|^2 / 2 for the energy kernel (which
// multiplies them by w)
template <int D>
__global__ __launch_bounds__(kBlock) void k_l2_match(const double *__restrict__ q, const double *__restrict__ tl,
                                                     int n, double w, double *__restrict__ p,
                                                     double *__restrict__ parts)
{
    __shared__ double red[kBlock];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double s = 0.0;
    if (i < n) {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            const double d = q[(size_t)c * n + i] - tl[(size_t)c * n + i];
            p[(size_t)c * n + i] = -w * d;
            s += d * d;
        }
    }
    const double t = block_sum(s, red);
    if (threadIdx.x == 0) parts[blockIdx.x] = 0.5 * t;
}

// ---------------------------------------------------------------- L-BFGS
// Compact form of the two-loop recursion (Cpp/src/getDirection.cpp:9-66).
// With a_i = s_i.g, b_i = y_i.g and the tables SY[i][j] = s_i.y_j,
// YY[i][j] = y_i.y_j of the stored pairs (slot-indexed, m x m), the two
// loops are
//   newest to oldest:  alpha_i = (a_i - sum_{q newer} alpha_q SY[i][q]) / SY[i][i]
//   oldest to newest:  beta_i  = (gamma (b_i - sum_q alpha_q YY[i][q])
//                                 + sum_{q older} (alpha_q - beta_q) SY[q][i]) / SY[i][i]
// (the same alpha and beta: s_i.q and y_i.r re-associated), gamma =
// SY[n][n] / YY[n][n] of the newest pair n (Cpp/src/LBFGS.cpp:69-81), and
// the direction is one pass
//   d = -gamma g + sum_j gamma alpha_j y_j - sum_j (alpha_j - beta_j) s_j.
// Slots 0..h-1 hold the h stored pairs; slot(k) = (newest - k) mod m is the
// k-th newest.

// Ring products: product k = y0 + blockIdx.y is combination c = k / h at
// slot j = k % h, <rings[c] + j nv, vecs[c]>; nb blocks per product
struct RingDotArgs {
    const double *ring[3];
    const double *vec[3];
};

__global__ __launch_bounds__(kBlock) void k_ring_dots_stage1(RingDotArgs args, int h, int y0, long long nv, int nb,
                                                             double *__restrict__ scratch)
{
    __shared__ double red[kBlock];
    const int k = y0 + blockIdx.y, c = k / h, j = k - c * h;
    const double *__restrict__ x = args.ring[c] + (size_t)j * nv;
    const double *__restrict__ y = args.vec[c];
    double s = 0.0;
    const long long stride = (long long)nb * kBlock;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < nv; i += stride) s += x[i] * y[i];
    const double t = block_sum(s, red);
    if (threadIdx.x == 0) scratch[(size_t)k * nb + blockIdx.x] = t;
}

// one block per result: sums nb partials in a fixed order
__global__ __launch_bounds__(kBlock) void k_stage2n(const double *__restrict__ scratch, int nb, int k0,
                                                    double *__restrict__ out)
{
    __shared__ double red[kBlock];
    const int k = k0 + blockIdx.x;
    const double *p = scratch + (size_t)k * nb;
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += kBlock) s += p[i];
    const double t = block_sum(s, red);
    if (threadIdx.x == 0) out[k] = t;
}

// One block: the recursion coefficients from ab = [a_0..a_{h-1}, b_0..b_{h-1}]
// (by slot).  work holds 3m doubles (alpha, beta, running sums); coef =
// [g, y slots 0..m-1, s slots 0..m-1].  Each step of a loop is a column
// update of the remaining sums, one barrier apart.
__global__ __launch_bounds__(kBlock) void k_lbfgs_coef(const double *__restrict__ ab, const double *__restrict__ SY,
                                                       const double *__restrict__ YY, int m, int h, int newest,
                                                       double *__restrict__ work, double *__restrict__ coef)
{
    const int tid = threadIdx.x;
    double *al = work, *be = work + m, *t = work + 2 * m;
    for (int j = tid; j < 1 + 2 * m; j += kBlock) coef[j] = 0.0;
    if (h == 0) {   // first iteration: d = -g
        __syncthreads();
        if (tid == 0) coef[0] = -1.0;
        return;
    }
    auto slot = [&](int k) {
        const int i = newest - k;
        return i < 0 ? i + m : i;
    };
    for (int i = tid; i < h; i += kBlock) t[i] = ab[i];
    __syncthreads();
    for (int k = 0; k < h; ++k) {
        const int i = slot(k);
        const double a = t[i] / SY[(size_t)i * m + i];
        for (int kk = k + 1 + tid; kk < h; kk += kBlock) {
            const int ii = slot(kk);
            t[ii] -= a * SY[(size_t)ii * m + i];
        }
        if (tid == 0) al[i] = a;
        __syncthreads();
    }
    const double gam = SY[(size_t)newest * m + newest] / YY[(size_t)newest * m + newest];
    for (int i = tid; i < h; i += kBlock) {
        double s = ab[h + i];
        for (int q = 0; q < h; ++q) s -= al[q] * YY[(size_t)i * m + q];
        t[i] = gam * s;
    }
    __syncthreads();
    for (int k = h - 1; k >= 0; --k) {
        const int i = slot(k);
        const double b = t[i] / SY[(size_t)i * m + i], c = al[i] - b;
        for (int kk = tid; kk < k; kk += kBlock) {
            const int ii = slot(kk);
            t[ii] += c * SY[(size_t)i * m + ii];
        }
        if (tid == 0) be[i] = b;
        __syncthreads();
    }
    if (tid == 0) coef[0] = -gam;
    for (int j = tid; j < h; j += kBlock) {
        coef[1 + j] = gam * al[j];
        coef[1 + m + j] = -(al[j] - be[j]);
    }
}

// d = coef_g g + sum_{j<h} (coef_y[j] Y_j + coef_s[j] S_j), and the first
// stage of g.d and d.d over the fixed kDotBlocks grid
__global__ __launch_bounds__(kBlock) void k_lbfgs_dir(double *__restrict__ d, const double *__restrict__ g,
                                                      const double *__restrict__ S, const double *__restrict__ Y,
                                                      long long nv, const double *__restrict__ coef, int h, int m,
                                                      double *__restrict__ scratch)
{
    __shared__ double red[kBlock];
    const double c0 = coef[0];
    double sgd = 0.0, sdd = 0.0;
    const long long stride = (long long)kDotBlocks * kBlock;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < nv; i += stride) {
        const double gi = g[i];
        double v = c0 * gi;
        for (int j = 0; j < h; ++j) v += coef[1 + j] * Y[(size_t)j * nv + i] + coef[1 + m + j] * S[(size_t)j * nv + i];
        d[i] = v;
        sgd += gi * v;
        sdd += v * v;
    }
    const double t1 = block_sum(sgd, red);
    __syncthreads();
    const double t2 = block_sum(sdd, red);
    if (threadIdx.x == 0) {
        scratch[blockIdx.x] = t1;
        scratch[kDotBlocks + blockIdx.x] = t2;
    }
}

// the newest pair's row and column of the tables: dots = [y_j.s_n | s_j.y_n |
// y_j.y_n] for the slots j < h (n = newest)
__global__ void k_lbfgs_store(const double *__restrict__ dots, int h, int n, int m, double *__restrict__ SY,
                              double *__restrict__ YY)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= h) return;
    SY[(size_t)n * m + j] = dots[j];
    SY[(size_t)j * m + n] = dots[h + j];
    YY[(size_t)n * m + j] = YY[(size_t)j * m + n] = dots[2 * h + j];
}

__global__ void k_try(double *__restrict__ xt, const double *__restrict__ x, double rel,
                      const double *__restrict__ d, long long n)
{
    long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) xt[i] = x[i] + rel * d[i];
}

// accept the trial point: s = xt - x, y = gt - g, x = xt, g = gt
__global__ void k_accept(double *__restrict__ x, double *__restrict__ g, const double *__restrict__ xt,
                         const double *__restrict__ gt, double *__restrict__ s, double *__restrict__ y,
                         long long n)
{
    long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        double xn = xt[i], gn = gt[i];
        s[i] = xn - x[i];
        y[i] = gn - g[i];
        x[i] = xn;
        g[i] = gn;
    }
}

inline int nblk(long long n) { return (int)((n + kBlock - 1) / kBlock); }

// rows x n permutation between the caller's and the internal vertex order
__global__ void k_permute(const double *__restrict__ src, double *__restrict__ dst, const int *__restrict__ perm,
                          int n, bool gather)
{
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const size_t row = (size_t)blockIdx.y * n;
    const int v = perm[s];
    if (gather) dst[row + s] = src[row + v];
    else dst[row + v] = src[row + s];
}

}  // namespace

void launch_dots(int nd, const double *const *xs, const double *const *ys, long long n, double *scratch,
                 double *out, hipStream_t st)
{
    if (nd <= 0) return;
    if (nd > kMaxDots) throw Error(kErrArg, "too many dots in one launch");
    DotArgs args{};
    for (int k = 0; k < nd; ++k) {
        args.x[k] = xs[k];
        args.y[k] = ys[k];
    }
    k_dots_stage1<<<dim3(kDotBlocks, nd), kBlock, 0, st>>>(args, n, scratch);
    LDDMM_LAUNCH_CHECK();
    k_stage2<<<nd, kBlock, 0, st>>>(scratch, out);
    LDDMM_LAUNCH_CHECK();
}

void launch_sum(const double *x, long long n, double *scratch, double *out, hipStream_t st)
{
    k_sum_stage1<<<kDotBlocks, kBlock, 0, st>>>(x, n, scratch);
    LDDMM_LAUNCH_CHECK();
    k_stage2<<<1, kBlock, 0, st>>>(scratch, out);
    LDDMM_LAUNCH_CHECK();
}

void launch_energy(const double *fwd_parts, int nf, const double *vf_parts, int nv, const double *tt, double h,
                   double w, double *out, hipStream_t st)
{
    k_energy<<<1, kBlock, 0, st>>>(fwd_parts, nf, vf_parts, nv, tt, h, w, out);
    LDDMM_LAUNCH_CHECK();
}

void launch_permute(const double *src, double *dst, const int *perm, int n, int rows, bool gather, hipStream_t st)
{
    for (int r0 = 0; r0 < rows && n > 0; r0 += 65535) {
        const int nr = std::min(rows - r0, 65535);
        const size_t off = (size_t)r0 * n;
        k_permute<<<dim3(ceil_div(n, kBlock), nr), kBlock, 0, st>>>(src + off, dst + off, perm, n, gather);
        LDDMM_LAUNCH_CHECK();
    }
}

int l2_match_blocks(int n) { return ceil_div(n, kBlock); }

void launch_l2_match(int dim, const double *q, const double *tl, int n, double w, double *p, double *parts,
                     hipStream_t st)
{
    if (dim == 2) k_l2_match<2><<<l2_match_blocks(n), kBlock, 0, st>>>(q, tl, n, w, p, parts);
    else k_l2_match<3><<<l2_match_blocks(n), kBlock, 0, st>>>(q, tl, n, w, p, parts);
    LDDMM_LAUNCH_CHECK();
}

int ring_dot_blocks(int nd) { return nd <= 16 ? kDotBlocks : std::max(16, 16 * kDotBlocks / nd); }

size_t ring_dot_scratch(int nd) { return (size_t)std::max(nd, 1) * ring_dot_blocks(std::max(nd, 1)); }

void launch_ring_dots(int ncombo, const double *const *rings, const double *const *vecs, int h, long long nv,
                      double *scratch, double *out, hipStream_t st)
{
    if (h <= 0 || ncombo <= 0) return;
    if (ncombo > 3) throw Error(kErrArg, "at most three ring products per launch");
    RingDotArgs a{};
    for (int c = 0; c < ncombo; ++c) {
        a.ring[c] = rings[c];
        a.vec[c] = vecs[c];
    }
    const int nd = ncombo * h, nb = ring_dot_blocks(nd);
    for (int y0 = 0; y0 < nd; y0 += 65535) {
        const int ny = std::min(nd - y0, 65535);
        k_ring_dots_stage1<<<dim3(nb, ny), kBlock, 0, st>>>(a, h, y0, nv, nb, scratch);
        LDDMM_LAUNCH_CHECK();
        k_stage2n<<<ny, kBlock, 0, st>>>(scratch, nb, y0, out);
        LDDMM_LAUNCH_CHECK();
    }
}

void launch_lbfgs_coef(const double *ab, const double *SY, const double *YY, int m, int h, int newest, double *work,
                       double *coef, hipStream_t st)
{
    k_lbfgs_coef<<<1, kBlock, 0, st>>>(ab, SY, YY, m, h, newest, work, coef);
    LDDMM_LAUNCH_CHECK();
}

void launch_lbfgs_dir(double *d, const double *g, const double *S, const double *Y, long long nv, const double *coef,
                      int h, int m, double *scratch, double *out, hipStream_t st)
{
    k_lbfgs_dir<<<kDotBlocks, kBlock, 0, st>>>(d, g, S, Y, nv, coef, h, m, scratch);
    LDDMM_LAUNCH_CHECK();
    k_stage2<<<2, kBlock, 0, st>>>(scratch, out);
    LDDMM_LAUNCH_CHECK();
}

void launch_lbfgs_store(const double *dots, int h, int newest, int m, double *SY, double *YY, hipStream_t st)
{
    if (h <= 0) return;
    k_lbfgs_store<<<ceil_div(h, 64), 64, 0, st>>>(dots, h, newest, m, SY, YY);
    LDDMM_LAUNCH_CHECK();
}

void vec_try(double *xt, const double *x, double rel, const double *d, long long n, hipStream_t st)
{
    k_try<<<nblk(n), kBlock, 0, st>>>(xt, x, rel, d, n);
    LDDMM_LAUNCH_CHECK();
}

void vec_accept(double *x, double *g, const double *xt, const double *gt, double *s, double *y, long long n,
                hipStream_t st)
{
    k_accept<<<nblk(n), kBlock, 0, st>>>(x, g, xt, gt, s, y, n);
    LDDMM_LAUNCH_CHECK();
}

}  // namespace lddmm
