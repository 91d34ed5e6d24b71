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
__global__ void k_copy_neg(double *__restrict__ d, const double *__restrict__ g, long long n)
{
    long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) d[i] = -g[i];
}

// first loop of the two-loop recursion: a = (s.d) / (s.y); d -= a y
__global__ void k_loop1(double *__restrict__ d, const double *__restrict__ y, const double *__restrict__ sd,
                        const double *__restrict__ sy, double *__restrict__ rec, long long n)
{
    const double a = sd[0] / sy[0];
    long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i == 0) rec[0] = a;
    if (i < n) d[i] -= a * y[i];
}

// d *= sy / yy (newest pair) -- H0 of Cpp/src/LBFGS.cpp:69-81
__global__ void k_scale_h0(double *__restrict__ d, const double *__restrict__ sy, const double *__restrict__ yy,
                           long long n)
{
    const double h0 = sy[0] / yy[0];
    long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) d[i] *= h0;
}

// second loop: b = (y.d) / (s.y); d += (a - b) s; the final negation of
// getDirection folded into the last pass
__global__ void k_loop2(double *__restrict__ d, const double *__restrict__ s, const double *__restrict__ yd,
                        const double *__restrict__ sy, const double *__restrict__ rec, int negate, long long n)
{
    const double c = rec[0] - yd[0] / sy[0];
    long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        double v = d[i] + c * s[i];
        d[i] = negate ? -v : v;
    }
}

__global__ void k_scale(double *__restrict__ d, double a, long long n)
{
    long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) d[i] *= a;
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

void vec_copy_neg(double *d, const double *g, long long n, hipStream_t st)
{
    k_copy_neg<<<nblk(n), kBlock, 0, st>>>(d, g, n);
    LDDMM_LAUNCH_CHECK();
}

void vec_loop1(double *d, const double *y, const double *sd, const double *sy, double *rec, long long n,
               hipStream_t st)
{
    k_loop1<<<nblk(n), kBlock, 0, st>>>(d, y, sd, sy, rec, n);
    LDDMM_LAUNCH_CHECK();
}

void vec_scale_h0(double *d, const double *sy, const double *yy, long long n, hipStream_t st)
{
    k_scale_h0<<<nblk(n), kBlock, 0, st>>>(d, sy, yy, n);
    LDDMM_LAUNCH_CHECK();
}

void vec_loop2(double *d, const double *s, const double *yd, const double *sy, const double *rec, bool negate,
               long long n, hipStream_t st)
{
    k_loop2<<<nblk(n), kBlock, 0, st>>>(d, s, yd, sy, rec, negate ? 1 : 0, n);
    LDDMM_LAUNCH_CHECK();
}

void vec_scale(double *d, double a, long long n, hipStream_t st)
{
    k_scale<<<nblk(n), kBlock, 0, st>>>(d, a, n);
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
