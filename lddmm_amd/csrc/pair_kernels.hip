// pair_kernels.hip -- all-pairs kernel sums of the LDDMM flow on gfx950.
//
// Every kernel here is a tiled all-pairs sum: a block of kBlock (=256,
// four wave64) threads owns kBlock target points i and one chunk of the
// source points j.  Sources are staged through LDS in tiles of kTileJ SoA
// coordinates; all 64 lanes of a wave read the same LDS word (broadcast,
// no bank conflicts) while each lane keeps its own target in registers.
// The source chunks (j-splits) give the chip enough workgroups at the
// configs' sizes (N = 10,242 -> 41 target tiles only); the per-split
// partial sums are reduced in a fixed order by a second, HBM-light kernel,
// so every result is bitwise repeatable.
//
// Reference semantics: multiplyKernel (Cpp/src/multiplyKernel.cpp:22-564),
// dqKernel (Cpp/src/dqKernel.cpp:31-591), computeKernel
// (Cpp/src/computeKernel.cpp:21-470) and the Euler / adjoint steps of
// objgrd (Cpp/src/objgrd.cpp:45-87).
#include "kernels.hpp"

namespace lddmm {

namespace {

template <int D>
__device__ __forceinline__ void load_pt(double (&x)[D], const double *__restrict__ p, int n, int i)
{
#pragma unroll
    for (int c = 0; c < D; ++c) x[c] = p[(size_t)c * n + i];
}

// ---------------------------------------------------------------------------
// v_i = sum_j k(qi_i, qj_j) a_j      (forward flow and K.alpha operators)
// ---------------------------------------------------------------------------
template <int D, int ORDER>
__global__ __launch_bounds__(kBlock) void k_kv(const double *__restrict__ qi, int ni,
                                               const double *__restrict__ qj, const double *__restrict__ a,
                                               int nj, int chunk, int tiles, double inv_w, double inv_w2,
                                               double *__restrict__ part)
{
    __shared__ double sq[D][kTileJ];
    __shared__ double sa[D][kTileJ];
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int tile = L % tiles, s = L / tiles;
    const int i = tile * kBlock + threadIdx.x;
    const int j0 = s * chunk, j1 = min(nj, j0 + chunk);
    double x[D], acc[D];
#pragma unroll
    for (int c = 0; c < D; ++c) {
        x[c] = i < ni ? qi[(size_t)c * ni + i] : 0.0;
        acc[c] = 0.0;
    }
    for (int jt = j0; jt < j1; jt += kTileJ) {
        const int cnt = min(kTileJ, j1 - jt);
        __syncthreads();
        if ((int)threadIdx.x < cnt) {
#pragma unroll
            for (int c = 0; c < D; ++c) {
                sq[c][threadIdx.x] = qj[(size_t)c * nj + jt + threadIdx.x];
                sa[c][threadIdx.x] = a[(size_t)c * nj + jt + threadIdx.x];
            }
        }
        __syncthreads();
#pragma unroll 4
        for (int t = 0; t < cnt; ++t) {
            double r2 = 0.0;
#pragma unroll
            for (int c = 0; c < D; ++c) {
                double df = x[c] - sq[c][t];
                r2 += df * df;
            }
            double k = KernelFn<ORDER>::value(r2, inv_w, inv_w2);
#pragma unroll
            for (int c = 0; c < D; ++c) acc[c] += k * sa[c][t];
        }
    }
    if (i < ni) {
#pragma unroll
        for (int c = 0; c < D; ++c) part[((size_t)s * D + c) * ni + i] = acc[c];
    }
}

// ---------------------------------------------------------------------------
// Fused backward pass of the adjoint (Cpp/src/objgrd.cpp:76-84): one kernel
// evaluation per pair feeds both K(q) amp and the dq sum of l^T K(q) r
// with amp = a - p, l = amp - a/2, r = a.
// ---------------------------------------------------------------------------
template <int D, int ORDER>
__global__ __launch_bounds__(kBlock) void k_bwd(const double *__restrict__ q, const double *__restrict__ a,
                                                const double *__restrict__ p, int n, int chunk, int tiles,
                                                double inv_w, double inv_w2, double *__restrict__ part)
{
    __shared__ double sq[D][kTileJ];
    __shared__ double sa[D][kTileJ];   // r_j = a_j
    __shared__ double sm[D][kTileJ];   // amp_j
    __shared__ double sl[D][kTileJ];   // l_j
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int tile = L % tiles, s = L / tiles;
    const int i = tile * kBlock + threadIdx.x;
    const int j0 = s * chunk, j1 = min(n, j0 + chunk);
    double x[D], ai[D], li[D], g[D], pd[D];
#pragma unroll
    for (int c = 0; c < D; ++c) {
        const bool in = i < n;
        x[c] = in ? q[(size_t)c * n + i] : 0.0;
        ai[c] = in ? a[(size_t)c * n + i] : 0.0;
        double amp = ai[c] - (in ? p[(size_t)c * n + i] : 0.0);
        li[c] = amp - 0.5 * ai[c];
        g[c] = 0.0;
        pd[c] = 0.0;
    }
    for (int jt = j0; jt < j1; jt += kTileJ) {
        const int cnt = min(kTileJ, j1 - jt);
        __syncthreads();
        if ((int)threadIdx.x < cnt) {
            const size_t jj = jt + threadIdx.x;
#pragma unroll
            for (int c = 0; c < D; ++c) {
                double aj = a[(size_t)c * n + jj];
                double amp = aj - p[(size_t)c * n + jj];
                sq[c][threadIdx.x] = q[(size_t)c * n + jj];
                sa[c][threadIdx.x] = aj;
                sm[c][threadIdx.x] = amp;
                sl[c][threadIdx.x] = amp - 0.5 * aj;
            }
        }
        __syncthreads();
#pragma unroll 2
        for (int t = 0; t < cnt; ++t) {
            double df[D], r2 = 0.0;
#pragma unroll
            for (int c = 0; c < D; ++c) {
                df[c] = x[c] - sq[c][t];
                r2 += df[c] * df[c];
            }
            double k, cc;
            KernelFn<ORDER>::eval(r2, inv_w, inv_w2, k, cc);
            double lr = 0.0;
#pragma unroll
            for (int c = 0; c < D; ++c) lr += li[c] * sa[c][t];
#pragma unroll
            for (int c = 0; c < D; ++c) lr += sl[c][t] * ai[c];
            double f = lr * cc;
#pragma unroll
            for (int c = 0; c < D; ++c) {
                g[c] += k * sm[c][t];
                pd[c] += f * df[c];
            }
        }
    }
    if (i < n) {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            part[((size_t)s * 2 * D + c) * n + i] = g[c];
            part[((size_t)s * 2 * D + D + c) * n + i] = pd[c];
        }
    }
}

// ---------------------------------------------------------------------------
// dq of l^T K(q) r over one point set (Cpp/src/dqKernel.cpp:31-305)
// ---------------------------------------------------------------------------
template <int D, int ORDER>
__global__ __launch_bounds__(kBlock) void k_dq_self(const double *__restrict__ q, const double *__restrict__ l,
                                                    const double *__restrict__ r, int n, int chunk, int tiles,
                                                    double inv_w, double inv_w2, double *__restrict__ part)
{
    __shared__ double sq[D][kTileJ];
    __shared__ double sl[D][kTileJ];
    __shared__ double sr[D][kTileJ];
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int tile = L % tiles, s = L / tiles;
    const int i = tile * kBlock + threadIdx.x;
    const int j0 = s * chunk, j1 = min(n, j0 + chunk);
    double x[D], li[D], ri[D], pd[D];
#pragma unroll
    for (int c = 0; c < D; ++c) {
        const bool in = i < n;
        x[c] = in ? q[(size_t)c * n + i] : 0.0;
        li[c] = in ? l[(size_t)c * n + i] : 0.0;
        ri[c] = in ? r[(size_t)c * n + i] : 0.0;
        pd[c] = 0.0;
    }
    for (int jt = j0; jt < j1; jt += kTileJ) {
        const int cnt = min(kTileJ, j1 - jt);
        __syncthreads();
        if ((int)threadIdx.x < cnt) {
            const size_t jj = jt + threadIdx.x;
#pragma unroll
            for (int c = 0; c < D; ++c) {
                sq[c][threadIdx.x] = q[(size_t)c * n + jj];
                sl[c][threadIdx.x] = l[(size_t)c * n + jj];
                sr[c][threadIdx.x] = r[(size_t)c * n + jj];
            }
        }
        __syncthreads();
#pragma unroll 2
        for (int t = 0; t < cnt; ++t) {
            double df[D], r2 = 0.0;
#pragma unroll
            for (int c = 0; c < D; ++c) {
                df[c] = x[c] - sq[c][t];
                r2 += df[c] * df[c];
            }
            double k, cc;
            KernelFn<ORDER>::eval(r2, inv_w, inv_w2, k, cc);
            double lr = 0.0;
#pragma unroll
            for (int c = 0; c < D; ++c) lr += li[c] * sr[c][t];
#pragma unroll
            for (int c = 0; c < D; ++c) lr += sl[c][t] * ri[c];
            double f = lr * cc;
#pragma unroll
            for (int c = 0; c < D; ++c) pd[c] += f * df[c];
        }
    }
    if (i < n) {
#pragma unroll
        for (int c = 0; c < D; ++c) part[((size_t)s * D + c) * n + i] = pd[c];
    }
}

// ---------------------------------------------------------------------------
// Cross dq: sum_j (u_i.v_j) c_ij (qi_i - qj_j) (Cpp/src/dqKernel.cpp:307-591;
// the qj-side gradient is the same sum with the roles of the sets swapped)
// ---------------------------------------------------------------------------
template <int D, int ORDER>
__global__ __launch_bounds__(kBlock) void k_dq_cross(const double *__restrict__ qi, const double *__restrict__ u,
                                                     int ni, const double *__restrict__ qj,
                                                     const double *__restrict__ v, int nj, int chunk, int tiles,
                                                     double inv_w, double inv_w2, double *__restrict__ part)
{
    __shared__ double sq[D][kTileJ];
    __shared__ double sv[D][kTileJ];
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int tile = L % tiles, s = L / tiles;
    const int i = tile * kBlock + threadIdx.x;
    const int j0 = s * chunk, j1 = min(nj, j0 + chunk);
    double x[D], ui[D], pd[D];
#pragma unroll
    for (int c = 0; c < D; ++c) {
        const bool in = i < ni;
        x[c] = in ? qi[(size_t)c * ni + i] : 0.0;
        ui[c] = in ? u[(size_t)c * ni + i] : 0.0;
        pd[c] = 0.0;
    }
    for (int jt = j0; jt < j1; jt += kTileJ) {
        const int cnt = min(kTileJ, j1 - jt);
        __syncthreads();
        if ((int)threadIdx.x < cnt) {
            const size_t jj = jt + threadIdx.x;
#pragma unroll
            for (int c = 0; c < D; ++c) {
                sq[c][threadIdx.x] = qj[(size_t)c * nj + jj];
                sv[c][threadIdx.x] = v[(size_t)c * nj + jj];
            }
        }
        __syncthreads();
#pragma unroll 2
        for (int t = 0; t < cnt; ++t) {
            double df[D], r2 = 0.0;
#pragma unroll
            for (int c = 0; c < D; ++c) {
                df[c] = x[c] - sq[c][t];
                r2 += df[c] * df[c];
            }
            double k, cc;
            KernelFn<ORDER>::eval(r2, inv_w, inv_w2, k, cc);
            double lr = 0.0;
#pragma unroll
            for (int c = 0; c < D; ++c) lr += ui[c] * sv[c][t];
            double f = lr * cc;
#pragma unroll
            for (int c = 0; c < D; ++c) pd[c] += f * df[c];
        }
    }
    if (i < ni) {
#pragma unroll
        for (int c = 0; c < D; ++c) part[((size_t)s * D + c) * ni + i] = pd[c];
    }
}

// ---------------------------------------------------------------------------
// Dense kernel matrix (Cpp/src/computeKernel.cpp:246-470 layout K[j*ni + i])
// ---------------------------------------------------------------------------
template <int D, int ORDER>
__global__ __launch_bounds__(kBlock) void k_dense(const double *__restrict__ qi, int ni,
                                                  const double *__restrict__ qj, int nj, double inv_w,
                                                  double inv_w2, double *__restrict__ K)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= ni) return;
    double x[D];
    load_pt<D>(x, qi, ni, i);
    for (int j = blockIdx.y; j < nj; j += gridDim.y) {
        double r2 = 0.0;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            double df = x[c] - qj[(size_t)c * nj + j];
            r2 += df * df;
        }
        K[(size_t)j * ni + i] = KernelFn<ORDER>::value(r2, inv_w, inv_w2);
    }
}

// ---------------------------------------------------------------------------
// split reductions
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_sum_splits(int comps, int splits, int n,
                                                       const double *__restrict__ part, double scale,
                                                       double *__restrict__ out)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= n) return;
    double s = 0.0;
    for (int k = 0; k < splits; ++k) s += part[((size_t)k * comps + c) * n + i];
    out[(size_t)c * n + i] = scale * s;
}

// fixed-shape tree over the 256 threads of a block; result valid in thread 0
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

template <int D>
__global__ __launch_bounds__(kBlock) void k_fwd_reduce(int splits, int n, const double *__restrict__ part,
                                                       const double *__restrict__ q,
                                                       const double *__restrict__ a, double h,
                                                       double *__restrict__ qnext, double *__restrict__ vout,
                                                       double *__restrict__ eparts)
{
    __shared__ double red[kBlock];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double e = 0.0;
    if (i < n) {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            double v = 0.0;
            for (int k = 0; k < splits; ++k) v += part[((size_t)k * D + c) * n + i];
            const size_t o = (size_t)c * n + i;
            e += a[o] * v;
            qnext[o] = q[o] + h * v;
            if (vout) vout[o] = v;
        }
    }
    double t = block_sum(e, red);
    if (threadIdx.x == 0) eparts[blockIdx.x] = t;
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_bwd_reduce(int splits, int n, const double *__restrict__ part,
                                                       double h, double *__restrict__ g, double *__restrict__ p)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int c = 0; c < D; ++c) {
        double gs = 0.0, ps = 0.0;
        for (int k = 0; k < splits; ++k) {
            gs += part[((size_t)k * 2 * D + c) * n + i];
            ps += part[((size_t)k * 2 * D + D + c) * n + i];
        }
        const size_t o = (size_t)c * n + i;
        g[o] = h * gs;
        p[o] = p[o] - h * ps;
    }
}

void check_dim_order(int dim, int order)
{
    if (dim != 2 && dim != 3) throw Error(kErrArg, "dimension must be 2 or 3");
    if (order < -1 || order > 4) throw Error(kErrArg, "knlOrder must be -1 (Gaussian) or 0..4 (Matern)");
    if (dim == 2 && order >= 0)
        throw Error(kErrUnsupported, "2D Matern kernels (Bessel K0/K1) are not implemented on the GPU path");
}

template <typename F>
void dispatch_dim_order(int dim, int order, F &&f)
{
    check_dim_order(dim, order);
    dispatch_order(order, [&](auto o) {
        if (dim == 2) f(std::integral_constant<int, 2>{}, o);
        else f(std::integral_constant<int, 3>{}, o);
    });
}

}  // namespace

void launch_kv(int dim, int order, const double *qi, int ni, const double *qj, const double *a, int nj,
               double width, const PairSplit &sp, double *part, hipStream_t st)
{
    if (ni <= 0) return;
    dispatch_dim_order(dim, order, [&](auto D, auto O) {
        k_kv<D(), O()><<<sp.blocks(), kBlock, 0, st>>>(qi, ni, qj, a, nj, sp.chunk, sp.tiles, 1.0 / width,
                                                       1.0 / (width * width), part);
    });
    LDDMM_LAUNCH_CHECK();
}

void launch_sum_splits(int comps, int splits, int n, const double *part, double scale, double *out,
                       hipStream_t st)
{
    if (n <= 0) return;
    dim3 grid(ceil_div(n, kBlock), comps);
    k_sum_splits<<<grid, kBlock, 0, st>>>(comps, splits, n, part, scale, out);
    LDDMM_LAUNCH_CHECK();
}

int fwd_reduce_blocks(int n) { return ceil_div(n, kBlock); }

void launch_fwd_reduce(int dim, int splits, int n, const double *part, const double *q, const double *a,
                       double h, double *qnext, double *vout, double *eparts, hipStream_t st)
{
    int nb = fwd_reduce_blocks(n);
    if (dim == 2) k_fwd_reduce<2><<<nb, kBlock, 0, st>>>(splits, n, part, q, a, h, qnext, vout, eparts);
    else k_fwd_reduce<3><<<nb, kBlock, 0, st>>>(splits, n, part, q, a, h, qnext, vout, eparts);
    LDDMM_LAUNCH_CHECK();
}

void launch_bwd_pairs(int dim, int order, const double *q, const double *a, const double *p, int n, double width,
                      const PairSplit &sp, double *part, hipStream_t st)
{
    dispatch_dim_order(dim, order, [&](auto D, auto O) {
        k_bwd<D(), O()><<<sp.blocks(), kBlock, 0, st>>>(q, a, p, n, sp.chunk, sp.tiles, 1.0 / width,
                                                        1.0 / (width * width), part);
    });
    LDDMM_LAUNCH_CHECK();
}

void launch_bwd_reduce(int dim, int splits, int n, const double *part, double h, double *g, double *p,
                       hipStream_t st)
{
    int nb = ceil_div(n, kBlock);
    if (dim == 2) k_bwd_reduce<2><<<nb, kBlock, 0, st>>>(splits, n, part, h, g, p);
    else k_bwd_reduce<3><<<nb, kBlock, 0, st>>>(splits, n, part, h, g, p);
    LDDMM_LAUNCH_CHECK();
}

void launch_dq_self(int dim, int order, const double *q, const double *l, const double *r, int n, double width,
                    const PairSplit &sp, double *part, hipStream_t st)
{
    dispatch_dim_order(dim, order, [&](auto D, auto O) {
        k_dq_self<D(), O()><<<sp.blocks(), kBlock, 0, st>>>(q, l, r, n, sp.chunk, sp.tiles, 1.0 / width,
                                                            1.0 / (width * width), part);
    });
    LDDMM_LAUNCH_CHECK();
}

void launch_dq_cross(int dim, int order, const double *qi, const double *u, int ni, const double *qj,
                     const double *v, int nj, double width, const PairSplit &sp, double *part, hipStream_t st)
{
    if (ni <= 0) return;
    dispatch_dim_order(dim, order, [&](auto D, auto O) {
        k_dq_cross<D(), O()><<<sp.blocks(), kBlock, 0, st>>>(qi, u, ni, qj, v, nj, sp.chunk, sp.tiles,
                                                             1.0 / width, 1.0 / (width * width), part);
    });
    LDDMM_LAUNCH_CHECK();
}

void launch_dense_kernel(int dim, int order, const double *qi, int ni, const double *qj, int nj, double width,
                         double *K, hipStream_t st)
{
    if (ni <= 0 || nj <= 0) return;
    dispatch_dim_order(dim, order, [&](auto D, auto O) {
        dim3 grid(ceil_div(ni, kBlock), nj < 65535 ? nj : 65535);
        k_dense<D(), O()><<<grid, kBlock, 0, st>>>(qi, ni, qj, nj, 1.0 / width, 1.0 / (width * width), K);
    });
    LDDMM_LAUNCH_CHECK();
}

}  // namespace lddmm
