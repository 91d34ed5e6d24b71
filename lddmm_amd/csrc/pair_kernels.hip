// pair_kernels.hip -- all-pairs kernel sums of the LDDMM flow on gfx950.
//
// Every sum here runs through the tiled driver of pair_driver.hpp; this
// file supplies the per-pair operators and the split reductions that
// finish them.  Reference semantics: multiplyKernel
// (Cpp/src/multiplyKernel.cpp:22-564), dqKernel (Cpp/src/dqKernel.cpp:31-591),
// computeKernel (Cpp/src/computeKernel.cpp:21-470) and the Euler / adjoint
// steps of objgrd (Cpp/src/objgrd.cpp:45-87).
#include <atomic>

#include "kernels.hpp"
#include "pair_driver.hpp"
#include "pair_ops.hpp"

namespace lddmm {

namespace {

using namespace ops;

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
#pragma unroll
    for (int c = 0; c < D; ++c) x[c] = qi[(size_t)c * ni + i];
    for (int j = blockIdx.y; j < nj; j += gridDim.y) {
        double r2 = 0.0;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            double df = x[c] - qj[(size_t)c * nj + j];
            r2 += df * df;
        }
        K[(size_t)j * ni + i] = KernelFnD<D, ORDER>::value(r2, inv_w, inv_w2);
    }
}

// ---------------------------------------------------------------------------
// split reductions (64-thread blocks: one target per lane, the split loop
// unrolled so several loads are in flight; the order of the sum is fixed)
// ---------------------------------------------------------------------------
constexpr int kRedBlock = 64;

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

template <int D>
__global__ __launch_bounds__(kRedBlock) void k_fwd_reduce(int splits, int n, const double *__restrict__ part,
                                                          const double *__restrict__ q,
                                                          const double *__restrict__ a, double h,
                                                          double *__restrict__ qnext, double *__restrict__ vout,
                                                          double *__restrict__ eparts)
{
    const int i = blockIdx.x * kRedBlock + threadIdx.x;
    double e = 0.0;
    if (i < n) {
        double v[D];
        sum_splits<D>(part, splits, n, i, v);
#pragma unroll
        for (int c = 0; c < D; ++c) {
            const size_t o = (size_t)c * n + i;
            e += a[o] * v[c];
            qnext[o] = q[o] + h * v[c];
            if (vout) vout[o] = v[c];
        }
    }
    e = wave_sum(e);
    if (threadIdx.x == 0) eparts[blockIdx.x] = e;
}

template <int D>
__global__ __launch_bounds__(kRedBlock) void k_bwd_reduce(int splits, int n, const double *__restrict__ part,
                                                          double h, double *__restrict__ g,
                                                          double *__restrict__ p)
{
    const int i = blockIdx.x * kRedBlock + threadIdx.x;
    if (i >= n) return;
    double v[2 * D];
    sum_splits<2 * D>(part, splits, n, i, v);
#pragma unroll
    for (int c = 0; c < D; ++c) {
        const size_t o = (size_t)c * n + i;
        g[o] = h * v[c];
        p[o] = p[o] - h * v[D + c];
    }
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_group_bounds(const double *__restrict__ x, int n, int ng,
                                                         double4 *__restrict__ out)
{
    const int g = blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (g < ng) group_bounds_wave<D>(x, n, g, threadIdx.x & 63, out);
}

void check_dim_order(int dim, int order)
{
    if (dim != 2 && dim != 3) throw Error(kErrArg, "dimension must be 2 or 3");
    if (order < -1 || order > 4) throw Error(kErrArg, "knlOrder must be -1 (Gaussian) or 0..4 (Matern)");
}

// f(Dim, Order, Prec)
template <typename F>
void dispatch_all(int dim, int order, int prec, F &&f)
{
    check_dim_order(dim, order);
    if (prec != kFP64 && prec != kFP32) throw Error(kErrArg, "unknown precision policy");
    dispatch_order(order, [&](auto o) {
        auto with_prec = [&](auto d) {
            if (prec == kFP32) f(d, o, PrecF32{});
            else f(d, o, PrecF64{});
        };
        if (dim == 2) with_prec(std::integral_constant<int, 2>{});
        else with_prec(std::integral_constant<int, 3>{});
    });
}

}  // namespace

int kv_targets_per_block(int prec) { return prec == kFP32 ? 64 * kKvR32 : 64; }
int bwd_targets_per_block(int prec, int order) { return prec == kFP32 ? 64 * bwd_r32(order) : 64; }

PairSplit plan_kv(int ni, int nj, int prec, int target_blocks)
{
    return make_split_ti(ni, nj, kv_targets_per_block(prec), target_blocks);
}

PairSplit plan_bwd(int ni, int nj, int prec, int order, int target_blocks)
{
    return make_split_ti(ni, nj, bwd_targets_per_block(prec, order), target_blocks);
}

PairSplit plan_single(int ni, int nj, int target_blocks) { return make_split_ti(ni, nj, 64, target_blocks); }

namespace {
std::atomic<bool> g_pair_culling{true}, g_pair_expansion{true};
std::atomic<int> g_cull_bits{kDefaultCullBits};
std::atomic<bool> g_pair_groups{true};
std::atomic<unsigned long long *> g_pair_counter{nullptr};
}

void set_pair_counter(unsigned long long *dev_counter) { g_pair_counter.store(dev_counter); }
unsigned long long *pair_counter() { return g_pair_counter.load(); }
void set_pair_culling(bool on) { g_pair_culling.store(on); }
bool pair_culling() { return g_pair_culling.load(); }
void set_pair_cull_bits(int bits)
{
    if (bits < 24 || bits > 126) throw Error(kErrArg, "the culling level must be 24..126 (kernel values below 2^-level)");
    g_cull_bits.store(bits);
}
int pair_cull_bits() { return g_pair_culling.load() ? g_cull_bits.load() : 0; }
void set_pair_groups(bool on) { g_pair_groups.store(on); }
bool pair_groups() { return g_pair_groups.load() && pair_cull_bits() > 0; }
void set_pair_expansion(bool on) { g_pair_expansion.store(on); }
bool pair_expansion() { return g_pair_expansion.load(); }

int group_count(int n) { return (n + 63) / 64; }

void launch_group_bounds(int dim, const double *x, int n, double4 *out, hipStream_t st)
{
    const int ng = group_count(n);
    if (ng <= 0) return;
    if (dim == 2) k_group_bounds<2><<<ceil_div(ng, kWaves), kBlock, 0, st>>>(x, n, ng, out);
    else k_group_bounds<3><<<ceil_div(ng, kWaves), kBlock, 0, st>>>(x, n, ng, out);
    LDDMM_LAUNCH_CHECK();
}

void launch_kv(int dim, int order, int prec, const double *qi, int ni, const double *qj, const double *a, int nj,
               double width, const PairSplit &sp, double *part, hipStream_t st, const double4 *grp)
{
    if (ni <= 0) return;
    dispatch_all(dim, order, prec, [&](auto Dc, auto Oc, auto Pc) {
        constexpr int R = decltype(Pc)::fp32 ? kKvR32 : 1;
        using Op = OpKV<Dc(), Oc(), decltype(Pc), R>;
        using real = typename Op::real;
        const double iw = 1.0 / width, iw2 = 1.0 / (width * width);
        typename Op::Args A{qi,
                            qj,
                            a,
                            ni,
                            nj,
                            (real)iw,
                            (real)iw2,
                            cull_r2(pair_cull_bits()),
                            expand_r2(pair_expansion()),
                            Op::KF::kappa(iw, iw2),
                            pair_counter(),
                            grp};
        launch_pairs<Op>(A, ni, nj, sp, 0, part, st);
    });
}

void launch_sum_splits(int comps, int splits, int n, const double *part, double scale, double *out,
                       hipStream_t st)
{
    if (n <= 0) return;
    dim3 grid(ceil_div(n, kBlock), comps);
    k_sum_splits<<<grid, kBlock, 0, st>>>(comps, splits, n, part, scale, out);
    LDDMM_LAUNCH_CHECK();
}

int fwd_reduce_blocks(int n) { return ceil_div(n, kRedBlock); }

void launch_fwd_reduce(int dim, int splits, int n, const double *part, const double *q, const double *a,
                       double h, double *qnext, double *vout, double *eparts, hipStream_t st)
{
    int nb = fwd_reduce_blocks(n);
    if (dim == 2) k_fwd_reduce<2><<<nb, kRedBlock, 0, st>>>(splits, n, part, q, a, h, qnext, vout, eparts);
    else k_fwd_reduce<3><<<nb, kRedBlock, 0, st>>>(splits, n, part, q, a, h, qnext, vout, eparts);
    LDDMM_LAUNCH_CHECK();
}

void launch_bwd_pairs(int dim, int order, int prec, const double *q, const double *a, const double *p, int n,
                      double width, const PairSplit &sp, double *part, hipStream_t st, const double4 *grp)
{
    dispatch_all(dim, order, prec, [&](auto Dc, auto Oc, auto Pc) {
        constexpr int R = decltype(Pc)::fp32 ? bwd_r32(Oc()) : 1;
        using Op = OpBwd<Dc(), Oc(), decltype(Pc), R>;
        using real = typename Op::real;
        const double iw = 1.0 / width, iw2 = 1.0 / (width * width), kap = Op::KF::kappa(iw, iw2);
        typename Op::Args A{q,   a,   p, n, (real)iw, (real)iw2, cull_r2(pair_cull_bits()), expand_r2(pair_expansion()),
                            kap, Op::KF::cmul(iw, iw2) / kap, pair_counter(), grp};
        launch_pairs<Op>(A, n, n, sp, 0, part, st);
    });
}

void launch_bwd_reduce(int dim, int splits, int n, const double *part, double h, double *g, double *p,
                       hipStream_t st)
{
    int nb = ceil_div(n, kRedBlock);
    if (dim == 2) k_bwd_reduce<2><<<nb, kRedBlock, 0, st>>>(splits, n, part, h, g, p);
    else k_bwd_reduce<3><<<nb, kRedBlock, 0, st>>>(splits, n, part, h, g, p);
    LDDMM_LAUNCH_CHECK();
}

void launch_dq_self(int dim, int order, int prec, const double *q, const double *l, const double *r, int n,
                    double width, const PairSplit &sp, double *part, hipStream_t st)
{
    dispatch_all(dim, order, prec, [&](auto Dc, auto Oc, auto Pc) {
        using Op = OpDqSelf<Dc(), Oc(), decltype(Pc)>;
        using real = typename Op::real;
        const double iw = 1.0 / width, iw2 = 1.0 / (width * width), kap = Op::KF::kappa(iw, iw2);
        typename Op::Args A{q, l, r, n, (real)iw, (real)iw2, kap, Op::KF::cmul(iw, iw2) / kap};
        launch_pairs<Op>(A, n, n, sp, 0, part, st);
    });
}

void launch_dq_cross(int dim, int order, int prec, const double *qi, const double *u, int ni, const double *qj,
                     const double *v, int nj, double width, const PairSplit &sp, double *part, hipStream_t st)
{
    if (ni <= 0) return;
    dispatch_all(dim, order, prec, [&](auto Dc, auto Oc, auto Pc) {
        using Op = OpDqCross<Dc(), Oc(), decltype(Pc)>;
        using real = typename Op::real;
        const double iw = 1.0 / width, iw2 = 1.0 / (width * width), kap = Op::KF::kappa(iw, iw2);
        typename Op::Args A{qi, u, qj, v, ni, nj, (real)iw, (real)iw2, kap, Op::KF::cmul(iw, iw2) / kap};
        launch_pairs<Op>(A, ni, nj, sp, 0, part, st);
    });
}

void launch_dense_kernel(int dim, int order, const double *qi, int ni, const double *qj, int nj, double width,
                         double *K, hipStream_t st)
{
    if (ni <= 0 || nj <= 0) return;
    dispatch_all(dim, order, kFP64, [&](auto Dc, auto Oc, auto) {
        dim3 grid(ceil_div(ni, kBlock), nj < 65535 ? nj : 65535);
        k_dense<Dc(), Oc()><<<grid, kBlock, 0, st>>>(qi, ni, qj, nj, 1.0 / width, 1.0 / (width * width), K);
    });
    LDDMM_LAUNCH_CHECK();
}

}  // namespace lddmm
