// varifold_ops.hpp -- per-pair operators of the varifold term for the tiled
// driver of pair_driver.hpp (Cpp/src/varifold3D.cpp:41-357,
// Cpp/src/varifold2D.cpp:33-216).  The factorisation of the per-pair
// gradient terms into the accumulated sums E, S, G, U is described in
// varifold_kernels.hip.
#pragma once

#include "pair_driver.hpp"

namespace lddmm {
namespace ops {

enum { kCenG = 0, kCenC = 1 };
enum { kDirB = 0, kDirL = 1, kDirO = 2, kDirU = 3 };

template <int CEN, typename real>
__device__ __forceinline__ real cen_value(real r2, real inv_cs2)
{
    if constexpr (CEN == kCenG) {
        if constexpr (sizeof(real) == 4) return __builtin_amdgcn_exp2f(r2 * (inv_cs2 * -1.4426950408889634f));
        else return exp(-r2 * inv_cs2);
    } else {
        return real(1) / (real(1) + r2 * inv_cs2);
    }
}

// FP32 position kernels on coordinates scaled by cen_kappa: 'G'
// exp2(-r2) (scaled by sqrt(ics2 log2 e)), 'C' 1 / (1 + r2) (scaled by
// sqrt(ics2)) with v_rcp_f32 (~1 ulp)
template <int CEN>
__device__ __forceinline__ float cen_value_scaled(float r2)
{
    if constexpr (CEN == kCenG) return __builtin_amdgcn_exp2f(-r2);
    else return __builtin_amdgcn_rcpf(1.0f + r2);
}

// 1 / x for x >= 1 in FP64: the hardware reciprocal v_rcp_f64 and two
// Newton steps y += y (1 - x y) (each doubles the correct bits; within an
// ulp), in place of the library division's scaling and fix-up sequence
__device__ __forceinline__ double rcp_ge1(double x)
{
    double y = __builtin_amdgcn_rcp(x);
    y = fma(y, fma(-x, y, 1.0), y);
    return fma(y, fma(-x, y, 1.0), y);
}

// FP64 position kernels on coordinates scaled by cen_kappa: 'G' 2^-r2 by
// exp2_neg, 'C' 1 / (1 + r2) by rcp_ge1
template <int CEN>
__device__ __forceinline__ double cen_value_scaled64(double r2)
{
    if constexpr (CEN == kCenG) return exp2_neg(r2);
    else return rcp_ge1(1.0 + r2);
}

template <int CEN>
__host__ __device__ inline double cen_kappa(double ics2)
{
    return CEN == kCenG ? sqrt(ics2 * 1.4426950408889634) : sqrt(ics2);
}

template <typename real>
__device__ __forceinline__ real ex(real x)
{
    if constexpr (sizeof(real) == 4) return __builtin_amdgcn_exp2f(x * 1.4426950408889634f);
    else return exp(x);
}

template <int DIR, typename real>
__device__ __forceinline__ real dir_value(real ang, real inv_ds2, real &dku)
{
    if constexpr (DIR == kDirB) {
        dku = real(2) * ang;
        return ang * ang;
    } else if constexpr (DIR == kDirL) {
        dku = real(1);
        return ang;
    } else if constexpr (DIR == kDirO) {
        real f = ex<real>(real(2) * (ang - real(1)) * inv_ds2);
        dku = real(2) * f * inv_ds2;
        return f;
    } else {
        real f = ex<real>(real(2) * (ang * ang - real(1)) * inv_ds2);
        dku = real(4) * ang * f * inv_ds2;
        return f;
    }
}

template <int D_, int CEN, int DIR, class P>
struct OpVf {
    static constexpr int D = D_, F = D_ + 1, OUT = 2 * D_ + 2, R = 1;
    using real = typename P::real;
    using lds_t = real;
    using crd_t = real;
    static constexpr bool kCull = P::fp32 && CEN == kCenG;
    static constexpr bool kExpand = kCull;   // expanded form (pair_driver.hpp)
    // anchored centres scaled by cen_kappa in both precisions
    static constexpr bool kAnchor = true;
    static constexpr int kCountSlot = 2;   // executed-pair counter slot
    struct Args {
        const double *ci, *ui;   // targets: deformed element centres / directions
        int ni;
        const double *cj, *uj, *aj;   // sources
        int nj;
        real ics2, ids2;
        float cut2;   // culling range (cull_r2); +inf: none
        float xr2;    // expanded-form radius bound (expand_r2)
        double kap;   // cen_kappa of the staged offsets
        unsigned long long *exec = nullptr;   // executed-pair count (pair_counter), or none
        const double4 *grp = nullptr;   // bounds of the aligned 64-source groups, or none
    };
    __device__ __forceinline__ static double crd_scale(const Args &A) { return A.kap; }
    struct Tgt {
        real u[D];
    };
    struct Acc {
        real e, s, g[D], u[D];
    };
    __device__ __forceinline__ static void load_src(const Args &A, int j, double (&x)[D], lds_t (&f)[F])
    {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            x[c] = A.cj[(size_t)c * A.nj + j];
            f[c] = (real)A.uj[(size_t)c * A.nj + j];
        }
        f[D] = (real)A.aj[j];
    }
    __device__ __forceinline__ static void load_tgt(const Args &A, int i, Tgt &t, double (&x)[D])
    {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            x[c] = i >= 0 ? A.ci[(size_t)c * A.ni + i] : 0.0;
            t.u[c] = i >= 0 ? (real)A.ui[(size_t)c * A.ni + i] : real(0);
        }
    }
    __device__ __forceinline__ static void zero(Acc &a)
    {
        a.e = a.s = real(0);
#pragma unroll
        for (int c = 0; c < D; ++c) a.g[c] = a.u[c] = real(0);
    }
    // FP32 with Binet: the factor 2 of dku = 2 ang moves to the flush
    static constexpr bool kBinet32 = P::fp32 && DIR == kDirB;
    __device__ __forceinline__ static void pair(const Args &A, const Tgt &t, const crd_t (&xi)[D],
                                                const crd_t (&xj)[D], const lds_t (&fj)[F], Acc &acc)
    {
        real df[D], r2 = real(0), ang = real(0);
#pragma unroll
        for (int c = 0; c < D; ++c) {
            df[c] = xi[c] - xj[c];
            r2 += df[c] * df[c];
            ang += t.u[c] * fj[c];
        }
        real kc;
        if constexpr (P::fp32) kc = cen_value_scaled<CEN>(r2);
        else kc = cen_value_scaled64<CEN>(r2);
        const real w = kc * fj[D];
        if constexpr (kBinet32) {
            // ku = ang^2: wk = (w ang) ang, u-sum of (w ang) u_j (times 2 at the flush)
            const real wa = w * ang;
            const real wk = wa * ang;
            const real gw = (CEN == kCenG) ? wk : wk * kc;
            acc.e += wk;
#pragma unroll
            for (int c = 0; c < D; ++c) {
                acc.g[c] += gw * df[c];
                acc.u[c] += wa * fj[c];
            }
        } else {
            real dku;
            const real ku = dir_value<DIR, real>(ang, A.ids2, dku);
            const real wk = w * ku;
            const real gw = (CEN == kCenG) ? wk : wk * kc;
            const real wd = w * dku;
            acc.e += wk;
            // Binet: wk - wd ang = w ang^2 - 2 w ang^2 = -wk, so S = -E
            if constexpr (DIR != kDirB) acc.s += wk - wd * ang;
#pragma unroll
            for (int c = 0; c < D; ++c) {
                acc.g[c] += gw * df[c];
                acc.u[c] += wd * fj[c];
            }
        }
    }
    __device__ __forceinline__ static void flush(const Args &A, const Acc &a, double (&tot)[OUT])
    {
        tot[0] += (double)a.e;
        tot[1] += DIR == kDirB ? -(double)a.e : (double)a.s;
        // the g-sum is over scaled differences (1 / kappa)
        const double gs = 1.0 / A.kap, us = kBinet32 ? 2.0 : 1.0;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            tot[2 + c] += (double)a.g[c] * gs;
            if constexpr (P::fp32) tot[2 + D + c] += (double)a.u[c] * us;
            else tot[2 + D + c] += a.u[c];
        }
    }
    // expanded form (Gaussian position kernel): kc = exp2(-t); the g-sum
    // sum_j wk (c_i - c_j) is c_i E - sum_j wk c_j with E the e-sum
    __device__ __forceinline__ static void pair_x(const Args &A, const Tgt &t, const float (&m)[D],
                                                  const float (&xj)[D + 1], const lds_t (&fj)[F], Acc &acc)
    {
        float s = xj[D], ang = 0.0f;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            s = __builtin_fmaf(m[c], xj[c], s);
            ang += t.u[c] * fj[c];
        }
        const float w = __builtin_amdgcn_exp2f(-s) * fj[D];
        if constexpr (kBinet32) {
            const float wa = w * ang;
            const float wk = wa * ang;
            acc.e += wk;
#pragma unroll
            for (int c = 0; c < D; ++c) {
                acc.g[c] += wk * xj[c];
                acc.u[c] += wa * fj[c];
            }
        } else {
            float dku;
            const float ku = dir_value<DIR, float>(ang, A.ids2, dku);
            const float wk = w * ku;
            const float wd = w * dku;
            acc.e += wk;
            if constexpr (DIR != kDirB) acc.s += wk - wd * ang;
#pragma unroll
            for (int c = 0; c < D; ++c) {
                acc.g[c] += wk * xj[c];
                acc.u[c] += wd * fj[c];
            }
        }
    }
    __device__ __forceinline__ static void flush_x(const Args &A, const Acc &a, double (&tot)[OUT],
                                                   const crd_t (&xi)[D], float sc)
    {
        const double s = (double)sc, e = (double)a.e * s;
        tot[0] += e;
        tot[1] += DIR == kDirB ? -e : (double)a.s * s;
        const double gs = s / A.kap, us = (kBinet32 ? 2.0 : 1.0) * s;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            tot[2 + c] += ((double)xi[c] * (double)a.e - (double)a.g[c]) * gs;
            tot[2 + D + c] += (double)a.u[c] * us;
        }
    }
};

// value-only: sum_j kc ku a_j (target-target term)
template <int D_, int CEN, int DIR>
struct OpVfValue {
    static constexpr int D = D_, F = D_ + 1, OUT = 1, R = 1;
    using real = double;
    using lds_t = double;
    using crd_t = double;
    static constexpr bool kCull = false;
    struct Args {
        const double *c, *u, *a;
        int n;
        double ics2, ids2;
    };
    struct Tgt {
        double u[D];
    };
    struct Acc {
        double e;
    };
    __device__ __forceinline__ static void load_src(const Args &A, int j, double (&x)[D], lds_t (&f)[F])
    {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            x[c] = A.c[(size_t)c * A.n + j];
            f[c] = A.u[(size_t)c * A.n + j];
        }
        f[D] = A.a[j];
    }
    __device__ __forceinline__ static void load_tgt(const Args &A, int i, Tgt &t, double (&x)[D])
    {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            x[c] = i >= 0 ? A.c[(size_t)c * A.n + i] : 0.0;
            t.u[c] = i >= 0 ? A.u[(size_t)c * A.n + i] : 0.0;
        }
    }
    __device__ __forceinline__ static void zero(Acc &a) { a.e = 0.0; }
    __device__ __forceinline__ static void pair(const Args &A, const Tgt &t, const crd_t (&xi)[D],
                                                const crd_t (&xj)[D], const lds_t (&fj)[F], Acc &acc)
    {
        double r2 = 0.0, ang = 0.0;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            double df = xi[c] - xj[c];
            r2 += df * df;
            ang += t.u[c] * fj[c];
        }
        double dku;
        acc.e += cen_value<CEN, double>(r2, A.ics2) * dir_value<DIR, double>(ang, A.ids2, dku) * fj[D];
    }
    __device__ __forceinline__ static void flush(const Args &, const Acc &a, double (&tot)[OUT]) { tot[0] += a.e; }
};

}  // namespace ops
}  // namespace lddmm
