// pair_ops.hpp -- per-pair operators of the point kernels, for the tiled
// driver of pair_driver.hpp (single-problem and batched launches).
// Reference semantics: multiplyKernel (Cpp/src/multiplyKernel.cpp:22-564),
// dqKernel (Cpp/src/dqKernel.cpp:31-591) and the adjoint step of objgrd
// (Cpp/src/objgrd.cpp:76-84).
#pragma once

#include "pair_driver.hpp"

namespace lddmm {
namespace ops {

// ---------------------------------------------------------------------------
// v_i = sum_j k(qi_i, qj_j) a_j   (forward flow, multiplyKernel)
// ---------------------------------------------------------------------------
template <int D_, int ORDER, class P, int R_>
struct OpKV {
    static constexpr int D = D_, F = D_, OUT = D_, R = R_;
    using real = typename P::real;
    using lds_t = real;
    using crd_t = real;
    static constexpr bool kCull = P::fp32 && ORDER == -1;
    using KF = KernelFnT<ORDER, real, D_>;
    struct Args {
        const double *qi, *qj, *a;
        int ni, nj;
        real iw, iw2;
        float cut2;   // culling range (cull_r2); +inf: none
        double kap;   // KF::kappa
    };
    __device__ __forceinline__ static double crd_scale(const Args &A) { return A.kap; }
    struct Tgt {
    };
    struct Acc {
        real v[D];
    };
    __device__ __forceinline__ static void load_src(const Args &A, int j, double (&x)[D], lds_t (&f)[F])
    {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            x[c] = A.qj[(size_t)c * A.nj + j];
            f[c] = (real)A.a[(size_t)c * A.nj + j];
        }
    }
    __device__ __forceinline__ static void load_tgt(const Args &A, int i, Tgt &, double (&x)[D])
    {
#pragma unroll
        for (int c = 0; c < D; ++c) x[c] = i >= 0 ? A.qi[(size_t)c * A.ni + i] : 0.0;
    }
    __device__ __forceinline__ static void zero(Acc &a)
    {
#pragma unroll
        for (int c = 0; c < D; ++c) a.v[c] = real(0);
    }
    __device__ __forceinline__ static void pair(const Args &A, const Tgt &, const crd_t (&xi)[D],
                                                const crd_t (&xj)[D], const lds_t (&fj)[F], Acc &acc)
    {
        real r2 = real(0);
#pragma unroll
        for (int c = 0; c < D; ++c) {
            real df = xi[c] - xj[c];
            r2 += df * df;
        }
        real k = KF::value(r2, A.iw, A.iw2);
#pragma unroll
        for (int c = 0; c < D; ++c) acc.v[c] += k * fj[c];
    }
    __device__ __forceinline__ static void flush(const Args &, const Acc &a, double (&tot)[OUT])
    {
#pragma unroll
        for (int c = 0; c < D; ++c) tot[c] += (double)a.v[c];
    }
};

// ---------------------------------------------------------------------------
// Fused backward pass of the adjoint (Cpp/src/objgrd.cpp:76-84): one kernel
// evaluation per pair feeds K(q) amp and the dq sum of l^T K(q) r with
// amp = a - p, l = amp - a/2, r = a.  Outputs: [K amp (D), dq (D)].
// R_ targets per lane (two in FP32, where registers allow).
// ---------------------------------------------------------------------------
template <int D_, int ORDER, class P, int R_ = 1>
struct OpBwd {
    static constexpr int D = D_, F = 3 * D_, OUT = 2 * D_, R = R_;
    using real = typename P::real;
    using lds_t = real;
    using crd_t = real;
    static constexpr bool kCull = P::fp32 && ORDER == -1;
    using KF = KernelFnT<ORDER, real, D_>;
    struct Args {
        const double *q, *a, *p;
        int n;
        real iw, iw2;
        float cut2;        // culling range (cull_r2); +inf: none
        double kap, pds;   // KF::kappa; dq flush factor KF::cmul / kappa
    };
    __device__ __forceinline__ static double crd_scale(const Args &A) { return A.kap; }
    struct Tgt {
        real a[D], l[D];
    };
    struct Acc {
        real g[D], pd[D];
    };
    __device__ __forceinline__ static void load_src(const Args &A, int j, double (&x)[D], lds_t (&f)[F])
    {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            const size_t o = (size_t)c * A.n + j;
            double aj = A.a[o], amp = aj - A.p[o];
            x[c] = A.q[o];
            f[c] = (real)aj;
            f[D + c] = (real)amp;
            f[2 * D + c] = (real)(amp - 0.5 * aj);
        }
    }
    __device__ __forceinline__ static void load_tgt(const Args &A, int i, Tgt &t, double (&x)[D])
    {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            const size_t o = (size_t)c * A.n + (i >= 0 ? i : 0);
            double ai = i >= 0 ? A.a[o] : 0.0, amp = ai - (i >= 0 ? A.p[o] : 0.0);
            x[c] = i >= 0 ? A.q[o] : 0.0;
            t.a[c] = (real)ai;
            t.l[c] = (real)(amp - 0.5 * ai);
        }
    }
    __device__ __forceinline__ static void zero(Acc &a)
    {
#pragma unroll
        for (int c = 0; c < D; ++c) a.g[c] = a.pd[c] = real(0);
    }
    __device__ __forceinline__ static void pair(const Args &A, const Tgt &t, const crd_t (&xi)[D],
                                                const crd_t (&xj)[D], const lds_t (&fj)[F], Acc &acc)
    {
        real df[D], r2 = real(0);
#pragma unroll
        for (int c = 0; c < D; ++c) {
            df[c] = xi[c] - xj[c];
            r2 += df[c] * df[c];
        }
        real k, cf;
        KF::eval(r2, A.iw, A.iw2, k, cf);
        real lr = real(0);
#pragma unroll
        for (int c = 0; c < D; ++c) lr += t.l[c] * fj[c];
#pragma unroll
        for (int c = 0; c < D; ++c) lr += fj[2 * D + c] * t.a[c];
        const real f = lr * cf;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            acc.g[c] += k * fj[D + c];
            acc.pd[c] += f * df[c];
        }
    }
    __device__ __forceinline__ static void flush(const Args &A, const Acc &a, double (&tot)[OUT])
    {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            tot[c] += (double)a.g[c];
            if constexpr (P::fp32) tot[D + c] += (double)a.pd[c] * A.pds;
            else tot[D + c] += a.pd[c];
        }
    }
};

// ---------------------------------------------------------------------------
// dq of l^T K(q) r over one point set (Cpp/src/dqKernel.cpp:31-305)
// ---------------------------------------------------------------------------
template <int D_, int ORDER, class P>
struct OpDqSelf {
    static constexpr int D = D_, F = 2 * D_, OUT = D_, R = 1;
    using real = typename P::real;
    using lds_t = real;
    using crd_t = real;
    static constexpr bool kCull = false;
    using KF = KernelFnT<ORDER, real, D_>;
    struct Args {
        const double *q, *l, *r;
        int n;
        real iw, iw2;
        double kap, pds;   // KF::kappa; flush factor KF::cmul / kappa
    };
    __device__ __forceinline__ static double crd_scale(const Args &A) { return A.kap; }
    struct Tgt {
        real l[D], r[D];
    };
    struct Acc {
        real pd[D];
    };
    __device__ __forceinline__ static void load_src(const Args &A, int j, double (&x)[D], lds_t (&f)[F])
    {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            const size_t o = (size_t)c * A.n + j;
            x[c] = A.q[o];
            f[c] = (real)A.l[o];
            f[D + c] = (real)A.r[o];
        }
    }
    __device__ __forceinline__ static void load_tgt(const Args &A, int i, Tgt &t, double (&x)[D])
    {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            const size_t o = (size_t)c * A.n + (i >= 0 ? i : 0);
            x[c] = i >= 0 ? A.q[o] : 0.0;
            t.l[c] = i >= 0 ? (real)A.l[o] : real(0);
            t.r[c] = i >= 0 ? (real)A.r[o] : real(0);
        }
    }
    __device__ __forceinline__ static void zero(Acc &a)
    {
#pragma unroll
        for (int c = 0; c < D; ++c) a.pd[c] = real(0);
    }
    __device__ __forceinline__ static void pair(const Args &A, const Tgt &t, const crd_t (&xi)[D],
                                                const crd_t (&xj)[D], const lds_t (&fj)[F], Acc &acc)
    {
        real df[D], r2 = real(0);
#pragma unroll
        for (int c = 0; c < D; ++c) {
            df[c] = xi[c] - xj[c];
            r2 += df[c] * df[c];
        }
        real k, cf;
        KF::eval(r2, A.iw, A.iw2, k, cf);
        real lr = real(0);
#pragma unroll
        for (int c = 0; c < D; ++c) lr += t.l[c] * fj[D + c];
#pragma unroll
        for (int c = 0; c < D; ++c) lr += fj[c] * t.r[c];
        const real f = lr * cf;
#pragma unroll
        for (int c = 0; c < D; ++c) acc.pd[c] += f * df[c];
    }
    __device__ __forceinline__ static void flush(const Args &A, const Acc &a, double (&tot)[OUT])
    {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            if constexpr (P::fp32) tot[c] += (double)a.pd[c] * A.pds;
            else tot[c] += a.pd[c];
        }
    }
};

// ---------------------------------------------------------------------------
// Cross dq: sum_j (u_i.v_j) c_ij (qi_i - qj_j) (Cpp/src/dqKernel.cpp:307-591;
// the qj-side gradient is the same sum with the roles of the sets swapped)
// ---------------------------------------------------------------------------
template <int D_, int ORDER, class P>
struct OpDqCross {
    static constexpr int D = D_, F = D_, OUT = D_, R = 1;
    using real = typename P::real;
    using lds_t = real;
    using crd_t = real;
    static constexpr bool kCull = false;
    using KF = KernelFnT<ORDER, real, D_>;
    struct Args {
        const double *qi, *u, *qj, *v;
        int ni, nj;
        real iw, iw2;
        double kap, pds;   // KF::kappa; flush factor KF::cmul / kappa
    };
    __device__ __forceinline__ static double crd_scale(const Args &A) { return A.kap; }
    struct Tgt {
        real u[D];
    };
    struct Acc {
        real pd[D];
    };
    __device__ __forceinline__ static void load_src(const Args &A, int j, double (&x)[D], lds_t (&f)[F])
    {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            x[c] = A.qj[(size_t)c * A.nj + j];
            f[c] = (real)A.v[(size_t)c * A.nj + j];
        }
    }
    __device__ __forceinline__ static void load_tgt(const Args &A, int i, Tgt &t, double (&x)[D])
    {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            x[c] = i >= 0 ? A.qi[(size_t)c * A.ni + i] : 0.0;
            t.u[c] = i >= 0 ? (real)A.u[(size_t)c * A.ni + i] : real(0);
        }
    }
    __device__ __forceinline__ static void zero(Acc &a)
    {
#pragma unroll
        for (int c = 0; c < D; ++c) a.pd[c] = real(0);
    }
    __device__ __forceinline__ static void pair(const Args &A, const Tgt &t, const crd_t (&xi)[D],
                                                const crd_t (&xj)[D], const lds_t (&fj)[F], Acc &acc)
    {
        real df[D], r2 = real(0);
#pragma unroll
        for (int c = 0; c < D; ++c) {
            df[c] = xi[c] - xj[c];
            r2 += df[c] * df[c];
        }
        real k, cf;
        KF::eval(r2, A.iw, A.iw2, k, cf);
        real lr = real(0);
#pragma unroll
        for (int c = 0; c < D; ++c) lr += t.u[c] * fj[c];
        const real f = lr * cf;
#pragma unroll
        for (int c = 0; c < D; ++c) acc.pd[c] += f * df[c];
    }
    __device__ __forceinline__ static void flush(const Args &A, const Acc &a, double (&tot)[OUT])
    {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            if constexpr (P::fp32) tot[c] += (double)a.pd[c] * A.pds;
            else tot[c] += a.pd[c];
        }
    }
};

}  // namespace ops
}  // namespace lddmm
