// pair_ops.hpp -- per-pair operators of the point kernels, for the tiled
// driver of pair_driver.hpp (single-problem and batched launches).
// Reference semantics: multiplyKernel (Cpp/src/multiplyKernel.cpp:22-564),
// dqKernel (Cpp/src/dqKernel.cpp:31-591) and the adjoint step of objgrd
// (Cpp/src/objgrd.cpp:76-84).
#pragma once

#include "pair_driver.hpp"

namespace lddmm {
namespace ops {

// targets per lane of the FP32 point-kernel launches (FP64: one); the split
// plans and the batched work items use 64 times these per block.  Four for
// K.a measured slower than two (larger tiles, fewer resident waves).  The
// fused adjoint takes two with the Gaussian kernel and one with the Matern
// kernels, whose longer pair term needs the registers (measured on the
// sphere->cube Matern adjoint: one 5% faster than two; on the Gaussian
// batches two 9-37% faster than one).
constexpr int kKvR32 = 2;
constexpr int bwd_r32(int order) { return order == -1 ? 2 : 1; }

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
    static constexpr bool kExpand = kCull;   // expanded form (pair_driver.hpp)
    static constexpr int kCountSlot = 0;     // executed-pair counter slot
    using KF = KernelFnT<ORDER, real, D_>;
    static constexpr bool kAnchor = KF::kScaled;   // scaled offsets (pair_driver.hpp)
    struct Args {
        const double *qi, *qj, *a;
        int ni, nj;
        real iw, iw2;
        float cut2;   // culling range (cull_r2); +inf: none
        float xr2;    // expanded-form radius bound (expand_r2)
        double kap;   // KF::kappa
        unsigned long long *exec = nullptr;   // executed-pair count (pair_counter), or none
        const double4 *grp = nullptr;   // bounds of the aligned 64-source groups, or none
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
    // expanded form: exp2(-t), t = |x_j|^2 - 2 x_i.x_j
    __device__ __forceinline__ static void pair_x(const Args &, const Tgt &, const float (&m)[D],
                                                  const float (&xj)[D + 1], const lds_t (&fj)[F], Acc &acc)
    {
        float t = xj[D];
#pragma unroll
        for (int c = 0; c < D; ++c) t = __builtin_fmaf(m[c], xj[c], t);
        const float k = __builtin_amdgcn_exp2f(-t);
#pragma unroll
        for (int c = 0; c < D; ++c) acc.v[c] += k * fj[c];
    }
    __device__ __forceinline__ static void flush_x(const Args &, const Acc &a, double (&tot)[OUT],
                                                   const crd_t (&)[D], float sc)
    {
#pragma unroll
        for (int c = 0; c < D; ++c) tot[c] += (double)a.v[c] * (double)sc;
    }
};

// ---------------------------------------------------------------------------
// Fused backward pass of the adjoint (Cpp/src/objgrd.cpp:76-84): one kernel
// evaluation per pair feeds K(q) amp and the dq sum of l^T K(q) r with
// amp = a - p, l = amp - a/2, r = a.  Outputs: [K amp (D), dq (D)].
// The dq weight of a pair, l_i.a_j + l_j.a_i, equals -p_i.a_j + a_i.amp_j,
// so a source carries a_j and amp_j (2D fields) and a target a_i and -p_i.
// R_ targets per lane (two in FP32, where registers allow).
// ---------------------------------------------------------------------------
template <int D_, int ORDER, class P, int R_ = 1>
struct OpBwd {
    static constexpr int D = D_, F = 2 * D_, OUT = 2 * D_, R = R_;
    using real = typename P::real;
    using lds_t = real;
    using crd_t = real;
    static constexpr bool kCull = P::fp32 && ORDER == -1;
    static constexpr bool kExpand = kCull;   // expanded form (pair_driver.hpp)
    static constexpr int kCountSlot = 1;     // executed-pair counter slot
    using KF = KernelFnT<ORDER, real, D_>;
    static constexpr bool kAnchor = KF::kScaled;   // scaled offsets (pair_driver.hpp)
    struct Args {
        const double *q, *a, *p;
        int n;
        real iw, iw2;
        float cut2;        // culling range (cull_r2); +inf: none
        float xr2;         // expanded-form radius bound (expand_r2)
        double kap, pds;   // KF::kappa; dq flush factor KF::cmul / kappa
        unsigned long long *exec = nullptr;   // executed-pair count (pair_counter), or none
        const double4 *grp = nullptr;   // bounds of the aligned 64-source groups, or none
    };
    __device__ __forceinline__ static double crd_scale(const Args &A) { return A.kap; }
    struct Tgt {
        real a[D], np[D];   // a_i, -p_i
    };
    struct Acc {
        real g[D], pd[D], sf;   // sf: expanded form's sum of the dq weights
    };
    __device__ __forceinline__ static void load_src(const Args &A, int j, double (&x)[D], lds_t (&f)[F])
    {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            const size_t o = (size_t)c * A.n + j;
            const double aj = A.a[o];
            x[c] = A.q[o];
            f[c] = (real)aj;
            f[D + c] = (real)(aj - A.p[o]);
        }
    }
    __device__ __forceinline__ static void load_tgt(const Args &A, int i, Tgt &t, double (&x)[D])
    {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            const size_t o = (size_t)c * A.n + (i >= 0 ? i : 0);
            x[c] = i >= 0 ? A.q[o] : 0.0;
            t.a[c] = i >= 0 ? (real)A.a[o] : real(0);
            t.np[c] = i >= 0 ? (real)(-A.p[o]) : real(0);
        }
    }
    __device__ __forceinline__ static void zero(Acc &a)
    {
#pragma unroll
        for (int c = 0; c < D; ++c) a.g[c] = a.pd[c] = real(0);
        a.sf = real(0);
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
        for (int c = 0; c < D; ++c) lr += t.np[c] * fj[c];
#pragma unroll
        for (int c = 0; c < D; ++c) lr += fj[D + c] * t.a[c];
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
            if constexpr (kAnchor) tot[D + c] += (double)a.pd[c] * A.pds;
            else tot[D + c] += a.pd[c];
        }
    }
    // expanded form: k = exp2(-t) (cf = k for the Gaussian); the dq sum
    // sum_j f (x_i - x_j) is carried as sum_j f and sum_j f x_j
    __device__ __forceinline__ static void pair_x(const Args &, const Tgt &t, const float (&m)[D],
                                                  const float (&xj)[D + 1], const lds_t (&fj)[F], Acc &acc)
    {
        float s = xj[D];
#pragma unroll
        for (int c = 0; c < D; ++c) s = __builtin_fmaf(m[c], xj[c], s);
        const float k = __builtin_amdgcn_exp2f(-s);
        float lr = 0.0f;
#pragma unroll
        for (int c = 0; c < D; ++c) lr += t.np[c] * fj[c];
#pragma unroll
        for (int c = 0; c < D; ++c) lr += fj[D + c] * t.a[c];
        const float f = lr * k;
        acc.sf += f;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            acc.g[c] += k * fj[D + c];
            acc.pd[c] += f * xj[c];
        }
    }
    __device__ __forceinline__ static void flush_x(const Args &A, const Acc &a, double (&tot)[OUT],
                                                   const crd_t (&xi)[D], float sc)
    {
        const double s = (double)sc;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            tot[c] += (double)a.g[c] * s;
            tot[D + c] += ((double)xi[c] * (double)a.sf - (double)a.pd[c]) * (s * A.pds);
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
    static constexpr bool kAnchor = KF::kScaled;   // scaled offsets (pair_driver.hpp)
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
            if constexpr (kAnchor) tot[c] += (double)a.pd[c] * A.pds;
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
    static constexpr bool kAnchor = KF::kScaled;   // scaled offsets (pair_driver.hpp)
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
            if constexpr (kAnchor) tot[c] += (double)a.pd[c] * A.pds;
            else tot[c] += a.pd[c];
        }
    }
};

}  // namespace ops
}  // namespace lddmm
