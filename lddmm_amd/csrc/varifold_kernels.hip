// varifold_kernels.hip -- varifold matching term and its vertex gradient.
//
// V(q) = sum_{i,j in D} kc ku a_i a_j - 2 sum_{i in D, j in T} kc ku a_i a_j
//        + sum_{i,j in T} kc ku a_i a_j
// over element centres c, unit directions u and volumes a
// (Cpp/src/varifold3D.cpp:259-357, Cpp/src/varifold2D.cpp:126-216).
// Position kernels kc: 'G' exp(-r^2/s^2), 'C' 1/(1+r^2/s^2); orientation
// kernels ku(ang = u_i.u_j): 'B' ang^2, 'L' ang, 'O' exp(2(ang-1)/s^2),
// 'U' exp(2(ang^2-1)/s^2) (Cpp/src/varifold3D.cpp:41-188).
//
// The pair pass (the tiled driver of pair_driver.hpp) accumulates, per
// deformed element i, four sums from which the reference's per-pair
// gradient terms factor out exactly:
//   E   = sum_j kc ku a_j                  -> energy a_i E
//   G   = sum_j kc^p ku a_j (c_i - c_j)    -> dc_i = -4 a_i G / s^2 (p = 1 'G', 2 'C')
//   U   = sum_j kc a_j ku'(ang) u_j
//   S   = sum_j kc a_j (ku - ku'(ang) ang) -> dd_i = 2 dv (U + S u_i)
// with dv = d a / d |n| = 1/2 in 3D (|n| = 2a) and 1 in 2D, which is the
// reference's (ku' / |n|)(u_j - ang u_i) a_i + ku dv u_i summed over j.
// The deformed-deformed and deformed-target sources are two launches
// writing disjoint split ranges; the finalize kernel applies the -2 / -1
// weights of the cross term.  The target-target term does not depend on q;
// it is evaluated once per problem, in FP64, and cached.
#include "kernels.hpp"
#include "pair_driver.hpp"

namespace lddmm {

namespace {

enum { kCenG = 0, kCenC = 1 };
enum { kDirB = 0, kDirL = 1, kDirO = 2, kDirU = 3 };

template <int CEN, typename real>
__device__ __forceinline__ real cen_value(real r2, real inv_cs2)
{
    if constexpr (CEN == kCenG) {
        if constexpr (sizeof(real) == 4) return __builtin_amdgcn_exp2f(-r2 * inv_cs2 * 1.4426950408889634f);
        else return exp(-r2 * inv_cs2);
    } else {
        return real(1) / (real(1) + r2 * inv_cs2);
    }
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
    struct Args {
        const double *ci, *ui;   // targets: deformed element centres / directions
        int ni;
        const double *cj, *uj, *aj;   // sources
        int nj;
        real ics2, ids2;
    };
    struct Tgt {
        double c[D];
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
    __device__ __forceinline__ static void load_tgt(const Args &A, int i, Tgt &t)
    {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            t.c[c] = i >= 0 ? A.ci[(size_t)c * A.ni + i] : 0.0;
            t.u[c] = i >= 0 ? (real)A.ui[(size_t)c * A.ni + i] : real(0);
        }
    }
    __device__ __forceinline__ static void zero(Acc &a)
    {
        a.e = a.s = real(0);
#pragma unroll
        for (int c = 0; c < D; ++c) a.g[c] = a.u[c] = real(0);
    }
    __device__ __forceinline__ static void pair(const Args &A, const Tgt &t, const double (&xj)[D],
                                                const lds_t (&fj)[F], Acc &acc)
    {
        real df[D], r2 = real(0), ang = real(0);
#pragma unroll
        for (int c = 0; c < D; ++c) {
            df[c] = (real)(t.c[c] - xj[c]);
            r2 += df[c] * df[c];
            ang += t.u[c] * fj[c];
        }
        const real kc = cen_value<CEN, real>(r2, A.ics2);
        real dku;
        const real ku = dir_value<DIR, real>(ang, A.ids2, dku);
        const real w = kc * fj[D];
        const real wk = w * ku;
        const real gw = (CEN == kCenG) ? wk : wk * kc;
        const real wd = w * dku;
        acc.e += wk;
        acc.s += wk - wd * ang;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            acc.g[c] += gw * df[c];
            acc.u[c] += wd * fj[c];
        }
    }
    __device__ __forceinline__ static void flush(const Acc &a, double (&tot)[OUT])
    {
        tot[0] += (double)a.e;
        tot[1] += (double)a.s;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            tot[2 + c] += (double)a.g[c];
            tot[2 + D + c] += (double)a.u[c];
        }
    }
};

// value-only: sum_j kc ku a_j (target-target term)
template <int D_, int CEN, int DIR>
struct OpVfValue {
    static constexpr int D = D_, F = D_ + 1, OUT = 1, R = 1;
    using real = double;
    using lds_t = double;
    struct Args {
        const double *c, *u, *a;
        int n;
        double ics2, ids2;
    };
    struct Tgt {
        double c[D], u[D];
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
    __device__ __forceinline__ static void load_tgt(const Args &A, int i, Tgt &t)
    {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            t.c[c] = i >= 0 ? A.c[(size_t)c * A.n + i] : 0.0;
            t.u[c] = i >= 0 ? A.u[(size_t)c * A.n + i] : 0.0;
        }
    }
    __device__ __forceinline__ static void zero(Acc &a) { a.e = 0.0; }
    __device__ __forceinline__ static void pair(const Args &A, const Tgt &t, const double (&xj)[D],
                                                const lds_t (&fj)[F], Acc &acc)
    {
        double r2 = 0.0, ang = 0.0;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            double df = t.c[c] - xj[c];
            r2 += df * df;
            ang += t.u[c] * fj[c];
        }
        double dku;
        acc.e += cen_value<CEN, double>(r2, A.ics2) * dir_value<DIR, double>(ang, A.ids2, dku) * fj[D];
    }
    __device__ __forceinline__ static void flush(const Acc &a, double (&tot)[OUT]) { tot[0] += a.e; }
};

template <int D>
__global__ __launch_bounds__(kBlock) void k_elm_geometry(const double *__restrict__ q, int nq,
                                                         const int *__restrict__ elm, int ne,
                                                         double *__restrict__ cen, double *__restrict__ dir,
                                                         double *__restrict__ vol)
{
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= ne) return;
    if constexpr (D == 2) {
        const int i0 = elm[e], i1 = elm[ne + e];
        double x0 = q[i0], y0 = q[nq + i0], x1 = q[i1], y1 = q[nq + i1];
        double ux = x1 - x0, uy = y1 - y0;
        double a = sqrt(ux * ux + uy * uy);
        cen[e] = 0.5 * (x0 + x1);
        cen[ne + e] = 0.5 * (y0 + y1);
        dir[e] = ux / a;
        dir[ne + e] = uy / a;
        vol[e] = a;
    } else {
        const int i0 = elm[e], i1 = elm[ne + e], i2 = elm[2 * ne + e];
        double p0[3], p1[3], p2[3], e1[3], e2[3], n[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            p0[c] = q[(size_t)c * nq + i0];
            p1[c] = q[(size_t)c * nq + i1];
            p2[c] = q[(size_t)c * nq + i2];
            e1[c] = p1[c] - p0[c];
            e2[c] = p2[c] - p0[c];
            cen[(size_t)c * ne + e] = (p0[c] + p1[c] + p2[c]) / 3.0;
        }
        n[0] = e1[1] * e2[2] - e1[2] * e2[1];
        n[1] = -e1[0] * e2[2] + e1[2] * e2[0];
        n[2] = e1[0] * e2[1] - e1[1] * e2[0];
        double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) dir[(size_t)c * ne + e] = n[c] / nn;
        vol[e] = 0.5 * nn;
    }
}

constexpr int kRedBlock = 64;

template <int D>
__global__ __launch_bounds__(kRedBlock) void k_vf_finalize(int sd, int splits, int ne,
                                                           const double *__restrict__ part,
                                                           const double *__restrict__ ddir,
                                                           const double *__restrict__ dvol, double gscale,
                                                           double *__restrict__ dc, double *__restrict__ dd,
                                                           double *__restrict__ eparts)
{
    constexpr int NC = 2 * D + 2;
    const int i = blockIdx.x * kRedBlock + threadIdx.x;
    double e = 0.0;
    if (i < ne) {
        double acc[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c] = 0.0;
        for (int s = 0; s < splits; ++s) {
            const double *o = part + (size_t)s * NC * ne + i;
            double t[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) t[c] = o[(size_t)c * ne];
            // cross term: energy weight -2, gradient sign -1
            const double we = s < sd ? 1.0 : -2.0, wg = s < sd ? 1.0 : -1.0;
            acc[0] += we * t[0];
#pragma unroll
            for (int c = 1; c < NC; ++c) acc[c] += wg * t[c];
        }
        const double a = dvol[i];
        const double dv = (D == 3) ? 0.5 : 1.0;
        e = a * acc[0];
#pragma unroll
        for (int c = 0; c < D; ++c) {
            dc[(size_t)c * ne + i] = gscale * a * acc[2 + c];
            dd[(size_t)c * ne + i] = 2.0 * dv * (acc[2 + D + c] + acc[1] * ddir[(size_t)c * ne + i]);
        }
    }
    e = wave_sum(e);
    if (threadIdx.x == 0) eparts[blockIdx.x] = e;
}

__global__ __launch_bounds__(kRedBlock) void k_vf_value_finalize(int splits, int n, const double *__restrict__ part,
                                                                 const double *__restrict__ vol,
                                                                 double *__restrict__ eparts)
{
    const int i = blockIdx.x * kRedBlock + threadIdx.x;
    double e = 0.0;
    if (i < n) {
        double s = 0.0;
        for (int k = 0; k < splits; ++k) s += part[(size_t)k * n + i];
        e = vol[i] * s;
    }
    e = wave_sum(e);
    if (threadIdx.x == 0) eparts[blockIdx.x] = e;
}

// Vertex gather (Cpp/src/varifold3D.cpp:359-391, Cpp/src/varifold2D.cpp:218-240):
// 3D dq_n += dc_e / 3 + dd_e x (q_a - q_b); 2D dq_n += dc_e / 2 + s dd_e.
template <int D>
__global__ __launch_bounds__(kBlock) void k_vf_gather(const double *__restrict__ q, int nq,
                                                      const int *__restrict__ ifo, const double *__restrict__ dc,
                                                      const double *__restrict__ dd, int ne, double scale,
                                                      double *__restrict__ out)
{
    const int n = blockIdx.x * kBlock + threadIdx.x;
    if (n >= nq) return;
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
            double ab[3], d[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                ab[c] = q[(size_t)c * nq + ia] - q[(size_t)c * nq + ib];
                d[c] = dd[(size_t)c * ne + e];
            }
            double x0 = d[1] * ab[2] - d[2] * ab[1];
            double x1 = -d[0] * ab[2] + d[2] * ab[0];
            double x2 = d[0] * ab[1] - d[1] * ab[0];
            g[0] += dc[e] / 3.0 + x0;
            g[1] += dc[ne + e] / 3.0 + x1;
            g[2] += dc[2 * (size_t)ne + e] / 3.0 + x2;
        }
    }
#pragma unroll
    for (int c = 0; c < D; ++c) out[(size_t)c * nq + n] = scale * g[c];
}

int cen_code(char t)
{
    if (t == 'G') return kCenG;
    if (t == 'C') return kCenC;
    throw Error(kErrArg, "cenKnlType must be 'G' or 'C'");
}

int dir_code(char t)
{
    switch (t) {
    case 'B': return kDirB;
    case 'L': return kDirL;
    case 'O': return kDirO;
    case 'U': return kDirU;
    }
    throw Error(kErrArg, "dirKnlType must be 'B', 'L', 'O' or 'U'");
}

// f(Dim, Cen, Dir)
template <typename F>
void dispatch_vf(int dim, const VfParams &vp, F &&f)
{
    if (dim != 2 && dim != 3) throw Error(kErrArg, "dimension must be 2 or 3");
    const int cc = cen_code(vp.cen_type), dc = dir_code(vp.dir_type);
    auto with_dim = [&](auto C, auto U) {
        if (dim == 2) f(std::integral_constant<int, 2>{}, C, U);
        else f(std::integral_constant<int, 3>{}, C, U);
    };
    auto with_dir = [&](auto C) {
        switch (dc) {
        case kDirB: with_dim(C, std::integral_constant<int, kDirB>{}); break;
        case kDirL: with_dim(C, std::integral_constant<int, kDirL>{}); break;
        case kDirO: with_dim(C, std::integral_constant<int, kDirO>{}); break;
        default: with_dim(C, std::integral_constant<int, kDirU>{}); break;
        }
    };
    if (cc == kCenG) with_dir(std::integral_constant<int, kCenG>{});
    else with_dir(std::integral_constant<int, kCenC>{});
}

double inv_sq(double w) { return w != 0.0 ? 1.0 / (w * w) : 0.0; }

}  // namespace

PairSplit plan_vf(int ni, int nj, int target_blocks) { return make_split_ti(ni, nj, 64, target_blocks); }

void launch_elm_geometry(int dim, const double *q, int nq, const int *elm, int ne, double *cen, double *dir,
                         double *vol, hipStream_t st)
{
    if (ne <= 0) return;
    if (dim == 2) k_elm_geometry<2><<<ceil_div(ne, kBlock), kBlock, 0, st>>>(q, nq, elm, ne, cen, dir, vol);
    else k_elm_geometry<3><<<ceil_div(ne, kBlock), kBlock, 0, st>>>(q, nq, elm, ne, cen, dir, vol);
    LDDMM_LAUNCH_CHECK();
}

void launch_vf_pairs(int dim, int prec, const VfParams &vp, const double *dcen, const double *ddir, const double *dvol,
                     int ne, const double *tcen, const double *tdir, const double *tvol, int nt, const PairSplit &spd,
                     const PairSplit &spt, double *part, hipStream_t st)
{
    if (ne <= 0) return;
    if (prec != kFP64 && prec != kFP32) throw Error(kErrArg, "unknown precision policy");
    dispatch_vf(dim, vp, [&](auto Dc, auto Cc, auto Uc) {
        auto go = [&](auto Pc) {
            using Op = OpVf<Dc(), Cc(), Uc(), decltype(Pc)>;
            using real = typename Op::real;
            const real ics2 = (real)inv_sq(vp.cen_width), ids2 = (real)inv_sq(vp.dir_width);
            typename Op::Args dd{dcen, ddir, ne, dcen, ddir, dvol, ne, ics2, ids2};
            launch_pairs<Op>(dd, ne, ne, spd, 0, part, st);
            if (nt > 0) {
                typename Op::Args dt{dcen, ddir, ne, tcen, tdir, tvol, nt, ics2, ids2};
                launch_pairs<Op>(dt, ne, nt, spt, spd.splits, part, st);
            }
        };
        if (prec == kFP32) go(PrecF32{});
        else go(PrecF64{});
    });
}

int vf_finalize_blocks(int ne) { return ceil_div(ne, kRedBlock); }

void launch_vf_finalize(int dim, const VfParams &vp, int sd, int splits, int ne, const double *part,
                        const double *ddir, const double *dvol, double *dc, double *dd, double *eparts,
                        hipStream_t st)
{
    if (ne <= 0) return;
    // dc_i = 2 a_i sum_j gk ku a_j (c_i - c_j) with gk = -2 kc^p / s^2
    const double gscale = -4.0 * inv_sq(vp.cen_width);
    const int nb = vf_finalize_blocks(ne);
    if (dim == 2)
        k_vf_finalize<2><<<nb, kRedBlock, 0, st>>>(sd, splits, ne, part, ddir, dvol, gscale, dc, dd, eparts);
    else k_vf_finalize<3><<<nb, kRedBlock, 0, st>>>(sd, splits, ne, part, ddir, dvol, gscale, dc, dd, eparts);
    LDDMM_LAUNCH_CHECK();
}

void launch_vf_value_pairs(int dim, const VfParams &vp, const double *cen, const double *dir, const double *vol,
                           int n, const PairSplit &sp, double *part, hipStream_t st)
{
    if (n <= 0) return;
    dispatch_vf(dim, vp, [&](auto Dc, auto Cc, auto Uc) {
        using Op = OpVfValue<Dc(), Cc(), Uc()>;
        typename Op::Args A{cen, dir, vol, n, inv_sq(vp.cen_width), inv_sq(vp.dir_width)};
        launch_pairs<Op>(A, n, n, sp, 0, part, st);
    });
}

int vf_value_finalize_blocks(int n) { return ceil_div(n, kRedBlock); }

void launch_vf_value_finalize(int splits, int n, const double *part, const double *vol, double *eparts,
                              hipStream_t st)
{
    if (n <= 0) return;
    k_vf_value_finalize<<<vf_value_finalize_blocks(n), kRedBlock, 0, st>>>(splits, n, part, vol, eparts);
    LDDMM_LAUNCH_CHECK();
}

void launch_vf_gather(int dim, const double *q, int nq, const int *ifo, int ifo_rows, const double *dc,
                      const double *dd, int ne, double scale, double *out, hipStream_t st)
{
    (void)ifo_rows;
    if (nq <= 0) return;
    if (dim == 2) k_vf_gather<2><<<ceil_div(nq, kBlock), kBlock, 0, st>>>(q, nq, ifo, dc, dd, ne, scale, out);
    else k_vf_gather<3><<<ceil_div(nq, kBlock), kBlock, 0, st>>>(q, nq, ifo, dc, dd, ne, scale, out);
    LDDMM_LAUNCH_CHECK();
}

}  // namespace lddmm
