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
#include "varifold_ops.hpp"

namespace lddmm {

namespace {

using namespace ops;

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
                     const PairSplit &spt, double *part, hipStream_t st, const double4 *grp_d, const double4 *grp_t)
{
    if (ne <= 0) return;
    if (prec != kFP64 && prec != kFP32) throw Error(kErrArg, "unknown precision policy");
    dispatch_vf(dim, vp, [&](auto Dc, auto Cc, auto Uc) {
        auto go = [&](auto Pc) {
            using Op = OpVf<Dc(), Cc(), Uc(), decltype(Pc)>;
            using real = typename Op::real;
            const real ics2 = (real)inv_sq(vp.cen_width), ids2 = (real)inv_sq(vp.dir_width);
            const float cut2 = cull_r2(pair_cull_bits()), xr2 = expand_r2(pair_expansion());
            const double kap = cen_kappa<Cc()>(inv_sq(vp.cen_width));
            typename Op::Args dd{dcen, ddir, ne, dcen, ddir, dvol, ne, ics2, ids2, cut2, xr2, kap, pair_counter(), grp_d};
            launch_pairs<Op>(dd, ne, ne, spd, 0, part, st);
            if (nt > 0) {
                typename Op::Args dt{dcen, ddir, ne, tcen, tdir, tvol, nt, ics2, ids2, cut2, xr2, kap, pair_counter(), grp_t};
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
