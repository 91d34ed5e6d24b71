// pair_driver.hpp -- the one tiled all-pairs loop every pair kernel uses.
//
// A block of 256 threads (4 wave64) owns TI = 64*R targets (lane l of
// every wave holds targets base + l + 64 r, r < R) and one chunk of the
// sources.  The chunk is staged through LDS 256 sources at a time, one
// 16-byte aligned record per source: the per-source fields in the policy's
// storage type, then the coordinates (crd_t, below), so one base address and
// immediate offsets reach every field (wide ds_read_b128/b64 loads, no
// per-field address arithmetic in the pair loop).  Wave w consumes the w-th
// contiguous quarter of each tile, so all 64 lanes of a wave read the same
// LDS words (broadcast, conflict-free) and every source is read once per
// wave for R targets.  Per tile, sums are
// carried in the policy's arithmetic type and flushed to FP64 totals
// (two-level accumulation: at most 64 pairs are summed in FP32 before the
// FP64 add); at the end the four waves' totals are combined in LDS in a
// fixed order and the block writes one partial per (split, component,
// target), which a following kernel reduces over the splits in a fixed
// order.  No atomics: results are bitwise repeatable.
//
// Coordinates reach the pair term as `crd_t`.  FP64 operators use the FP64
// positions.  FP32 operators use FP32 offsets from an anchor a, the FP64
// centroid of the block's targets: targets are converted once, sources
// once per block when they are staged, so a pair forms its D differences
// with D FP32 subtractions (instead of D FP64 subtractions and D
// conversions).  The rounding error of a difference is then
// 2^-25 (|x_i - a| + |x_j - a| + |x_i - x_j|) per component instead of
// 2^-25 |x_i - x_j|; with the spatially ordered problems of Objective and
// Batch (compact 64/128-point tiles) |x_i - a| is at most a tile radius.
// The offsets are scaled by the kernel's kappa (KernelFnT) in FP64 before
// the rounding, which changes none of these bounds.
//
// An operator Op supplies
//   D, F, OUT, R                       dimensions, LDS fields, outputs, targets/lane
//   Args                               kernel arguments
//   lds_t, crd_t, Tgt, Acc             LDS field type, coordinate type, per-target
//                                      registers, per-tile sums
//   load_src(args, j, x[D], f[F])      source j's FP64 coordinates and fields
//   load_tgt(args, i, Tgt&, x[D])      target i's registers and FP64 coordinates
//                                      (i < 0: padding lane, coordinates 0)
//   pair(args, Tgt&, xi[D], xj[D], fj[F], Acc&)   one pair term
//   zero(Acc&), flush(args, const Acc&, double tot[OUT])
//   crd_scale(args)                    FP32 only: kappa of the staged offsets
//   kCull, args.cut2                   culling (below)
//
// Executed pairs: with args.exec set (device counters, pair_counter() in
// kernels.hpp), every culling operator adds the pairs of each source range it
// does not skip to its slot Op::kCountSlot, so pair rates can count computed
// pairs only.
//
// Culling (FP32 Gaussian kernels, kCull): a wave skips its range of staged
// sources when every kernel value exp2(-r2) (scaled coordinates, KernelFnT)
// between them and the block's targets is below 2^-48, i.e. r2 > 48.  Per
// block the driver takes the target tile's radius about the anchor (rt) and
// per wave's range the smallest distance of a source from the anchor (ms);
// when ms - rt exceeds sqrt(48) (with a 1e-4 relative margin for the FP32
// rounding of both bounds and of r2: args.cut2 = 48 (1 + 1e-4)), no pair of
// the two can reach it.  Every term of a pair is a kernel value times factors
// of the problem's own scale, so a skipped range changes a target's sum by
// less than 2^-48 times the sum of its terms' factors: 24 binary orders below
// the FP32 unit roundoff of the sums, and in practice at the FP64 rounding of
// the totals (culled and full sums agree to a few FP64 ulps).  With the
// spatially ordered problems this removes the far tiles of narrow kernels
// (the varifold's Gaussian of width 0.1 on unit-sized shapes skips most of
// its pairs).
//
// Expanded form (FP32 Gaussian operators with kExpand): with n_i = |x_i|^2
// and t = |x_j|^2 - 2 x_i.x_j, r2 = n_i + t and exp2(-r2) =
// exp2(-n_i) exp2(-t).  The staged record carries |x_j|^2 after the
// coordinates, the lanes hold m = -2 x_i, a pair forms t with D FMAs
// (instead of D subtractions and D multiply-adds) and the operator
// applies exp2(-n_i) once per flush:
//   pair_x(args, Tgt&, m[D], xj[D+1], fj[F], Acc&)
//   flush_x(args, const Acc&, double tot[OUT], xi[D], exp2(-n_i))
// Differences x_i - x_j in the pair terms become x_i sum_j w - sum_j w x_j,
// completed at the flush.  The FP32 rounding of t is relative to
// |x_j|^2 + 2 |x_i||x_j| rather than to r2, so the expanded form is taken
// only by blocks of small target radius, rt^2 <= args.xr2 = 8 (the kernel
// falls to 1/256 over that radius): there its error stays within a small
// factor of the direct form's, and exp2(-t) <= exp2(n_i) <= 256.  Wider
// tiles (small problems, narrow kernels) keep the direct form above.
// Culling in the expanded form bounds t >= ms (ms - 2 rt) instead (then
// exp2(-t) < 2^-48, and the factor exp2(-n_i) <= 1 keeps the bound).
#pragma once

#include "common.hpp"

namespace lddmm {

// Precision policies.  Coordinates are FP64 in HBM; the pair arithmetic
// and the coordinates staged in LDS are in `real` (FP32 as anchored
// offsets, above).
struct PrecF64 {
    using real = double;
    static constexpr bool fp32 = false;
};
struct PrecF32 {
    using real = float;
    static constexpr bool fp32 = true;
};

// Kernel family in the policy's arithmetic.  The pair terms see the squared
// distance r2 of coordinates scaled by kappa(iw, iw2) (the driver scales the
// staged FP32 offsets; FP64 coordinates are unscaled, kappa = 1), and
//   value(r2, iw, iw2) = k,   eval(r2, iw, iw2, k, cf)
// with the gradient factor c of grad_{q_i} k = c (q_i - q_j) equal to
// cmul(iw, iw2) * cf.  Operators carry cmul / kappa into the FP64 flush of
// their sums of cf (q_i - q_j)_scaled, so neither the width nor the scale
// costs an instruction per pair.  D selects the 2D Matern family (Bessel
// factors).
template <int ORDER, typename real, int D = 3>
struct KernelFnT;

// FP64, Gaussian (2D, 3D) and 3D Matern: the coordinates are scaled as in
// FP32 below (anchored offsets times kappa, formed in FP64), so the kernel is
// k = 2^-r2 (Gaussian) or a polynomial in s = sqrt(r2) times 2^-s (Matern),
// with exp2_neg in place of the library exp and no per-pair multiply by the
// width; the gradient factors' constants move to the flush (cmul / kappa).
// Rounding: a scaled difference carries 2^-53 (|x_i - a| + |x_j - a|) kappa
// per component instead of 2^-53 |x_i - x_j|; 2^-s is about 1 ulp.  FP64,
// 2D Matern: the reference's formulas on unscaled coordinates (kappa = cmul
// = 1; the Bessel factors need d itself).
template <int ORDER, int D>
struct KernelFnT<ORDER, double, D> {
    static constexpr bool kScaled = !(D == 2 && ORDER >= 0);
    static constexpr double kLog2e = 1.4426950408889634, kLn2 = 0.6931471805599453;
    static constexpr double L1 = kLn2, L2 = kLn2 * kLn2, L3 = kLn2 * kLn2 * kLn2, L4 = kLn2 * kLn2 * kLn2 * kLn2;
    __host__ __device__ static double kappa(double iw, double iw2)
    {
        if constexpr (!kScaled) return 1.0;
        else return ORDER == -1 ? sqrt(iw2 * kLog2e) : iw * kLog2e;
    }
    __host__ __device__ static double cmul(double, double iw2)
    {
        if constexpr (!kScaled) return 1.0;
        switch (ORDER) {
        case -1: return -2.0 * iw2;
        case 0: return 0.0;   // order 0 has no dq (Cpp/src/dqKernel.cpp:39)
        case 1: return -iw2;
        case 2: return -iw2 / 3.0;
        case 3: return -iw2 * 3.0 / 15.0;
        default: return -iw2 * 15.0 / 105.0;
        }
    }
    // P_1 .. P_4 of d = s ln 2 (Cpp/src/computeKernel.cpp:170-234)
    __device__ __forceinline__ static double poly(int p, double s)
    {
        switch (p) {
        case 1: return fma(s, L1, 1.0);
        case 2: return fma(s, fma(s, L2 * (1.0 / 3.0), L1), 1.0);
        case 3: return fma(s, fma(s, fma(s, L3 * (1.0 / 15.0), L2 * (6.0 / 15.0)), L1), 1.0);
        default:
            return fma(s, fma(s, fma(s, fma(s, L4 * (1.0 / 105.0), L3 * (10.0 / 105.0)), L2 * (45.0 / 105.0)), L1),
                       1.0);
        }
    }
    __device__ __forceinline__ static double value(double r2, double iw, double iw2)
    {
        if constexpr (!kScaled) {
            return KernelFnD<D, ORDER>::value(r2, iw, iw2);
        } else if constexpr (ORDER == -1) {
            return exp2_neg(r2);
        } else {
            const double s = sqrt_r2(r2), e = exp2_neg(s);
            if constexpr (ORDER == 0) return e;
            else return poly(ORDER, s) * e;
        }
    }
    __device__ __forceinline__ static void eval(double r2, double iw, double iw2, double &k, double &cf)
    {
        if constexpr (!kScaled) {
            KernelFnD<D, ORDER>::eval(r2, iw, iw2, k, cf);
        } else if constexpr (ORDER == -1) {
            k = exp2_neg(r2);
            cf = k;
        } else {
            const double s = sqrt_r2(r2), e = exp2_neg(s);
            if constexpr (ORDER == 0) {
                k = e;
                cf = 0.0;
            } else if constexpr (ORDER == 1) {
                k = poly(1, s) * e;
                cf = e;
            } else {
                k = poly(ORDER, s) * e;
                cf = poly(ORDER - 1, s) * e;
            }
        }
    }
};

// FP32 Bessel factors p0 = x^2 K0(x), p1 = x K1(x) of the 2D Matern kernels.
// 0 < x <= 1: the ascending series (A&S 9.6.13, 9.6.11) in t = x^2/4,
//   K0 = -ln(x/2) I0 + sum_k psi(k+1) t^k/(k!)^2,
//   x K1 = 1 + ln(x/2) x I1 - t sum_k (psi(k+1) + psi(k+2)) t^k/(k!(k+1)!),
// x I1 = 2t sum_k t^k/(k!(k+1)!), through k = 5 (the next terms are below
// 1e-9 at t = 1/4); the coefficients are those closed forms, rounded to FP32.
// x > 1: p1 = sqrt(x) f1(u) e^-x and p0 = x sqrt(x) f0(u) e^-x with
// f_n(u) = sqrt(x) e^x K_n(x), u = 1/x, polynomials of degree 10 in u fitted
// by lddmm_amd/fit_bessel_f32.py (relative error 5e-9; 9e-8 with FP32
// Horner).  x below 2^-60 gives the limits p0 = 0, p1 = 1.  Against scipy
// over x in (0, 90], absolute error <= 1.7e-7 (tests/test_bessel_f32.py).
__device__ __forceinline__ void bessel_p01_f32(float x, float &p0, float &p1)
{
    if (x <= 1.0f) {
        const float t = 0.25f * x * x;
        const float lg = (__builtin_amdgcn_logf(x) - 1.0f) * 0.693147181f;   // ln(x/2)
        const float i0 = 1.0f + t * (1.0f + t * (2.5e-1f + t * (2.777777778e-2f + t * (1.736111111e-3f +
                                                                                      t * 6.944444444e-5f))));
        const float s0 = -5.772156649e-1f +
                         t * (4.227843351e-1f + t * (2.306960838e-1f + t * (3.489215746e-2f + t * (2.614787619e-3f +
                                                                                                 t * 1.184803936e-4f))));
        const float i1 = 1.0f + t * (5.0e-1f + t * (8.333333333e-2f + t * (6.944444444e-3f + t * (3.472222222e-4f +
                                                                                              t * 1.157407407e-5f))));
        const float s1 = -1.544313298e-1f +
                         t * (6.727843351e-1f + t * (1.815751670e-1f + t * (1.918218984e-2f + t * (1.115359492e-3f +
                                                                                                 t * 4.142247689e-5f))));
        const bool tiny = !(x >= 8.673617380e-19f);   // 2^-60, and x = 0 (log2 0 = -inf)
        p0 = tiny ? 0.0f : 4.0f * t * (s0 - lg * i0);
        p1 = tiny ? 1.0f : 1.0f + t * (2.0f * lg * i1 - s1);
    } else {
        const float u = __builtin_amdgcn_rcpf(x);
        const float f0 =
            1.253314137e+00f +
            u * (-1.566625983e-01f +
                 u * (8.804495633e-02f +
                      u * (-9.030350298e-02f +
                           u * (1.252581924e-01f +
                                u * (-1.841103584e-01f +
                                     u * (2.345902026e-01f +
                                          u * (-2.255823016e-01f +
                                               u * (1.470358670e-01f + u * (-5.692518502e-02f + u * 9.803697467e-03f)))))))));
        const float f1 =
            1.253314137e+00f +
            u * (4.699910283e-01f +
                 u * (-1.467864066e-01f +
                      u * (1.268443316e-01f +
                           u * (-1.633526385e-01f +
                                u * (2.327853292e-01f +
                                     u * (-2.936871648e-01f +
                                          u * (2.822842300e-01f +
                                               u * (-1.846503466e-01f + u * (7.185639441e-02f + u * -1.244540792e-02f)))))))));
        const float se = __builtin_amdgcn_sqrtf(x) * __builtin_amdgcn_exp2f(-x * 1.442695041f);
        p1 = se * f1;
        p0 = x * se * f0;
    }
}

// FP32, 2D Matern (Cpp/src/computeKernel.cpp:60-149, Cpp/src/dqKernel.cpp:
// 91-185): coordinates scaled by kappa = 1/sigma, so d = sqrt(r2) and
// d^2 = r2; k and the gradient factor from bessel_p01_f32, the constant of
// each order's gradient factor (and -1/sigma^2) moved into cmul.
template <int ORDER>
struct KernelFnT<ORDER, float, 2> {
    static constexpr bool kScaled = true;
    __host__ __device__ static double kappa(double iw, double) { return iw; }
    __host__ __device__ static double cmul(double, double iw2)
    {
        switch (ORDER) {
        case 0: return 0.0;   // order 0 has no dq (Cpp/src/dqKernel.cpp:39)
        case 1: return -0.5 * iw2;
        case 2: return -0.125 * iw2;
        case 3: return -iw2 / 48.0;
        default: return -iw2 / 384.0;
        }
    }
    __device__ __forceinline__ static void eval(float r2, float, float, float &k, float &cf)
    {
        float p0, p1;
        bessel_p01_f32(__builtin_amdgcn_sqrtf(r2), p0, p1);
        if constexpr (ORDER == 0) {
            k = p1;
            cf = 0.0f;
        } else if constexpr (ORDER == 1) {
            k = 0.5f * p0 + p1;
            cf = p1;
        } else if constexpr (ORDER == 2) {
            k = 0.5f * p0 + (1.0f + 0.125f * r2) * p1;
            cf = p0 + 2.0f * p1;
        } else if constexpr (ORDER == 3) {
            k = ((24.0f + r2) * p0 + (48.0f + 8.0f * r2) * p1) * (1.0f / 48.0f);
            cf = 4.0f * p0 + (8.0f + r2) * p1;
        } else {
            k = ((192.0f + 12.0f * r2) * p0 + (384.0f + r2 * (72.0f + r2)) * p1) * (1.0f / 384.0f);
            cf = (24.0f + r2) * p0 + (48.0f + 8.0f * r2) * p1;
        }
    }
    __device__ __forceinline__ static float value(float r2, float iw, float iw2)
    {
        float k, cf;
        eval(r2, iw, iw2, k, cf);
        return k;
    }
};

// FP32, Gaussian and 3D Matern: exp via v_exp_f32 on a base-2 argument,
// sqrt via v_sqrt_f32 (both ~1 ulp).  Gaussian: kappa^2 = iw2 log2 e, so
// k = exp2(-r2) and cf = k (cmul = -2 iw2).  Matern order p:
// kappa = iw log2 e, s = sqrt(r2) = d log2 e with d = |q_i - q_j| iw,
// e = exp2(-s) = e^-d, k = P_p(d) e and cf = P_{p-1}(d) e with the
// polynomials of Cpp/src/computeKernel.cpp:170-234 and
// Cpp/src/dqKernel.cpp:215-291 written in s (coefficients times ln 2^m).
template <int ORDER>
struct KernelFnT<ORDER, float, 3> {
    static constexpr bool kScaled = true;
    static constexpr double kLog2e = 1.4426950408889634, kLn2 = 0.6931471805599453;
    static constexpr float L1 = (float)kLn2, L2 = (float)(kLn2 * kLn2), L3 = (float)(kLn2 * kLn2 * kLn2),
                           L4 = (float)(kLn2 * kLn2 * kLn2 * kLn2);
    __host__ __device__ static double kappa(double iw, double iw2)
    {
        return ORDER == -1 ? sqrt(iw2 * kLog2e) : iw * kLog2e;
    }
    __host__ __device__ static double cmul(double, double iw2)
    {
        switch (ORDER) {
        case -1: return -2.0 * iw2;
        case 0: return 0.0;   // order 0 has no dq (Cpp/src/dqKernel.cpp:39)
        case 1: return -iw2;
        case 2: return -iw2 / 3.0;
        case 3: return -iw2 * 3.0 / 15.0;
        default: return -iw2 * 15.0 / 105.0;
        }
    }
    // P_1 .. P_4 of d = s ln 2
    __device__ __forceinline__ static float poly(int p, float s)
    {
        switch (p) {
        case 1: return 1.0f + s * L1;
        case 2: return 1.0f + s * (L1 + s * (L2 * (1.0f / 3.0f)));
        case 3: return 1.0f + s * (L1 + s * (L2 * (6.0f / 15.0f) + s * (L3 * (1.0f / 15.0f))));
        default:
            return 1.0f +
                   s * (L1 + s * (L2 * (45.0f / 105.0f) + s * (L3 * (10.0f / 105.0f) + s * (L4 * (1.0f / 105.0f)))));
        }
    }
    __device__ __forceinline__ static float value(float r2, float, float)
    {
        if constexpr (ORDER == -1) {
            return __builtin_amdgcn_exp2f(-r2);
        } else {
            const float s = __builtin_amdgcn_sqrtf(r2);
            const float e = __builtin_amdgcn_exp2f(-s);
            if constexpr (ORDER == 0) return e;
            else return poly(ORDER, s) * e;
        }
    }
    __device__ __forceinline__ static void eval(float r2, float, float, float &k, float &cf)
    {
        if constexpr (ORDER == -1) {
            k = __builtin_amdgcn_exp2f(-r2);
            cf = k;
        } else {
            const float s = __builtin_amdgcn_sqrtf(r2);
            const float e = __builtin_amdgcn_exp2f(-s);
            if constexpr (ORDER == 0) {
                k = e;
                cf = 0.0f;
            } else if constexpr (ORDER == 1) {
                k = poly(1, s) * e;
                cf = e;
            } else {
                k = poly(ORDER, s) * e;
                cf = poly(ORDER - 1, s) * e;
            }
        }
    }
};

// the Gaussian is the same function in 2D and 3D
template <>
struct KernelFnT<-1, float, 2> : KernelFnT<-1, float, 3> {
};

constexpr int kWaves = kBlock / 64;

// args.cut2 of a culling operator (FP32 Gaussian on coordinates scaled so
// that k = exp2(-r2)): skipped pairs have k < 2^-bits (48 by default,
// pair_cull_bits); a 1e-4 relative margin covers the FP32 rounding of the
// bounds; bits <= 0 gives +inf, culling off
__host__ __device__ inline float cull_r2(int bits)
{
    return bits > 0 ? (float)(bits * (1.0 + 1e-4)) : __builtin_inff();
}

// args.xr2 of an expanding operator: the largest squared target radius
// (scaled units) of a block that takes the expanded form; -1: never
__host__ __device__ inline float expand_r2(bool on) { return on ? 8.0f : -1.0f; }

template <class Op, class = void>
struct ExpandOf : std::false_type {
};
// operators whose coordinates are anchored and scaled offsets: every FP32
// operator, and FP64 operators that declare kAnchor
template <class Op, class = void>
struct AnchorOf : std::integral_constant<bool, (sizeof(typename Op::crd_t) < sizeof(double))> {
};
template <class Op>
struct AnchorOf<Op, std::void_t<decltype(Op::kAnchor)>> : std::integral_constant<bool, Op::kAnchor> {
};
template <class Op>
struct ExpandOf<Op, std::void_t<decltype(Op::kExpand)>> : std::integral_constant<bool, Op::kExpand> {
};

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ float wave_min(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ double wave_max_d(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Bounds of the aligned 64-point groups of an SoA point set (x[c*n + i]):
// out[g] = (centroid, radius) of points 64g .. 64g+63, in FP64 and unscaled
// coordinates.  One wave per group (g from the caller); every lane returns.
// The culling operators skip whole groups whose sphere is out of range of a
// block's target tile (pairs_block), so the four waves stay busy.
template <int D>
__device__ __forceinline__ void group_bounds_wave(const double *__restrict__ x, int n, int g, int lane,
                                                  double4 *__restrict__ out)
{
    const int i = g * 64 + lane;
    const bool ok = i < n;
    const double cnt = (double)min(64, n - g * 64);
    double c[3] = {0.0, 0.0, 0.0}, v[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < D; ++k) {
        v[k] = ok ? x[(size_t)k * n + i] : 0.0;
        c[k] = wave_sum(v[k]) / cnt;
    }
    double r2 = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) r2 += (v[k] - c[k]) * (v[k] - c[k]);
    const double r = sqrt(wave_max_d(ok ? r2 : 0.0));
    if (lane == 0) out[g] = make_double4(c[0], c[1], c[2], r);
}

// LDS bytes of one block: the staging tiles and the final wave reduction
// share one buffer.
template <class Op>
struct PairSmem {
    // record of one source: fields at 0, coordinates at `xoff`
    static constexpr size_t csize = sizeof(typename Op::crd_t);
    static constexpr size_t fbytes = sizeof(typename Op::lds_t) * (Op::F > 0 ? Op::F : 1);
    static constexpr size_t xoff = (fbytes + csize - 1) & ~(csize - 1);
    // coordinates, and |x_j|^2 for the expanded form
    static constexpr int DX = Op::D + (ExpandOf<Op>::value ? 1 : 0);
    static constexpr size_t rec = (xoff + csize * DX + 15) & ~size_t(15);
    static constexpr size_t stage = rec * kTileJ;
    static constexpr size_t cull = Op::kCull ? sizeof(float) * kTileJ : 0;   // source distances
    // surviving-group list of a window of kBlock groups and the per-wave counts
    static constexpr size_t grp = Op::kCull ? sizeof(int) * (kBlock + kWaves) : 0;
    static constexpr size_t red = sizeof(double) * kWaves * Op::OUT * 64 * Op::R;
    static constexpr size_t bytes = stage + cull + grp > red ? stage + cull + grp : red;
};

// One block's work: targets [tile*TI, tile*TI + TI) against sources
// [j0, j1); writes out[c*ni + i] for its targets.  Each staged tile of
// sources is split into four contiguous ranges, one per wave (with the
// spatial order a range is a compact patch, which the culling test uses).
template <class Op>
__device__ __forceinline__ void pairs_block(const typename Op::Args &args, int ni, int tile, int j0, int j1,
                                            double *__restrict__ out, char *smem)
{
    constexpr int D = Op::D, F = Op::F, OUT = Op::OUT, R = Op::R, TI = 64 * R;
    using lds_t = typename Op::lds_t;
    using crd_t = typename Op::crd_t;
    constexpr bool kAnchor = AnchorOf<Op>::value;
    using SM = PairSmem<Op>;
    auto red = reinterpret_cast<double (*)[OUT][TI]>(smem);
    // the wave index is uniform: as a scalar it keeps the source loop's
    // counter and bounds on the scalar unit
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

    typename Op::Tgt tg[R];
    double xt[R][D];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = tile * TI + r * 64 + lane;
        Op::load_tgt(args, i < ni ? i : -1, tg[r], xt[r]);   // padding lanes: coordinates 0
    }
    // anchor of the FP32 coordinates: the centroid of the tile's targets
    // (every wave holds all of them and forms the same sum)
    double anc[D];
    if constexpr (kAnchor) {
        const double inv = 1.0 / (double)min(TI, ni - tile * TI);
#pragma unroll
        for (int c = 0; c < D; ++c) {
            double v = 0.0;
#pragma unroll
            for (int r = 0; r < R; ++r) v += xt[r][c];
            anc[c] = wave_sum(v) * inv;
        }
    }
    double kap = 1.0;
    if constexpr (kAnchor) kap = Op::crd_scale(args);
    auto to_crd = [&](double v, int c) -> crd_t {
        if constexpr (kAnchor) return (crd_t)((v - anc[c]) * kap);
        else return v;
    };

    constexpr bool kExp = ExpandOf<Op>::value;
    crd_t xi[R][D];
    float xm[R][D], sc[R];   // expanded form: -2 x_i and exp2(-|x_i|^2)
    double tot[R][OUT];
    float rt2 = 0.0f;   // squared target tile radius about the anchor (culling)
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const bool real_lane = tile * TI + r * 64 + lane < ni;
        float d2 = 0.0f;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            xi[r][c] = to_crd(xt[r][c], c);
            d2 += (float)xi[r][c] * (float)xi[r][c];
        }
        if (real_lane) rt2 = fmaxf(rt2, d2);
#pragma unroll
        for (int c = 0; c < D; ++c) xm[r][c] = real_lane ? -2.0f * (float)xi[r][c] : 0.0f;
        sc[r] = real_lane ? __builtin_amdgcn_exp2f(-d2) : 1.0f;
#pragma unroll
        for (int c = 0; c < OUT; ++c) tot[r][c] = 0.0;
    }
    float rt = 0.0f, rtw2 = 0.0f;
    if constexpr (Op::kCull || kExp) {
        rtw2 = wave_max(rt2);
        rt = __builtin_sqrtf(rtw2);
    }
    float *s2buf = reinterpret_cast<float *>(smem + SM::stage);   // squared source distances to the anchor

    // staging of source j into record slot `slot` (and its squared distance
    // to the anchor for the range test)
    auto stage = [&](auto form, int j, int slot) {
        constexpr bool E = decltype(form)::value;
        double x[D];
        lds_t f[F > 0 ? F : 1];
        Op::load_src(args, j, x, f);
        char *rec = smem + SM::rec * slot;
#pragma unroll
        for (int c = 0; c < F; ++c) reinterpret_cast<lds_t *>(rec)[c] = f[c];
        float s2 = 0.0f;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            const crd_t v = to_crd(x[c], c);
            reinterpret_cast<crd_t *>(rec + SM::xoff)[c] = v;
            s2 += (float)v * (float)v;
        }
        if constexpr (E) reinterpret_cast<crd_t *>(rec + SM::xoff)[D] = s2;
        if constexpr (Op::kCull) s2buf[slot] = s2;
    };
    // the wave's pairs with the staged records [t0, t1): range test (culling
    // operators), the pair terms in the policy's type, the flush into FP64
    auto range = [&](auto form, int t0, int t1) {
        constexpr bool E = decltype(form)::value;
        using Ox = std::conditional_t<E, Op, Op>;   // dependent: pair_x/flush_x only when E
        if (t0 >= t1) return;
        if constexpr (Op::kCull) {
            const float ms2 = wave_min(t0 + lane < t1 ? s2buf[t0 + lane] : __builtin_inff());
            const float ms = __builtin_sqrtf(ms2);
            bool skip;
            if constexpr (E) {
                const float gap = ms - 2.0f * rt;
                skip = gap > 0.0f && ms * gap > (float)args.cut2;
            } else {
                const float gap = ms - rt;
                skip = gap > 0.0f && gap * gap > (float)args.cut2;
            }
            if (__builtin_amdgcn_readfirstlane((int)skip)) return;
            // diagnostic count of the pairs a culling operator computes
            if (args.exec != nullptr && lane == 0)
                atomicAdd(args.exec + Op::kCountSlot,
                          (unsigned long long)(t1 - t0) * (unsigned long long)min(TI, ni - tile * TI));
        }
        typename Op::Acc acc[R];
#pragma unroll
        for (int r = 0; r < R; ++r) Op::zero(acc[r]);
#pragma unroll 2
        for (int t = t0; t < t1; ++t) {
            crd_t xj[E ? D + 1 : D];
            lds_t fj[F > 0 ? F : 1];
            const char *rec = smem + SM::rec * t;
#pragma unroll
            for (int c = 0; c < F; ++c) fj[c] = reinterpret_cast<const lds_t *>(rec)[c];
#pragma unroll
            for (int c = 0; c < (E ? D + 1 : D); ++c) xj[c] = reinterpret_cast<const crd_t *>(rec + SM::xoff)[c];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if constexpr (E) Ox::pair_x(args, tg[r], xm[r], xj, fj, acc[r]);
                else Op::pair(args, tg[r], xi[r], xj, fj, acc[r]);
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if constexpr (E) Ox::flush_x(args, acc[r], tot[r], xi[r], sc[r]);
            else Op::flush(args, acc[r], tot[r]);
        }
    };
    // one sweep over the chunk's sources in the direct (E false) or the
    // expanded form: consecutive tiles of kTileJ sources, wave w taking the
    // w-th quarter of each.  The culling operators keep the quarters at 64
    // sources also in a partial last tile, so that every FP32 range sum is
    // over one aligned 64-source group, as in the grouped sweep below: with
    // and without culling the same groups give the same FP32 sums.
    auto sweep = [&](auto form) {
        for (int jt = j0; jt < j1; jt += kTileJ) {
            const int cnt = min(kTileJ, j1 - jt);
            __syncthreads();
            if ((int)threadIdx.x < cnt) stage(form, jt + threadIdx.x, threadIdx.x);
            __syncthreads();
            const int share = Op::kCull ? 64 : (cnt + kWaves - 1) / kWaves;   // <= 64
            const int t0 = wave * share;
            range(form, t0, min(t0 + share, cnt));
        }
    };
    // The grouped sweep of the culling operators, given the bounds of the
    // aligned 64-source groups (args.grp): a window of up to kBlock groups is
    // tested against the block's targets (one group per thread: the distance
    // from the anchor to the group's sphere, less the tile radius, against
    // the culling range), the surviving groups are listed in order, and they
    // are staged four at a time, one per wave, so every staged tile is full
    // of sources in range and the waves share the work evenly.  Inside a
    // group the per-wave range test still applies.
    auto sweep_groups = [&](auto form) {
      if constexpr (Op::kCull) {
        constexpr bool E = decltype(form)::value;
        int *glist = reinterpret_cast<int *>(smem + SM::stage + SM::cull);
        int *wcnt = glist + kBlock;
        const int gA = j0 >> 6, gB = (j1 + 63) >> 6;
        for (int w0 = gA; w0 < gB; w0 += kBlock) {
            const int wn = min(kBlock, gB - w0);
            __syncthreads();   // the previous window's list and tiles are consumed
            bool live = false;
            if ((int)threadIdx.x < wn) {
                const double4 b = args.grp[w0 + threadIdx.x];
                const double bc[3] = {b.x, b.y, b.z};
                double d2 = 0.0;
#pragma unroll
                for (int c = 0; c < D; ++c) d2 += (bc[c] - anc[c]) * (bc[c] - anc[c]);
                const float ms = (float)(fmax(0.0, sqrt(d2) - b.w) * kap);
                if constexpr (E) {
                    const float gap = ms - 2.0f * rt;
                    live = !(gap > 0.0f && ms * gap > (float)args.cut2);
                } else {
                    const float gap = ms - rt;
                    live = !(gap > 0.0f && gap * gap > (float)args.cut2);
                }
            }
            const unsigned long long bal = __ballot(live);
            if (lane == 0) wcnt[wave] = __popcll(bal);
            __syncthreads();
            int base = 0, total = 0;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                base += w < wave ? wcnt[w] : 0;
                total += wcnt[w];
            }
            if (live)
                glist[base + __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32),
                                                       __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u))] =
                    w0 + threadIdx.x;
            for (int k = 0; k < total; k += kWaves) {
                const int ng = min(kWaves, total - k);
                __syncthreads();   // the list is written; the previous tile is consumed
                if (wave < ng) {
                    const int j = glist[k + wave] * 64 + lane;
                    if (j < j1) stage(form, j, threadIdx.x);
                }
                __syncthreads();
                if (wave < ng) {
                    const int g0 = glist[k + wave] * 64;
                    range(form, 64 * wave, 64 * wave + min(64, j1 - g0));
                }
            }
        }
      }
    };
    auto run = [&](auto form) {
        if constexpr (Op::kCull) {
            if (args.grp != nullptr && (j0 & 63) == 0) {
                sweep_groups(form);
                return;
            }
        }
        sweep(form);
    };
    if constexpr (kExp) {
        // rtw2 is the same in every wave (all hold the block's targets)
        if (__builtin_amdgcn_readfirstlane((int)(rtw2 <= args.xr2))) run(std::true_type{});
        else run(std::false_type{});
    } else {
        run(std::false_type{});
    }

    // fixed-order combination of the four waves' totals
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int c = 0; c < OUT; ++c) red[wave][c][r * 64 + lane] = tot[r][c];
    __syncthreads();
    for (int idx = threadIdx.x; idx < OUT * TI; idx += kBlock) {
        const int c = idx / TI, il = idx - c * TI;
        double v = red[0][c][il];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) v += red[w][c][il];
        const int i = tile * TI + il;
        if (i < ni) out[(size_t)c * ni + i] = v;
    }
}

template <class Op>
__global__ __launch_bounds__(kBlock) void k_pairs(const typename Op::Args args, int ni, int nj, int chunk, int tiles,
                                                  int split_offset, double *__restrict__ part)
{
    __shared__ __align__(16) char smem[PairSmem<Op>::bytes];
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int tile = L % tiles, s = L / tiles;
    const int j0 = s * chunk, j1 = min(nj, j0 + chunk);
    pairs_block<Op>(args, ni, tile, j0, j1, part + (size_t)(s + split_offset) * Op::OUT * ni, smem);
}

// Split plan: target tiles of TI, source chunks (multiples of 64) so that
// the grid is about `target_blocks` workgroups, with at least one staged
// tile (256 sources) per chunk: below that the per-block staging, flush and
// split-reduction work outweighs the finer balance (small problems).
inline PairSplit make_split_ti(int ni, int nj, int ti, int target_blocks)
{
    PairSplit s;
    s.tiles = ceil_div(std::max(ni, 1), ti);
    int want = ceil_div(target_blocks, s.tiles);
    if (want < 1) want = 1;
    int chunk = ceil_div(std::max(nj, 1), want);
    chunk = ((chunk + 63) / 64) * 64;
    if (chunk < kTileJ) chunk = kTileJ;
    s.chunk = chunk;
    s.splits = nj > 0 ? ceil_div(nj, chunk) : 1;
    return s;
}

template <class Op>
void launch_pairs(const typename Op::Args &args, int ni, int nj, const PairSplit &sp, int split_offset, double *part,
                  hipStream_t st)
{
    if (ni <= 0) return;
    k_pairs<Op><<<sp.blocks(), kBlock, 0, st>>>(args, ni, nj, sp.chunk, sp.tiles, split_offset, part);
    LDDMM_LAUNCH_CHECK();
}

}  // namespace lddmm
