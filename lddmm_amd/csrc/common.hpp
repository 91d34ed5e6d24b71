// common.hpp -- shared definitions of the LDDMM HIP library (gfx950).
//
// Device layout (the reference's CUDA layout, CUDA/src/include/matvec.h:293-309):
// a point set is SoA, x[0..N) y[0..N) [z[0..N)], a time stack is
// (T-1) consecutive SoA slices, element tables are E x V column-major
// (vertex v of element e at e + v*E) and the adjacency table is
// N x rows column-major (entry k of vertex n at n + k*N).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <type_traits>
#include <stdexcept>
#include <string>

namespace lddmm {

// Status codes of the C-ABI (0 = ok, 1..3 = line-search codes).
enum : int {
    kOk = 0,
    kErrArg = -1,
    kErrAlloc = -2,
    kErrHip = -3,
    kErrUnsupported = -4,
};

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define LDDMM_HIP(call)                                                                          \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            throw ::lddmm::Error(::lddmm::kErrHip, std::string(#call) + ": " + hipGetErrorString(e_) + \
                                                       " (" __FILE__ ":" + std::to_string(__LINE__) + ")"); \
    } while (0)

#define LDDMM_LAUNCH_CHECK() LDDMM_HIP(hipGetLastError())

// Precision policies of the pair sums (the L-BFGS state, energies and all
// reductions across tiles are FP64 in both).
enum : int {
    kFP64 = 0,   // FP64 throughout: the reference's arithmetic
    kFP32 = 1,   // FP32 pair terms, FP64 coordinate differences and cross-tile sums
};

constexpr int kBlock = 256;       // threads per block of every pair kernel (4 wave64)
constexpr int kTileJ = 256;       // j points staged in LDS per tile
constexpr int kDotBlocks = 512;   // fixed grid of the deterministic dot reduction

// Kernel order: -1 Gaussian, 0..4 Matern (nu = order + 1/2 in 3D).
// All pair kernels depend on the squared distance r2 only.  For each
// order the device function returns k and the factor c of
// grad_{q_i} k = c (q_i - q_j) (Cpp/src/computeKernel.cpp:36-37,170-234;
// Cpp/src/dqKernel.cpp:57,215-291).  Order 0 has no gradient in the
// reference (Cpp/src/dqKernel.cpp:39, no case 0), so c = 0 there.
// sqrt(r2) for the Matern distance (r2 >= 0): Goldschmidt's square-root
// iteration (R. E. Goldschmidt, MIT thesis 1964; P. Markstein, "IA-64 and
// Elementary Functions", 2000, ch. 9) on the hardware reciprocal square
// root v_rsq_f64: g = a y, h = y/2 converge to sqrt(a) and 1/(2 sqrt(a)),
// and a residual correction d = a - g^2, g += d h gives the FP64 root to
// within an ulp (the library adds a second correction for correct rounding).
// The library sqrt adds the scaling of inputs below 2^-767 and the class
// tests of 0 and inf (9 more instructions per pair); here 1e-300 is added
// to r2 instead, so a point with itself (r2 = 0) gets a distance of
// 1e-150 / sigma, which rounds away in every Matern factor (1 + d, e^-d):
// the kernel values and gradient factors are those of d = 0.  From 2^-767
// up the addition leaves r2 unchanged and the sequence is the library's
// refinement, so the distance is the same to an ulp; a NaN r2 stays NaN.
__device__ __forceinline__ double sqrt_r2(double a)
{
    a += 1e-300;
    const double y = __builtin_amdgcn_rsq(a);
    double g = a * y, h = 0.5 * y;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    const double d = fma(-g, g, a);
    return fma(d, h, g);
}

// 2^-t for t >= 0 in FP64, the exponential of the pair kernels on scaled
// coordinates: n = rint(t) and f = n - t in [-1/2, 1/2] (exact), 2^f by its
// degree-10 minimax polynomial (Remez exchange on the relative error over
// [-1/2, 1/2]: 2.12e-16; 2.8e-16 with the coefficients rounded to FP64),
// then 2^-n by ldexp, which rounds to the subnormal range and to 0 beyond
// it.  n is clamped at 2000 before the conversion (every result is 0
// there).  Within 3 ulp of 2^-t over the normal range; a NaN t stays NaN,
// and so does t = +inf.  The library exp adds the argument
// reduction by ln 2 and the overflow / underflow selects: 15 instructions
// here against 21.
__device__ __forceinline__ double exp2_neg(double t)
{
    const double n = __builtin_rint(t);
    const double f = n - t;
    double p = 7.037279131748084279406e-9;
    p = fma(p, f, 1.020853790328922734484e-7);
    p = fma(p, f, 1.321566283403617560016e-6);
    p = fma(p, f, 1.525265811655012437259e-5);
    p = fma(p, f, 1.540352996113170006663e-4);
    p = fma(p, f, 1.333355822856152522105e-3);
    p = fma(p, f, 9.618129108034592386175e-3);
    p = fma(p, f, 5.550410866445883284577e-2);
    p = fma(p, f, 2.402265069590876778604e-1);
    p = fma(p, f, 6.931471805599497750728e-1);
    p = fma(p, f, 1.0);
    return __builtin_amdgcn_ldexp(p, (int)fmax(-n, -2000.0));
}

template <int ORDER>
struct KernelFn {
    // inv_w = 1/sigma, inv_w2 = 1/sigma^2
    __device__ __forceinline__ static void eval(double r2, double inv_w, double inv_w2, double &k, double &c)
    {
        if constexpr (ORDER == -1) {
            k = exp(-r2 * inv_w2);
            c = -2.0 * k * inv_w2;
        } else {
            double d = sqrt_r2(r2) * inv_w;
            double e = exp(-d);
            if constexpr (ORDER == 0) {
                k = e;
                c = 0.0;
            } else if constexpr (ORDER == 1) {
                k = (1.0 + d) * e;
                c = -e * inv_w2;
            } else if constexpr (ORDER == 2) {
                k = (3.0 + d * (3.0 + d)) * (1.0 / 3.0) * e;
                c = -(1.0 + d) * e * (inv_w2 * (1.0 / 3.0));
            } else if constexpr (ORDER == 3) {
                k = (15.0 + d * (15.0 + d * (6.0 + d))) * (1.0 / 15.0) * e;
                c = -(3.0 + d * (3.0 + d)) * e * (inv_w2 * (1.0 / 15.0));
            } else {
                k = (105.0 + d * (105.0 + d * (45.0 + d * (10.0 + d)))) * (1.0 / 105.0) * e;
                c = -(15.0 + d * (15.0 + d * (6.0 + d))) * e * (inv_w2 * (1.0 / 105.0));
            }
        }
    }
    __device__ __forceinline__ static double value(double r2, double inv_w, double inv_w2)
    {
        if constexpr (ORDER == -1) {
            return exp(-r2 * inv_w2);
        } else {
            double d = sqrt_r2(r2) * inv_w;
            double e = exp(-d);
            if constexpr (ORDER == 0) return e;
            else if constexpr (ORDER == 1) return (1.0 + d) * e;
            else if constexpr (ORDER == 2) return (3.0 + d * (3.0 + d)) * (1.0 / 3.0) * e;
            else if constexpr (ORDER == 3) return (15.0 + d * (15.0 + d * (6.0 + d))) * (1.0 / 15.0) * e;
            else return (105.0 + d * (105.0 + d * (45.0 + d * (10.0 + d)))) * (1.0 / 105.0) * e;
        }
    }
};

// ---------------------------------------------------------------------------
// 2D Matern kernels need p0 = x^2 K0(x) and p1 = x K1(x) (modified Bessel
// functions of the second kind).  FP64 on the device, the oracle's method:
// the ascending series (A&S 9.6.13, 9.6.11) for 0 < x <= 2,
//   K0 = -ln(x/2) I0 + sum_k psi(k+1) t^k/(k!)^2,  t = x^2/4,
//   K1 = 1/x + ln(x/2) I1 - (x/4) sum_k (psi(k+1) + psi(k+2)) t^k/(k!(k+1)!),
// and for x > 2 Steed's continued fraction for K1/K0 with the
// Thompson-Barnett normalising sum (Thompson & Barnett, J. Comput. Phys. 64
// (1986) 490; Temme 1975): K0 = sqrt(pi/(2x)) e^-x / S, K1 = K0 (x + 1/2 - h)/x.
// Limits p0(0) = 0, p1(0) = 1 and p0 = p1 = 0 beyond x = 700 as in the
// reference (Cpp/src/polybesselk.cpp:13-17, 54-58).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void bessel_p01(double x, double &p0, double &p1)
{
    if (x < 2.220446049250313e-16) {
        p0 = 0.0;
        p1 = 1.0;
        return;
    }
    if (x > 700.0) {
        p0 = 0.0;
        p1 = 0.0;
        return;
    }
    double k0, k1;
    if (x <= 2.0) {
        const double t = 0.25 * x * x, lg = log(0.5 * x);
        double i0 = 0.0, i1 = 0.0, s0 = 0.0, s1 = 0.0, term0 = 1.0, term1 = 1.0, hk = 0.0;
        for (int k = 0; k < 40; ++k) {
            const double psi1 = -0.57721566490153286061 + hk;
            const double psi2 = psi1 + 1.0 / (double)(k + 1);
            i0 += term0;
            i1 += term1;
            s0 += psi1 * term0;
            s1 += (psi1 + psi2) * term1;
            if (term0 < 1e-18 * i0 && k > 2) break;
            hk += 1.0 / (double)(k + 1);
            term0 *= t / ((double)(k + 1) * (double)(k + 1));
            term1 *= t / ((double)(k + 1) * (double)(k + 2));
        }
        k0 = -lg * i0 + s0;
        k1 = 1.0 / x + lg * (0.5 * x * i1) - 0.25 * x * s1;
    } else {
        const double a1 = 0.25;
        double b = 2.0 * (1.0 + x), d = 1.0 / b, h = d, dh = d;
        double qm1 = 0.0, qc = 1.0, a = -a1, c = a1, q = a1, s = 1.0 + q * dh;
        for (int i = 2; i < 400; ++i) {
            a -= 2.0 * (double)(i - 1);
            c = -a * c / (double)i;
            const double qn = (qm1 - b * qc) / a;
            qm1 = qc;
            qc = qn;
            q += c * qn;
            b += 2.0;
            d = 1.0 / (b + a * d);
            dh = (b * d - 1.0) * dh;
            h += dh;
            const double ds = q * dh;
            s += ds;
            if (fabs(ds) < 1e-17 * fabs(s)) break;
        }
        h *= a1;
        k0 = sqrt(1.5707963267948966192 / x) * exp(-x) / s;
        k1 = k0 * (x + 0.5 - h) / x;
    }
    p0 = x * x * k0;
    p1 = x * k1;
}

// Kernel value k and gradient factor c (grad_{q_i} k = c (q_i - q_j)) of
// the 2D Matern family, order p (nu = p + 1), d = |q_i - q_j|/sigma
// (Cpp/src/computeKernel.cpp:60,83,105,127,149; Cpp/src/dqKernel.cpp:91,
// 122,153,184; order 0 has no derivative in the reference: c = 0).
template <int ORDER>
struct KernelFn2D {
    __device__ __forceinline__ static void eval(double r2, double inv_w, double inv_w2, double &k, double &c)
    {
        const double d = sqrt_r2(r2) * inv_w, d2 = r2 * inv_w2;
        double p0, p1;
        bessel_p01(d, p0, p1);
        if constexpr (ORDER == 0) {
            k = p1;
            c = 0.0;
        } else if constexpr (ORDER == 1) {
            k = 0.5 * (p0 + 2.0 * p1);
            c = -p1 * (0.5 * inv_w2);
        } else if constexpr (ORDER == 2) {
            k = (4.0 * p0 + (8.0 + d2) * p1) * 0.125;
            c = -(p0 + 2.0 * p1) * (0.125 * inv_w2);
        } else if constexpr (ORDER == 3) {
            k = ((24.0 + d2) * p0 + 8.0 * (6.0 + d2) * p1) * (1.0 / 48.0);
            c = -(4.0 * p0 + (8.0 + d2) * p1) * (inv_w2 / 48.0);
        } else {
            k = (12.0 * (16.0 + d2) * p0 + (384.0 + d2 * (72.0 + d2)) * p1) * (1.0 / 384.0);
            c = -((24.0 + d2) * p0 + 8.0 * (6.0 + d2) * p1) * (inv_w2 / 384.0);
        }
    }
    __device__ __forceinline__ static double value(double r2, double inv_w, double inv_w2)
    {
        double k, c;
        eval(r2, inv_w, inv_w2, k, c);
        return k;
    }
};

// The kernel family of a dimension: 2D Matern orders through the Bessel
// factors, Gaussian (both dimensions) and 3D Matern through KernelFn.
template <int D, int ORDER>
using KernelFnD = std::conditional_t<(D == 2 && ORDER >= 0), KernelFn2D<ORDER>, KernelFn<ORDER>>;

// Dispatch a runtime order onto a template lambda f(std::integral_constant<int, O>).
template <typename F>
inline void dispatch_order(int order, F &&f)
{
    switch (order) {
    case -1: f(std::integral_constant<int, -1>{}); break;
    case 0: f(std::integral_constant<int, 0>{}); break;
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 3: f(std::integral_constant<int, 3>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    default: throw Error(kErrArg, "knlOrder must be -1 (Gaussian) or 0..4 (Matern)");
    }
}

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

// Sum over the 64 lanes of a wave by an xor butterfly: every lane ends
// with the same, order-fixed result.
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// j-split of an all-pairs sum over ni targets and nj sources: every block
// owns kBlock targets and a chunk of sources; chunks are multiples of 64 so
// the chip gets at least ~target_blocks workgroups.
struct PairSplit {
    int tiles;   // i tiles
    int chunk;   // sources per split
    int splits;  // number of j chunks (S)
    int blocks() const { return tiles * splits; }
};

inline PairSplit make_split(int ni, int nj, int target_blocks)
{
    PairSplit s;
    s.tiles = ceil_div(ni, kBlock);
    int want = ceil_div(target_blocks, s.tiles);
    if (want < 1) want = 1;
    int chunk = ceil_div(nj, want);
    chunk = ((chunk + 63) / 64) * 64;
    if (chunk < 64) chunk = 64;
    s.chunk = chunk;
    s.splits = nj > 0 ? ceil_div(nj, chunk) : 1;
    return s;
}

// Blocks are dispatched round-robin over the 8 XCDs of the chip (block b on
// XCD b % 8).  This maps the hardware block id to a logical id such that
// consecutive logical ids -- which share a source chunk -- run on one XCD
// and hit the same L2.
__device__ __forceinline__ int xcd_remap(int b, int nb)
{
    constexpr int kXcd = 8;
    int per = nb / kXcd;           // full rounds
    int rem = nb - per * kXcd;     // blocks in the last partial round
    int x = b % kXcd, k = b / kXcd;
    (void)rem;
    if (k >= per) return b;        // the tail keeps its id
    return x * per + k;
}

}  // namespace lddmm
