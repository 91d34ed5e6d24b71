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
template <int ORDER>
struct KernelFn {
    // inv_w = 1/sigma, inv_w2 = 1/sigma^2
    __device__ __forceinline__ static void eval(double r2, double inv_w, double inv_w2, double &k, double &c)
    {
        if constexpr (ORDER == -1) {
            k = exp(-r2 * inv_w2);
            c = -2.0 * k * inv_w2;
        } else {
            double d = sqrt(r2) * inv_w;
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
            double d = sqrt(r2) * inv_w;
            double e = exp(-d);
            if constexpr (ORDER == 0) return e;
            else if constexpr (ORDER == 1) return (1.0 + d) * e;
            else if constexpr (ORDER == 2) return (3.0 + d * (3.0 + d)) * (1.0 / 3.0) * e;
            else if constexpr (ORDER == 3) return (15.0 + d * (15.0 + d * (6.0 + d))) * (1.0 / 15.0) * e;
            else return (105.0 + d * (105.0 + d * (45.0 + d * (10.0 + d)))) * (1.0 / 105.0) * e;
        }
    }
};

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
