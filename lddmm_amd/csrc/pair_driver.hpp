// pair_driver.hpp -- the one tiled all-pairs loop every pair kernel uses.
//
// A block of 256 threads (4 wave64) owns TI = 64*R targets (lane l of
// every wave holds targets base + l + 64 r, r < R) and one chunk of the
// sources.  The chunk is staged through LDS 256 sources at a time, one
// 16-byte aligned record per source: the per-source fields in the policy's
// storage type, then the coordinates as FP64, so one base address and
// immediate offsets reach every field (wide ds_read_b128/b64 loads, no
// per-field address arithmetic in the pair loop).  Wave w consumes sources
// t = w, w+4, ... of each tile, so all 64 lanes of a wave read the same LDS
// words (broadcast, conflict-free) and every source is read once per wave
// for R targets.  Per tile, sums are
// carried in the policy's arithmetic type and flushed to FP64 totals
// (two-level accumulation: at most 64 pairs are summed in FP32 before the
// FP64 add); at the end the four waves' totals are combined in LDS in a
// fixed order and the block writes one partial per (split, component,
// target), which a following kernel reduces over the splits in a fixed
// order.  No atomics: results are bitwise repeatable.
//
// An operator Op supplies
//   D, F, OUT, R                       dimensions, LDS fields, outputs, targets/lane
//   Args                               kernel arguments
//   lds_t, Tgt, Acc                    LDS field type, per-target registers, per-tile sums
//   load_src(args, j, x[D], f[F])      source j's coordinates and fields
//   load_tgt(args, i, Tgt&)            target i (i < 0: padding lane)
//   pair(args, Tgt&, xj[D], fj[F], Acc&)    one pair term
//   zero(Acc&), flush(const Acc&, double tot[OUT])
#pragma once

#include "common.hpp"

namespace lddmm {

// Precision policies.  Coordinates are always FP64 in HBM and LDS; each
// coordinate difference is formed in FP64 and then rounded to the
// arithmetic type, so FP32 pair terms see |q_i - q_j| to FP32 relative
// accuracy instead of |q| to FP32 accuracy.
struct PrecF64 {
    using real = double;
    static constexpr bool fp32 = false;
};
struct PrecF32 {
    using real = float;
    static constexpr bool fp32 = true;
};

// Kernel family in the policy's arithmetic: value k and factor c of
// grad_{q_i} k = c (q_i - q_j), from the squared distance.  D selects the
// 2D Matern family (Bessel factors, evaluated in FP64 under both policies).
template <int ORDER, typename real, int D = 3>
struct KernelFnT;

template <int ORDER, int D>
struct KernelFnT<ORDER, double, D> {
    __device__ __forceinline__ static double value(double r2, double iw, double iw2)
    {
        return KernelFnD<D, ORDER>::value(r2, iw, iw2);
    }
    __device__ __forceinline__ static void eval(double r2, double iw, double iw2, double &k, double &c)
    {
        KernelFnD<D, ORDER>::eval(r2, iw, iw2, k, c);
    }
};

// FP32: exp via v_exp_f32 on a base-2 argument, sqrt via v_sqrt_f32 (both
// ~1 ulp).  Every width-dependent factor is written as a product of
// per-launch constants (-iw2 log2(e), -2 iw2, the Matern polynomial
// coefficients divided by their leading constant) so the compiler hoists
// them out of the pair loop: the Gaussian costs one multiply and one
// v_exp_f32 per pair for k, one more multiply for c.
// 2D Matern under the FP32 policy: the Bessel factors and the polynomial in
// FP64 from the FP32 squared distance, rounded to FP32.
template <int ORDER>
struct KernelFnT<ORDER, float, 2> {
    __device__ __forceinline__ static float value(float r2, float iw, float iw2)
    {
        return (float)KernelFnD<2, ORDER>::value((double)r2, (double)iw, (double)iw2);
    }
    __device__ __forceinline__ static void eval(float r2, float iw, float iw2, float &k, float &c)
    {
        double kd, cd;
        KernelFnD<2, ORDER>::eval((double)r2, (double)iw, (double)iw2, kd, cd);
        k = (float)kd;
        c = (float)cd;
    }
};
template <int ORDER>
struct KernelFnT<ORDER, float, 3> {
    static constexpr float kLog2e = 1.4426950408889634f;
    __device__ __forceinline__ static float value(float r2, float iw, float iw2)
    {
        if constexpr (ORDER == -1) {
            return __builtin_amdgcn_exp2f(r2 * (iw2 * -kLog2e));
        } else {
            const float s = __builtin_amdgcn_sqrtf(r2);
            const float d = s * iw;
            const float e = __builtin_amdgcn_exp2f(s * (iw * -kLog2e));
            if constexpr (ORDER == 0) return e;
            else if constexpr (ORDER == 1) return (1.0f + d) * e;
            else if constexpr (ORDER == 2) return (1.0f + d * (1.0f + d * (1.0f / 3.0f))) * e;
            else if constexpr (ORDER == 3)
                return (1.0f + d * (1.0f + d * (6.0f / 15.0f + d * (1.0f / 15.0f)))) * e;
            else
                return (1.0f + d * (1.0f + d * (45.0f / 105.0f + d * (10.0f / 105.0f + d * (1.0f / 105.0f))))) * e;
        }
    }
    __device__ __forceinline__ static void eval(float r2, float iw, float iw2, float &k, float &c)
    {
        if constexpr (ORDER == -1) {
            k = __builtin_amdgcn_exp2f(r2 * (iw2 * -kLog2e));
            c = k * (-2.0f * iw2);
        } else {
            const float s = __builtin_amdgcn_sqrtf(r2);
            const float d = s * iw;
            const float e = __builtin_amdgcn_exp2f(s * (iw * -kLog2e));
            if constexpr (ORDER == 0) {
                k = e;
                c = 0.0f;
            } else if constexpr (ORDER == 1) {
                k = (1.0f + d) * e;
                c = e * -iw2;
            } else if constexpr (ORDER == 2) {
                k = (1.0f + d * (1.0f + d * (1.0f / 3.0f))) * e;
                c = (1.0f + d) * e * (iw2 * (-1.0f / 3.0f));
            } else if constexpr (ORDER == 3) {
                k = (1.0f + d * (1.0f + d * (6.0f / 15.0f + d * (1.0f / 15.0f)))) * e;
                c = (1.0f + d * (1.0f + d * (1.0f / 3.0f))) * e * (iw2 * (-3.0f / 15.0f));
            } else {
                k = (1.0f + d * (1.0f + d * (45.0f / 105.0f + d * (10.0f / 105.0f + d * (1.0f / 105.0f))))) * e;
                c = (1.0f + d * (1.0f + d * (6.0f / 15.0f + d * (1.0f / 15.0f)))) * e * (iw2 * (-15.0f / 105.0f));
            }
        }
    }
};

// the Gaussian is the same function in 2D and 3D
template <>
struct KernelFnT<-1, float, 2> : KernelFnT<-1, float, 3> {
};

constexpr int kWaves = kBlock / 64;

// LDS bytes of one block: the staging tiles and the final wave reduction
// share one buffer.
template <class Op>
struct PairSmem {
    // record of one source: fields at 0, coordinates at `xoff`
    static constexpr size_t fbytes = sizeof(typename Op::lds_t) * (Op::F > 0 ? Op::F : 1);
    static constexpr size_t xoff = (fbytes + 7) & ~size_t(7);
    static constexpr size_t rec = (xoff + sizeof(double) * Op::D + 15) & ~size_t(15);
    static constexpr size_t stage = rec * kTileJ;
    static constexpr size_t red = sizeof(double) * kWaves * Op::OUT * 64 * Op::R;
    static constexpr size_t bytes = stage > red ? stage : red;
};

// One block's work: targets [tile*TI, tile*TI + TI) against sources
// [j0, j1); writes out[c*ni + i] for its targets.
template <class Op>
__device__ __forceinline__ void pairs_block(const typename Op::Args &args, int ni, int tile, int j0, int j1,
                                            double *__restrict__ out, char *smem)
{
    constexpr int D = Op::D, F = Op::F, OUT = Op::OUT, R = Op::R, TI = 64 * R;
    using lds_t = typename Op::lds_t;
    using SM = PairSmem<Op>;
    auto red = reinterpret_cast<double (*)[OUT][TI]>(smem);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;

    typename Op::Tgt tg[R];
    double tot[R][OUT];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = tile * TI + r * 64 + lane;
        Op::load_tgt(args, i < ni ? i : -1, tg[r]);
#pragma unroll
        for (int c = 0; c < OUT; ++c) tot[r][c] = 0.0;
    }

    for (int jt = j0; jt < j1; jt += kTileJ) {
        const int cnt = min(kTileJ, j1 - jt);
        __syncthreads();
        if ((int)threadIdx.x < cnt) {
            double x[D];
            lds_t f[F > 0 ? F : 1];
            Op::load_src(args, jt + threadIdx.x, x, f);
            char *rec = smem + SM::rec * threadIdx.x;
#pragma unroll
            for (int c = 0; c < F; ++c) reinterpret_cast<lds_t *>(rec)[c] = f[c];
#pragma unroll
            for (int c = 0; c < D; ++c) reinterpret_cast<double *>(rec + SM::xoff)[c] = x[c];
        }
        __syncthreads();
        typename Op::Acc acc[R];
#pragma unroll
        for (int r = 0; r < R; ++r) Op::zero(acc[r]);
#pragma unroll 2
        for (int t = wave; t < cnt; t += kWaves) {
            double xj[D];
            lds_t fj[F > 0 ? F : 1];
            const char *rec = smem + SM::rec * t;
#pragma unroll
            for (int c = 0; c < F; ++c) fj[c] = reinterpret_cast<const lds_t *>(rec)[c];
#pragma unroll
            for (int c = 0; c < D; ++c) xj[c] = reinterpret_cast<const double *>(rec + SM::xoff)[c];
#pragma unroll
            for (int r = 0; r < R; ++r) Op::pair(args, tg[r], xj, fj, acc[r]);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) Op::flush(acc[r], tot[r]);
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
// the grid is about `target_blocks` workgroups.
inline PairSplit make_split_ti(int ni, int nj, int ti, int target_blocks)
{
    PairSplit s;
    s.tiles = ceil_div(std::max(ni, 1), ti);
    int want = ceil_div(target_blocks, s.tiles);
    if (want < 1) want = 1;
    int chunk = ceil_div(std::max(nj, 1), want);
    chunk = ((chunk + 63) / 64) * 64;
    if (chunk < 64) chunk = 64;
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
