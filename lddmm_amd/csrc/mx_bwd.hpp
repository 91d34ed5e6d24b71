// mx_bwd.hpp -- the batched FP32-policy adjoint of the Gaussian kernel in 3D
// on the matrix cores (included by batch.hip after kb_bwd).
//
// The fused adjoint (OpBwd in pair_ops.hpp, after Cpp/src/objgrd.cpp:63-85)
// sums, for every target i,
//     g_i  = sum_j k_ij w_j,                        w_j = a_j - p_j
//     dq_i = sum_j k_ij (u_i . v_j) (x_i - x_j),    u_i = (-p_i, a_i), v_j = (a_j, w_j)
// with k_ij = exp2(-|x_i - x_j|^2) in scaled coordinates.  Written as
// products of the kernel tile K = (k_ij) with 24 columns per source,
//     w (3), a (3), w (x) x (9), a (x) x (9),
// every sum is K times a column block: g_i = (K w)_i, and with
// P = K (a, w) and Q = K (a (x) x, w (x) x),
//     sf_i = -p_i . P_a + a_i . P_w,   dq_i = x_i sf_i - (-p_i . Q_a + a_i . Q_w).
//
// Per block: 256 targets, one wave per 64 (four row tiles of 16).  Per 32
// sources and row tile, the exponents t_ij = |x_j'|^2 - 2 y_i . x_j' - S come
// from two v_mfma_f32_16x16x4f32 (sources x (x', |x'|^2 - S) against targets
// x (-2 y, 1); x' and y are offsets from the wave's centroid of its targets,
// so the expanded form keeps its error bound of |y| <= rt), the kernel values
// k' = exp2(-t) = 2^S exp2(-|x_j - x_i|^2 + |y_i|^2) from v_exp_f32, and the
// products with the columns from v_mfma_f32_16x16x32_f16 on a split of both
// factors into FP16 high and low halves (hi.hi + hi.lo + lo.hi: 22 bits of
// each product, FP32 accumulation; columns 0-15 take three MFMAs per row
// tile, and columns 16-23, whose hi and lo halves share the second tile of
// 16, two: (hi + lo) of k' against both halves); the low half of k' is formed by one
// mixed-precision FMA per value (k' minus its FP16 high half, exact).  The
// wave's scale 2^S keeps k' below 2^15, and power-of-two column scales from
// each staged tile's maxima keep the columns below 2^14, so both halves stay
// normal down to 2^-29 of the largest value.  A tile's column scales come
// from its four groups' data, without a pass over the sources: the group
// bounds of q_t bound the offsets, and the group maxima of |alpha_t| and
// |alpha_t - p| (kb_group_maxima, launched before each step) bound the
// vectors.  The FP32 sums run over the whole source chunk: when a staged
// tile's column scales differ from the previous tile's, the sums are
// rescaled by the (power-of-two, exact) ratio.
// exp2(-|y_i|^2), the scales and the combination with the target's own
// vectors are applied once per block in FP64, one thread per target.
//
// The wave's exponent operand (x', |x'|^2 - S) of every staged source is
// formed once per staged tile, by the staging thread for all four waves,
// into per-wave rows of LDS (component-major, padded so that the 64 lanes'
// reads hit distinct banks).
//
// The chunk's aligned 64-source groups are tested against the block's
// targets (group bounds of the culling, BatchView grpv) before staging, the
// surviving groups are staged four to a tile, and each wave skips the staged
// groups whose kernel values against all its targets are below the culling
// level (the exact bound on r^2, as the direct form).
//
// A block whose wave tiles are wider than the expanded form's bound (small
// problems, narrow kernels), a run with the expanded form or the source
// groups switched off, takes the vector path instead (pairs_block, one
// target per lane, four tiles of 64 targets).
#pragma once

typedef float mx_v4f __attribute__((ext_vector_type(4)));
typedef _Float16 mx_v8h __attribute__((ext_vector_type(8)));
typedef __fp16 mx_h2 __attribute__((ext_vector_type(2)));

constexpr int kMxTI = 256;          // targets per block
constexpr int kMxTJ = 256;          // sources per staged tile
constexpr int kMxRow = kMxTJ + 8;   // halves per column row (16-byte pad: the rows of one read hit distinct banks)
constexpr int kMxCols = 24;         // columns per source (two tiles of 16; the second holds 16-23 as hi and lo halves)
constexpr int kMxAP = kMxTJ + 4;    // floats per component row of a wave's exponent operand

struct MxSmem {
    _Float16 colh[kMxCols][kMxRow];   // columns, high halves: [column][source]
    _Float16 coll[kMxCols][kMxRow];   // low halves
    float aop[kWaves][4][kMxAP];    // per wave: x' (3) and |x'|^2 - S of the staged sources
    float wfr[kWaves][4];           // per wave: its centroid (block frame, scaled) and S
    double dred[kWaves][4];         // block reductions
    int glist[kMxTJ];               // surviving groups of a window
    int wcnt[kWaves];
    int flag[kWaves];
    float wr2[kWaves];
};

using OpBwdMxFallback = OpBwd<3, -1, PrecF32, 1>;
constexpr size_t kMxEpilogueBytes = sizeof(double) * kWaves * 64 * kMxCols;   // FP64 column sums per target
constexpr size_t mx_max(size_t x, size_t y) { return x > y ? x : y; }
constexpr size_t kMxSmemBytes = mx_max(mx_max(sizeof(MxSmem), PairSmem<OpBwdMxFallback>::bytes), kMxEpilogueBytes);

// 2^(top - e) for m = f 2^e, f in [1/2, 1): m times it lies below 2^top;
// `keep` when m is zero (no constraint) or not finite
__device__ __forceinline__ float mx_pow2_below(float m, int top, float keep)
{
    if (!(m > 0.0f) || !(m < 3.0e38f)) return keep;
    int e;
    (void)frexpf(m, &e);
    return ldexpf(1.0f, top - e);
}

// e minus the low or the high FP16 half of h, in FP32 (v_fma_mix_f32: the
// FP16 operand converted on the fly); exact when that half is e truncated
// to FP16
__device__ __forceinline__ float mx_sub_h16_lo(float e, mx_h2 h)
{
    float r;
    asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h), "v"(e));
    return r;
}
__device__ __forceinline__ float mx_sub_h16_hi(float e, mx_h2 h)
{
    float r;
    asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h), "v"(e));
    return r;
}

// totals of three values over the block: one barrier; S.dred is free again
// after the caller's next barrier
__device__ __forceinline__ void mx_block_sum3(double (&v)[3], MxSmem &S, int lane, int wave, bool max_op)
{
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] = max_op ? wave_max_d(v[c]) : wave_sum(v[c]);
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < 3; ++c) S.dred[wave][c] = v[c];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double t = S.dred[0][c];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) t = max_op ? fmax(t, S.dred[w][c]) : t + S.dred[w][c];
        v[c] = t;
    }
}

template <int D>
__global__ __launch_bounds__(kBlock, 3) void kb_bwd_mx(BatchView v, int t, const int4 *__restrict__ items)
{
    static_assert(D == 3, "matrix-core adjoint: 3D only");
    using KF = KernelFnT<-1, float, 3>;
    __shared__ __align__(16) char smem[kMxSmemBytes];
    MxSmem &S = *reinterpret_cast<MxSmem *>(smem);
    const int4 it = items[xcd_remap(blockIdx.x, gridDim.x)];
    const BatchDesc d = v.desc[it.x];
    const int n = d.n, tile = it.y;
    const double *q = v.nde + (size_t)v.T * 3 * d.vo + (size_t)t * 3 * n;
    const double *a = v.alp + (size_t)(v.T - 1) * 3 * d.vo + (size_t)t * 3 * n;
    const double *p = v.padj + (size_t)3 * d.vo;
    const double4 *grp = v.gpts ? v.grpv + d.gvo + (size_t)t * ((n + 63) / 64) : nullptr;
    const double kap = KF::kappa(d.iw, d.iw2);
    const double pds = KF::cmul(d.iw, d.iw2) / kap;
    const int j0 = it.z * d.chunk, j1 = min(n, j0 + d.chunk);
    double *out = v.part + d.po + (size_t)it.z * 6 * n;
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tcount = min(kMxTI, n - tile * kMxTI);
    const int wcount = max(0, min(64, tcount - 64 * wave));

    // block anchor: centroid of the block's targets (FP64)
    const int ti = tile * kMxTI + threadIdx.x;
    const bool real = ti < n;
    double qi[3], anc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) anc[c] = qi[c] = real ? q[(size_t)c * n + ti] : 0.0;
    mx_block_sum3(anc, S, lane, wave, false);
#pragma unroll
    for (int c = 0; c < 3; ++c) anc[c] /= (double)tcount;
    float xh[3], dw[3], y[3], y2 = 0.0f, xh2 = 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        xh[c] = real ? (float)((qi[c] - anc[c]) * kap) : 0.0f;
        xh2 += xh[c] * xh[c];
    }
    // the wave's frame: offsets from the centroid of its targets
#pragma unroll
    for (int c = 0; c < 3; ++c) dw[c] = (float)(wave_sum(real ? (double)xh[c] : 0.0) / (double)max(wcount, 1));
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        y[c] = real ? xh[c] - dw[c] : 0.0f;
        y2 += y[c] * y[c];
    }
    const float rtw2 = wave_max(real ? y2 : 0.0f);
    const bool ok = wcount == 0 || rtw2 <= expand_r2(v.expand != 0);
    // the wave's kernel scale: k' = 2^S exp2(-|x_j - x_i|^2 + |y_i|^2) <= 2^(S + rt^2) <= 2^15
    const float Sk = floorf(15.0f - rtw2);
    const float wr2 = wave_max(real ? xh2 : 0.0f);   // squared wave radius about the block anchor
    __syncthreads();   // dred reads done
    if (lane == 0) {
        S.flag[wave] = ok ? 1 : 0;
        S.wr2[wave] = wr2;
#pragma unroll
        for (int c = 0; c < 3; ++c) S.wfr[wave][c] = dw[c];
        S.wfr[wave][3] = Sk;
    }
    __syncthreads();
    const bool all_ok = S.flag[0] && S.flag[1] && S.flag[2] && S.flag[3] && grp != nullptr && v.grpm != nullptr &&
                        (j0 & 63) == 0;
    if (!all_ok) {
        __syncthreads();   // flags read before the vector path reuses the memory
        using Op = OpBwdMxFallback;
        typename Op::Args A{q, a, p, n, (float)d.iw, (float)d.iw2, cull_r2(v.cull), expand_r2(v.expand != 0), kap,
                            pds, v.exec, grp};
        for (int s = 0; s < kMxTI / 64 && (4 * tile + s) * 64 < n; ++s) {
            pairs_block<Op>(A, n, 4 * tile + s, j0, j1, out, smem);
            __syncthreads();
        }
        return;
    }
    const float rtb = __builtin_sqrtf(fmaxf(fmaxf(S.wr2[0], S.wr2[1]), fmaxf(S.wr2[2], S.wr2[3])));

    const double k2 = ldexp(1.0, -(int)Sk);
    // culling: the wave's centroid (unscaled) and radius (scaled)
    double aw[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) aw[c] = anc[c] + (double)dw[c] / kap;
    const float rtw = __builtin_sqrtf(rtw2), cut2 = cull_r2(v.cull), kapf = (float)kap;
    const int nn = lane & 15, g = lane >> 4;
    // target operand of the exponent products, per row tile: (-2 y, 1)
    float bs[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        const int src = 16 * mt + nn;
        const float y0 = __shfl(y[0], src, 64), y1 = __shfl(y[1], src, 64), yz = __shfl(y[2], src, 64);
        bs[mt] = g == 0 ? -2.0f * y0 : g == 1 ? -2.0f * y1 : g == 2 ? -2.0f * yz : 1.0f;
    }
    // lane (m, k) of the source operand holds component k of source
    // 8(m/4) + m%4 (+4 for the second product) of the K-block
    const int js0 = 8 * (nn >> 2) + (nn & 3), co0 = 8 * g;

    mx_v4f C[4][2];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) C[mt][0] = C[mt][1] = mx_v4f{0.0f, 0.0f, 0.0f, 0.0f};
    const mx_v4f zero4 = {0.0f, 0.0f, 0.0f, 0.0f};
    // column scales of the sums in C: offsets, a, a - p (powers of two)
    float sx = 1.0f, sa = 1.0f, sw = 1.0f;
    // factor of column `col`'s sum for the scales (x, a, w)
    auto colf = [](int col, float fx, float fa, float fw) -> float {
        return col < 3 ? fw : col < 6 ? fa : col < 15 ? fw * fx : col < 24 ? fa * fx : 0.0f;
    };
    unsigned long long executed = 0;

    // per group of q_t (this problem, this step): the largest |alpha_t| and |alpha_t - p|
    const float2 *gmax = v.grpm + d.gvo / v.T;

    // One staged tile: slot s holds source 64 glist[k + s/64] + s%64 of the
    // surviving-group list (ng groups from position k).
    auto tile_run = [&](int k, int ng) {
        __syncthreads();   // the previous tile is consumed; the group list is written
        // lane qq < ng reads quarter qq's group once: its valid sources, the
        // bound of its offsets from the block anchor, its maxima of |a| and
        // |a - p|, and the wave's culling test against it; the wave combines
        // the four lanes' values (uniform)
        int vq = 0;
        float xq = 0.0f, aq = 0.0f, wq = 0.0f;
        bool keep = false;
        if (lane < ng) {
            const int gi = S.glist[k + lane];
            vq = min(64, j1 - 64 * gi);
            const double4 b = grp[gi];
            const double dx = b.x - anc[0], dy = b.y - anc[1], dz = b.z - anc[2];
            xq = (__builtin_sqrtf((float)(dx * dx + dy * dy + dz * dz)) + (float)b.w) * kapf;
            const float2 m = gmax[gi];
            aq = m.x;
            wq = m.y;
            const double ex = b.x - aw[0], ey = b.y - aw[1], ez = b.z - aw[2];
            const float ms = fmaxf(0.0f, __builtin_sqrtf((float)(ex * ex + ey * ey + ez * ez)) - (float)b.w) * kapf;
            const float gap = ms - rtw;
            keep = !(gap > 0.0f && gap * gap > cut2);
        }
        unsigned vpack = 0u;
        float xb = 0.0f, ab = 0.0f, wb = 0.0f;
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) {
            vpack |= (unsigned)__builtin_amdgcn_readlane(vq, qq) << (8 * qq);
            xb = fmaxf(xb, __int_as_float(__builtin_amdgcn_readlane(__float_as_int(xq), qq)));
            ab = fmaxf(ab, __int_as_float(__builtin_amdgcn_readlane(__float_as_int(aq), qq)));
            wb = fmaxf(wb, __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wq), qq)));
        }
        const unsigned live = wcount > 0 ? (unsigned)__ballot(keep) & 0xFu : 0u;   // bit qq: quarter qq
        // column scales of this tile: kept while the tile fits them with at
        // most 8 bits of headroom unused, else the tile's own; the sums so
        // far follow a change exactly (powers of two)
        auto keep_or = [](float cur, float want) { return want < cur || want > 256.0f * cur ? want : cur; };
        const float nx = keep_or(sx, mx_pow2_below(xb, 0, sx)), na = keep_or(sa, mx_pow2_below(ab, 14, sa)),
                    nw = keep_or(sw, mx_pow2_below(wb, 14, sw));
        if (nx != sx || na != sa || nw != sw) {
            const float fx = nx / sx, fa = na / sa, fw = nw / sw;
            const float f0 = colf(nn, fx, fa, fw), f1 = colf(16 + (nn & 7), fx, fa, fw);
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                C[mt][0] *= f0;
                C[mt][1] *= f1;
            }
            sx = nx;
            sa = na;
            sw = nw;
        }
        const int tj = threadIdx.x;
        const int jsrc =
            lane < (int)((vpack >> (8 * wave)) & 0xFFu) ? 64 * __builtin_amdgcn_readfirstlane(S.glist[k + wave]) + lane : -1;
        // staging: the source's offsets and vectors (zero for an empty slot)
        float xs[3] = {0.0f, 0.0f, 0.0f}, av[3] = {0.0f, 0.0f, 0.0f}, wv[3] = {0.0f, 0.0f, 0.0f};
        if (jsrc >= 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double ac = a[(size_t)c * n + jsrc], pc = p[(size_t)c * n + jsrc];
                xs[c] = (float)((q[(size_t)c * n + jsrc] - anc[c]) * kap);
                av[c] = (float)ac;
                wv[c] = (float)(ac - pc);
            }
        }
        // the wave exponent operands of this source, for every wave
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const float x0 = xs[0] - S.wfr[w][0], x1 = xs[1] - S.wfr[w][1], x2 = xs[2] - S.wfr[w][2];
            S.aop[w][0][tj] = x0;
            S.aop[w][1][tj] = x1;
            S.aop[w][2][tj] = x2;
            S.aop[w][3][tj] = __builtin_fmaf(x0, x0, __builtin_fmaf(x1, x1, __builtin_fmaf(x2, x2, -S.wfr[w][3])));
        }
        // the 24 scaled columns as FP16 halves, both rounded toward zero: hi
        // so that v - hi is exact in FP32, lo within 2^-22 of v
        {
            float xsc[3], wsc[3], asc[3], col[kMxCols];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                xsc[c] = xs[c] * sx;
                wsc[c] = wv[c] * sw;
                asc[c] = av[c] * sa;
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                col[c] = wsc[c];
                col[3 + c] = asc[c];
#pragma unroll
                for (int e = 0; e < 3; ++e) {
                    col[6 + 3 * c + e] = wsc[c] * xsc[e];
                    col[15 + 3 * c + e] = asc[c] * xsc[e];
                }
            }
#pragma unroll
            for (int kk = 0; kk < kMxCols; kk += 2) {
                const mx_h2 h = __builtin_amdgcn_cvt_pkrtz(col[kk], col[kk + 1]);
                const mx_h2 l = __builtin_amdgcn_cvt_pkrtz(mx_sub_h16_lo(col[kk], h), mx_sub_h16_hi(col[kk + 1], h));
                S.colh[kk][tj] = h[0];
                S.colh[kk + 1][tj] = h[1];
                S.coll[kk][tj] = l[0];
                S.coll[kk + 1][tj] = l[1];
            }
        }
        __syncthreads();   // columns and exponent operands are staged
        if (live == 0u) return;
        // the live K-blocks of the tile as a bit mask (uniform)
        unsigned kmask = 0u;
#pragma unroll
        for (int kb = 0; kb < 8; ++kb)
            if (((live >> (kb >> 1)) & 1u) && (int)((vpack >> (8 * (kb >> 1))) & 0xFFu) > 32 * (kb & 1))
                kmask |= 1u << kb;
        // second column tile: lanes 0-7 take the hi halves of columns 16-23 and
        // lanes 8-15 their lo halves, so that (ah + al) times it gives every
        // product of the split in two MFMAs; the two lanes of a column are added
        // at the end
        const _Float16(*col1)[kMxRow] = nn < 8 ? S.colh : S.coll;
        const int row1 = 16 + (nn & 7);
        while (kmask != 0u) {
            const int kb = __builtin_ctz(kmask);
            kmask &= kmask - 1u;
            const int vk = min(32, (int)((vpack >> (8 * (kb >> 1))) & 0xFFu) - 32 * (kb & 1));
            executed += (unsigned long long)vk * (unsigned long long)wcount;
            const float A0 = S.aop[wave][g][kb * 32 + js0], A1 = S.aop[wave][g][kb * 32 + js0 + 4];
            const mx_v8h h0c = *reinterpret_cast<const mx_v8h *>(&S.colh[nn][kb * 32 + co0]);
            const mx_v8h l0c = *reinterpret_cast<const mx_v8h *>(&S.coll[nn][kb * 32 + co0]);
            const mx_v8h b1c = *reinterpret_cast<const mx_v8h *>(&col1[row1][kb * 32 + co0]);
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                const mx_v4f s0 = __builtin_amdgcn_mfma_f32_16x16x4f32(A0, bs[mt], zero4, 0, 0, 0);
                const mx_v4f s1 = __builtin_amdgcn_mfma_f32_16x16x4f32(A1, bs[mt], zero4, 0, 0, 0);
                // exponents: lane (n, g) has target n against sources 8g .. 8g+7;
                // k = hi + lo: hi truncated to FP16 (so k - hi is exact in
                // FP32), lo rounded to nearest
                mx_v8h ah, al;
#pragma unroll
                for (int r = 0; r < 4; r += 2) {
                    const float e0 = __builtin_amdgcn_exp2f(-s0[r]), e1 = __builtin_amdgcn_exp2f(-s0[r + 1]);
                    const float f0 = __builtin_amdgcn_exp2f(-s1[r]), f1 = __builtin_amdgcn_exp2f(-s1[r + 1]);
                    const mx_h2 he = __builtin_amdgcn_cvt_pkrtz(e0, e1), hf = __builtin_amdgcn_cvt_pkrtz(f0, f1);
                    ah[r] = he[0];
                    ah[r + 1] = he[1];
                    ah[4 + r] = hf[0];
                    ah[5 + r] = hf[1];
                    al[r] = (_Float16)mx_sub_h16_lo(e0, he);
                    al[r + 1] = (_Float16)mx_sub_h16_hi(e1, he);
                    al[4 + r] = (_Float16)mx_sub_h16_lo(f0, hf);
                    al[5 + r] = (_Float16)mx_sub_h16_hi(f1, hf);
                }
                C[mt][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, h0c, C[mt][0], 0, 0, 0);
                C[mt][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, b1c, C[mt][1], 0, 0, 0);
                C[mt][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, l0c, C[mt][0], 0, 0, 0);
                C[mt][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, b1c, C[mt][1], 0, 0, 0);
                C[mt][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, h0c, C[mt][0], 0, 0, 0);
            }
        }
    };

    {
        // the chunk's groups against the block (anchor, radius rtb), a window
        // of up to kBlock groups at a time, survivors listed in order
        const int gA = j0 >> 6, gB = (j1 + 63) >> 6;
        for (int w0 = gA; w0 < gB; w0 += kBlock) {
            const int wn = min(kBlock, gB - w0);
            __syncthreads();   // the previous window's list and tiles are consumed
            bool glive = false;
            if ((int)threadIdx.x < wn) {
                const double4 b = grp[w0 + threadIdx.x];
                const double dx = b.x - anc[0], dy = b.y - anc[1], dz = b.z - anc[2];
                const float ms = fmaxf(0.0f, __builtin_sqrtf((float)(dx * dx + dy * dy + dz * dz)) - (float)b.w) * kapf;
                const float gap = ms - rtb;
                glive = !(gap > 0.0f && gap * gap > cut2);
            }
            const unsigned long long bal = __ballot(glive);
            if (lane == 0) S.wcnt[wave] = __popcll(bal);
            __syncthreads();
            int base = 0, total = 0;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                base += w < wave ? S.wcnt[w] : 0;
                total += S.wcnt[w];
            }
            if (glive)
                S.glist[base + __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u))] =
                    w0 + threadIdx.x;
            for (int k = 0; k < total; k += 4) tile_run(k, min(4, total - k));
        }
    }
    if (v.exec != nullptr && lane == 0 && executed != 0) atomicAdd(v.exec + 1, executed);

    // FP64 factors of the lane's two columns: the inverse column scales and 2^-S
    auto cinv = [&](int col) -> double {
        if (col < 3) return k2 / sw;
        if (col < 6) return k2 / sa;
        if (col < 15) return k2 / ((double)sw * sx);
        if (col < 24) return k2 / ((double)sa * sx);
        return 0.0;
    };
    const double cs0 = cinv(nn), cs1 = cinv(16 + (nn & 7));
    // combination with the targets' own vectors, one thread per target:
    // lane (n, g) holds column n of targets 16 mt + 4 g + r, and the hi (n < 8)
    // or lo (n >= 8) part of column 16 + n % 8
    static_assert(kMxEpilogueBytes <= kMxSmemBytes, "epilogue buffer");
    double *E = reinterpret_cast<double *>(smem);   // [wave][target][column], 48 KiB
    __syncthreads();   // the last tile is consumed
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int tl = 16 * mt + 4 * g + r;
            E[((size_t)wave * 64 + tl) * kMxCols + nn] = (double)C[mt][0][r] * cs0;
            const double c1 = (double)C[mt][1][r] * cs1;
            const double c1t = c1 + __shfl_xor(c1, 8, 64);   // the column's hi and lo parts
            if (nn < 8) E[((size_t)wave * 64 + tl) * kMxCols + 16 + nn] = c1t;
        }
    __syncthreads();
    if (!real) return;
    const double *V = E + (size_t)threadIdx.x * kMxCols;
    double ai[3], pi[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        ai[c] = a[(size_t)c * n + ti];
        pi[c] = p[(size_t)c * n + ti];
    }
    double sf = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) sf += ai[c] * V[c] - pi[c] * V[3 + c];
    const double ei = (double)__builtin_amdgcn_exp2f(-y2);
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        double pd = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) pd += ai[c] * V[6 + 3 * c + e] - pi[c] * V[15 + 3 * c + e];
        out[(size_t)e * n + ti] = V[e] * ei;
        out[(size_t)(3 + e) * n + ti] = ((double)xh[e] * sf - pd) * (ei * pds);
    }
}
