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
// sources and row tile, the exponents t_ij = |x_j|^2 - 2 y_i . x_j - S come
// from two v_mfma_f32_16x16x4f32 (sources x (x, |x|^2 - S) against targets
// x (-2 y, 1); y are the targets' offsets from the wave's centroid, so the
// expanded form keeps its error bound of |y| <= rt), the kernel values
// k' = exp2(-t) = 2^S exp2(-|x_j|^2 + 2 y_i . x_j) from v_exp_f32, and the
// products with the columns from v_mfma_f32_16x16x32_f16 on a split of both
// factors into FP16 high and low halves (hi.hi + hi.lo + lo.hi: 22 bits of
// each product, FP32 accumulation).  The wave's scale 2^S keeps k' below
// 2^15, and power-of-two column scales from the chunk's maxima keep the
// columns below 2^14, so both halves stay normal down to 2^-29 of the
// largest value; the scales are powers of two from each staged tile's
// maxima.  The FP32 sums of one staged tile (256 sources) are added to FP64
// totals with the scales removed; exp2(-|y_i|^2) and the combination with
// the target's own vectors happen once per block in FP64, one thread per
// target.  With the group bounds of the culling (BatchView grpv), a wave
// skips each 64-source group whose kernel values against all its targets
// are below the culling level (the exact bound on r^2, as the direct form).
//
// A block whose wave tiles are wider than the expanded form's bound (small
// problems, narrow kernels), or a run with the expanded form switched off,
// takes the vector path (pairs_block with four targets per lane) instead.
#pragma once

typedef float mx_v4f __attribute__((ext_vector_type(4)));
typedef _Float16 mx_v8h __attribute__((ext_vector_type(8)));

constexpr int kMxTI = 256;          // targets per block
constexpr int kMxTJ = 256;          // sources per staged tile
constexpr int kMxRow = kMxTJ + 8;   // halves per column row (16-byte pad: the rows of one read hit distinct banks)
constexpr int kMxCols = 24;         // columns per source (two tiles of 16, the last 8 zero)

struct MxSmem {
    float4 crd[kMxTJ];                  // block-anchored scaled FP32 offsets of the staged sources
    _Float16 colh[2][16][kMxRow];       // columns, high halves: [column tile][column][source]
    _Float16 coll[2][16][kMxRow];       // low halves
    double dred[kWaves][4];             // block reductions
    int flag[kWaves];
};

using OpBwdMxFallback = OpBwd<3, -1, PrecF32, 4>;
constexpr size_t kMxSmemBytes =
    sizeof(MxSmem) > PairSmem<OpBwdMxFallback>::bytes ? sizeof(MxSmem) : PairSmem<OpBwdMxFallback>::bytes;

// 2^(top - e) for m = f 2^e, f in [1/2, 1): m times it lies below 2^top
__device__ __forceinline__ float mx_pow2_below(float m, int top)
{
    if (!(m > 0.0f) || !(m < 3.0e38f)) return 1.0f;
    int e;
    (void)frexpf(m, &e);
    return ldexpf(1.0f, top - e);
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
__global__ __launch_bounds__(kBlock, 2) void kb_bwd_mx(BatchView v, int t, const int4 *__restrict__ items)
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
    float xh[3], dw[3], y[3], y2 = 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) xh[c] = real ? (float)((qi[c] - anc[c]) * kap) : 0.0f;
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
    __syncthreads();   // dred reads done
    if (lane == 0) S.flag[wave] = ok ? 1 : 0;
    __syncthreads();
    const bool all_ok = S.flag[0] && S.flag[1] && S.flag[2] && S.flag[3];
    if (!all_ok) {
        __syncthreads();   // flags read before the vector path reuses the memory
        using Op = OpBwdMxFallback;
        typename Op::Args A{q, a, p, n, (float)d.iw, (float)d.iw2, cull_r2(v.cull), expand_r2(v.expand != 0), kap,
                            pds, v.exec, grp};
        pairs_block<Op>(A, n, tile, j0, j1, out, smem);
        return;
    }

    // the wave's kernel scale: k' = 2^S exp2(-|x_j|^2 + 2 y.x_j) <= 2^(S + rt^2) <= 2^15
    const float Sk = floorf(15.0f - rtw2);
    const double k2 = ldexp(1.0, -(int)Sk);
    // culling: the wave's centroid (unscaled) and radius (scaled)
    double aw[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) aw[c] = anc[c] + (double)dw[c] / kap;
    const float rtw = __builtin_sqrtf(rtw2), cut2 = cull_r2(v.cull);
    const int nn = lane & 15, g = lane >> 4;
    // target operand of the exponent products, per row tile: (-2 y, 1)
    float bs[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        const int src = 16 * mt + nn;
        const float y0 = __shfl(y[0], src, 64), y1 = __shfl(y[1], src, 64), yz = __shfl(y[2], src, 64);
        bs[mt] = g == 0 ? -2.0f * y0 : g == 1 ? -2.0f * y1 : g == 2 ? -2.0f * yz : 1.0f;
    }
    // source operand of the exponent products: lane (m, k) holds component
    // k of source 8(m/4) + m%4 (+4 for the second half), in the wave's frame
    // (x' = x - dw, and |x'|^2 - S); the selection is arithmetic, so the
    // loop has no branch
    const float sel0 = g == 0 ? 1.0f : 0.0f, sel1 = g == 1 ? 1.0f : 0.0f, sel2 = g == 2 ? 1.0f : 0.0f,
                sel3 = g == 3 ? 1.0f : 0.0f;
    auto aval = [&](const float4 cc) {
        const float x0 = cc.x - dw[0], x1 = cc.y - dw[1], x2 = cc.z - dw[2];
        const float n2 = __builtin_fmaf(x0, x0, __builtin_fmaf(x1, x1, __builtin_fmaf(x2, x2, -Sk)));
        return __builtin_fmaf(sel0, x0, __builtin_fmaf(sel1, x1, __builtin_fmaf(sel2, x2, sel3 * n2)));
    };
    const int js0 = 8 * (nn >> 2) + (nn & 3), co0 = 8 * g;

    mx_v4f C[4][2];
    double acc[4][2][4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            C[mt][nt] = mx_v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[mt][nt][r] = 0.0;
        }
    const mx_v4f zero4 = {0.0f, 0.0f, 0.0f, 0.0f};
    unsigned long long executed = 0;

    for (int jt = j0; jt < j1; jt += kMxTJ) {
        const int cnt = min(kMxTJ, j1 - jt);
        __syncthreads();   // the previous tile is consumed, dred free
        // staging, first pass: the source's offsets and vectors, and the
        // tile's maxima of |x|, |a|, |a - p|
        const int tj = threadIdx.x;
        float xs[3] = {0.0f, 0.0f, 0.0f}, av[3] = {0.0f, 0.0f, 0.0f}, wv[3] = {0.0f, 0.0f, 0.0f};
        double mx3[3] = {0.0, 0.0, 0.0};
        if (tj < cnt) {
            const int j = jt + tj;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double ac = a[(size_t)c * n + j], pc = p[(size_t)c * n + j];
                xs[c] = (float)((q[(size_t)c * n + j] - anc[c]) * kap);
                av[c] = (float)ac;
                wv[c] = (float)(ac - pc);
                mx3[0] = fmax(mx3[0], (double)fabsf(xs[c]));
                mx3[1] = fmax(mx3[1], (double)fabsf(av[c]));
                mx3[2] = fmax(mx3[2], (double)fabsf(wv[c]));
            }
        }
        S.crd[tj] = make_float4(xs[0], xs[1], xs[2], 0.0f);
        mx_block_sum3(mx3, S, lane, wave, true);
        const float sx = mx_pow2_below((float)mx3[0], 0), sa = mx_pow2_below((float)mx3[1], 14),
                    sw = mx_pow2_below((float)mx3[2], 14);
        // second pass: the 24 scaled columns as FP16 halves
        {
            float col[kMxCols];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                col[c] = wv[c] * sw;
                col[3 + c] = av[c] * sa;
#pragma unroll
                for (int e = 0; e < 3; ++e) {
                    col[6 + 3 * c + e] = (wv[c] * sw) * (xs[e] * sx);
                    col[15 + 3 * c + e] = (av[c] * sa) * (xs[e] * sx);
                }
            }
#pragma unroll
            for (int k = 0; k < 32; ++k) {
                const float vv = k < kMxCols ? col[k < kMxCols ? k : 0] : 0.0f;
                const _Float16 h = (_Float16)vv;
                S.colh[k >> 4][k & 15][tj] = h;
                S.coll[k >> 4][k & 15][tj] = (_Float16)(vv - (float)h);
            }
        }
        // FP64 factor of the lane's two columns: the inverse column scale and 2^-S
        auto cinv = [&](int col) -> double {
            if (col < 3) return k2 / sw;
            if (col < 6) return k2 / sa;
            if (col < 15) return k2 / ((double)sw * sx);
            if (col < 24) return k2 / ((double)sa * sx);
            return 0.0;
        };
        const double cs0 = cinv(nn), cs1 = cinv(16 + nn);
        // the wave's live 64-source groups of the tile (bit q: group q)
        unsigned live = 0xFu;
        if (grp != nullptr && wcount > 0) {
            bool keep = true;
            if (lane < 4) {
                const int gi = (jt >> 6) + lane;
                if (64 * lane < cnt) {
                    const double4 b = grp[gi];
                    const double dx = b.x - aw[0], dy = b.y - aw[1], dz = b.z - aw[2];
                    const float ms = (float)(fmax(0.0, sqrt(dx * dx + dy * dy + dz * dz) - b.w) * kap);
                    const float gap = ms - rtw;
                    keep = !(gap > 0.0f && gap * gap > cut2);
                }
            }
            live = (unsigned)__ballot(lane < 4 && keep) & 0xFu;
        }
        __syncthreads();
        if (wcount > 0 && live != 0u) {
            const int nkb = (cnt + 31) >> 5;
            // the live K-blocks of the tile as a bit mask (uniform)
            unsigned kmask = 0u;
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (k < nkb && ((live >> (k >> 1)) & 1u)) kmask |= 1u << k;
            while (kmask != 0u) {
                const int kb = __builtin_ctz(kmask);
                kmask &= kmask - 1u;
                executed += (unsigned long long)min(32, cnt - 32 * kb) * (unsigned long long)wcount;
                const float A0 = aval(S.crd[kb * 32 + js0]), A1 = aval(S.crd[kb * 32 + js0 + 4]);
                // all eight exponent products of the K-block first, so the
                // exponentials of one row tile overlap the products of the next
                mx_v4f s0[4], s1[4];
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) {
                    s0[mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(A0, bs[mt], zero4, 0, 0, 0);
                    s1[mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(A1, bs[mt], zero4, 0, 0, 0);
                }
                const mx_v8h h0c = *reinterpret_cast<const mx_v8h *>(&S.colh[0][nn][kb * 32 + co0]);
                const mx_v8h l0c = *reinterpret_cast<const mx_v8h *>(&S.coll[0][nn][kb * 32 + co0]);
                const mx_v8h h1c = *reinterpret_cast<const mx_v8h *>(&S.colh[1][nn][kb * 32 + co0]);
                const mx_v8h l1c = *reinterpret_cast<const mx_v8h *>(&S.coll[1][nn][kb * 32 + co0]);
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) {
                    // exponents: lane (n, g) has target n against sources 8g .. 8g+7;
                    // k = hi + lo: hi truncated to FP16 (so k - hi is exact in
                    // FP32), lo rounded to nearest
                    mx_v8h ah, al;
#pragma unroll
                    for (int r = 0; r < 4; r += 2) {
                        const float e0 = __builtin_amdgcn_exp2f(-s0[mt][r]), e1 = __builtin_amdgcn_exp2f(-s0[mt][r + 1]);
                        const float f0 = __builtin_amdgcn_exp2f(-s1[mt][r]), f1 = __builtin_amdgcn_exp2f(-s1[mt][r + 1]);
                        const auto he = __builtin_amdgcn_cvt_pkrtz(e0, e1), hf = __builtin_amdgcn_cvt_pkrtz(f0, f1);
                        ah[r] = he[0];
                        ah[r + 1] = he[1];
                        ah[4 + r] = hf[0];
                        ah[5 + r] = hf[1];
                        al[r] = (_Float16)(e0 - (float)he[0]);
                        al[r + 1] = (_Float16)(e1 - (float)he[1]);
                        al[4 + r] = (_Float16)(f0 - (float)hf[0]);
                        al[5 + r] = (_Float16)(f1 - (float)hf[1]);
                    }
                    C[mt][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, h0c, C[mt][0], 0, 0, 0);
                    C[mt][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, h1c, C[mt][1], 0, 0, 0);
                    C[mt][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, l0c, C[mt][0], 0, 0, 0);
                    C[mt][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, l1c, C[mt][1], 0, 0, 0);
                    C[mt][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, h0c, C[mt][0], 0, 0, 0);
                    C[mt][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, h1c, C[mt][1], 0, 0, 0);
                }
            }
            // the tile's FP32 sums into the FP64 totals, scales removed
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    acc[mt][0][r] += (double)C[mt][0][r] * cs0;
                    acc[mt][1][r] += (double)C[mt][1][r] * cs1;
                }
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) C[mt][0] = C[mt][1] = zero4;
        }
    }
    if (v.exec != nullptr && lane == 0 && executed != 0) atomicAdd(v.exec + 1, executed);

    // combination with the targets' own vectors, one thread per target:
    // lane (n, g) holds columns n and 16 + n of targets 16 mt + 4 g + r
    static_assert(sizeof(double) * kWaves * 64 * kMxCols <= kMxSmemBytes, "epilogue buffer");
    double *E = reinterpret_cast<double *>(smem);   // [wave][target][column], 48 KiB
    __syncthreads();   // the last tile is consumed
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int tl = 16 * mt + 4 * g + r;
            E[((size_t)wave * 64 + tl) * kMxCols + nn] = acc[mt][0][r];
            if (nn < kMxCols - 16) E[((size_t)wave * 64 + tl) * kMxCols + 16 + nn] = acc[mt][1][r];
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
