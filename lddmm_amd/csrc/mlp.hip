// mlp.hip -- point features and the BF16-MFMA multilayer perceptron of the
// learned warm start.
//
// GEMM: C (M x N, FP32) = op(A) op(B) with A, B in BF16, on
// v_mfma_f32_32x32x16_bf16.  A block is 4 wave64; wave w owns the 32 x 32
// output tile at rows blockIdx.y*128 + 32w, columns blockIdx.x*32, and
// walks K in steps of 16.  Operand lane layout of the instruction: lane l
// holds A[l % 32][8 (l / 32) + v] and B[8 (l / 32) + v][l % 32], v < 8;
// accumulator element e of lane l is D[(e % 4) + 4 (l / 32) + 8 (e / 4)][l % 32].
// Long reductions (weight gradients, K = rows) are split over blockIdx.z
// into FP32 partials summed in a fixed order afterwards.
#include <cmath>
#include <cstring>

#include "mlp.hpp"
#include "pair_driver.hpp"

namespace lddmm {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

template <bool TA, bool TB>
__global__ __launch_bounds__(kBlock) void k_gemm(const __bf16 *__restrict__ A, int lda, const __bf16 *__restrict__ B,
                                                 int ldb, float *__restrict__ C, int ldc, int M, int N, int K,
                                                 int kchunk, long long split_stride)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m0 = blockIdx.y * 128 + wave * 32;
    if (m0 >= M) return;
    const int n0 = blockIdx.x * 32;
    const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
    const int row = m0 + (lane & 31), col = n0 + (lane & 31), kk = 8 * (lane >> 5);
    f32x16 acc;
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
    for (int k = kb; k < ke; k += 16) {
        bf16x8 a, b;
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            a[v] = TA ? A[(size_t)(k + kk + v) * lda + row] : A[(size_t)row * lda + k + kk + v];
            b[v] = TB ? B[(size_t)col * ldb + k + kk + v] : B[(size_t)(k + kk + v) * ldb + col];
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    float *Cs = C + (size_t)blockIdx.z * split_stride;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int i = (e & 3) + 4 * (lane >> 5) + 8 * (e >> 2);
        Cs[(size_t)(m0 + i) * ldc + n0 + (lane & 31)] = acc[e];
    }
}

__global__ void k_sum_parts(const float *__restrict__ part, int splits, long long n, float *__restrict__ out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int k = 0; k < splits; ++k) s += part[(size_t)k * n + i];
    out[i] = s;
}

__global__ void k_cast_bf16(const float *__restrict__ x, __bf16 *__restrict__ y, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (__bf16)x[i];
}

// H = act(C + b) in place, plus the BF16 copy fed to the next GEMM
__global__ void k_bias_act(float *__restrict__ h, const float *__restrict__ b, int rows, int cols, int relu,
                           __bf16 *__restrict__ hb)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)rows * cols) return;
    float v = h[i] + b[i % cols];
    if (relu) v = v > 0.0f ? v : 0.0f;
    h[i] = v;
    if (hb) hb[i] = (__bf16)v;
}

// mean squared error over the first `outs` columns of the valid rows;
// dz = d loss / d out, per-block loss partials in FP64
__global__ __launch_bounds__(kBlock) void k_loss(const float *__restrict__ out, const float *__restrict__ y,
                                                 int valid, int rows, int outs, float scale,
                                                 float *__restrict__ dz, __bf16 *__restrict__ dzb,
                                                 double *__restrict__ parts)
{
    __shared__ double red[kBlock];
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    double l = 0.0;
    if (i < (long long)rows * kMlpOut) {
        const int r = (int)(i / kMlpOut), c = (int)(i % kMlpOut);
        float g = 0.0f;
        if (r < valid && c < outs) {
            const float d = out[i] - y[(size_t)r * outs + c];
            l = (double)d * d;
            g = 2.0f * scale * d;
        }
        dz[i] = g;
        dzb[i] = (__bf16)g;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) parts[blockIdx.x] = red[0];
}

// dz may alias dh (each element is read before it is written by the same thread)
__global__ void k_relu_back(const float *dh, const float *__restrict__ h, long long n, float *dz,
                            __bf16 *__restrict__ dzb)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = h[i] > 0.0f ? dh[i] : 0.0f;
    dz[i] = g;
    dzb[i] = (__bf16)g;
}

// db[c] = sum over rows of dz[r][c], one block per column, fixed order
__global__ __launch_bounds__(kBlock) void k_colsum(const float *__restrict__ dz, int rows, int cols,
                                                   float *__restrict__ db)
{
    __shared__ float red[kBlock];
    const int c = blockIdx.x;
    float s = 0.0f;
    for (int r = threadIdx.x; r < rows; r += kBlock) s += dz[(size_t)r * cols + c];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] = red[0];
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int n, float lr, float b1, float b2, float eps, float bc1, float bc2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

__global__ void k_take_cols(const float *__restrict__ out, int rows, int outs, float *__restrict__ y)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)rows * outs) return;
    const int r = (int)(i / outs), c = (int)(i % outs);
    y[i] = out[(size_t)r * kMlpOut + c];
}

__global__ void k_pad_rows(const float *__restrict__ x, int rows, int cap, __bf16 *__restrict__ xb)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)cap * kFeat) return;
    const int r = (int)(i / kFeat);
    xb[i] = r < rows ? (__bf16)x[i] : (__bf16)0.0f;
}

inline int nb(long long n, int b = 256) { return (int)((n + b - 1) / b); }

// ------------------------------------------------------------- features
// moments of a source set j about each target vertex i (FP64):
// [m0, m1 (D), mu (D)] with w = a_j exp(-|x_i - c_j|^2 / s^2),
// m1 = sum w (c_j - x_i) / s, mu = sum w (u_j . n_i) u_j
template <int D_>
struct OpMoments {
    static constexpr int D = D_, F = D_ + 1, OUT = 1 + 2 * D_, R = 1;
    using real = double;
    using lds_t = double;
    using crd_t = double;
    static constexpr bool kCull = false;
    struct Args {
        const double *x, *nrm;
        int ni;
        const double *c, *u, *a;
        int nj;
        double is2, is;
    };
    struct Tgt {
        double n[D];
    };
    struct Acc {
        double m0, m1[D], mu[D];
    };
    __device__ __forceinline__ static void load_src(const Args &A, int j, double (&x)[D], lds_t (&f)[F])
    {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            x[c] = A.c[(size_t)c * A.nj + j];
            f[c] = A.u[(size_t)c * A.nj + j];
        }
        f[D] = A.a[j];
    }
    __device__ __forceinline__ static void load_tgt(const Args &A, int i, Tgt &t, double (&x)[D])
    {
#pragma unroll
        for (int c = 0; c < D; ++c) {
            x[c] = i >= 0 ? A.x[(size_t)c * A.ni + i] : 0.0;
            t.n[c] = i >= 0 ? A.nrm[(size_t)c * A.ni + i] : 0.0;
        }
    }
    __device__ __forceinline__ static void zero(Acc &a)
    {
        a.m0 = 0.0;
#pragma unroll
        for (int c = 0; c < D; ++c) a.m1[c] = a.mu[c] = 0.0;
    }
    __device__ __forceinline__ static void pair(const Args &A, const Tgt &t, const crd_t (&xi)[D],
                                                const crd_t (&xj)[D], const lds_t (&fj)[F], Acc &acc)
    {
        double df[D], r2 = 0.0, un = 0.0;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            df[c] = xj[c] - xi[c];
            r2 += df[c] * df[c];
            un += fj[c] * t.n[c];
        }
        const double w = fj[D] * exp(-r2 * A.is2);
        acc.m0 += w;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            acc.m1[c] += w * df[c] * A.is;
            acc.mu[c] += w * un * fj[c];
        }
    }
    __device__ __forceinline__ static void flush(const Args &, const Acc &a, double (&tot)[OUT])
    {
        tot[0] += a.m0;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            tot[1 + c] += a.m1[c];
            tot[1 + D + c] += a.mu[c];
        }
    }
};

__global__ void k_sum_moment_splits(const double *__restrict__ part, int splits, int comps, int n,
                                    double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= n) return;
    double s = 0.0;
    for (int k = 0; k < splits; ++k) s += part[((size_t)k * comps + c) * n + i];
    out[(size_t)c * n + i] = s;
}

}  // namespace

// ======================================================================
void point_features(int dim, const double *q, int n, const int *elm, int ne, const double *tc, const double *tu,
                    const double *ta, int nt, double knl_width, float *feat, hipStream_t st)
{
    if (dim != 3) throw Error(kErrUnsupported, "point features are defined for 3D surfaces");
    if (n <= 0 || ne <= 0) throw Error(kErrArg, "empty source surface");
    // area-weighted vertex normals (host)
    std::vector<double> nrm((size_t)3 * n, 0.0);
    for (int e = 0; e < ne; ++e) {
        const int *v = elm + (size_t)3 * e;
        for (int k = 0; k < 3; ++k)
            if (v[k] < 0 || v[k] >= n) throw Error(kErrArg, "element table holds a vertex outside [0, N)");
        const double *p0 = q + 3 * (size_t)v[0], *p1 = q + 3 * (size_t)v[1], *p2 = q + 3 * (size_t)v[2];
        double a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]}, b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
        double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
        for (int k = 0; k < 3; ++k)
            for (int d = 0; d < 3; ++d) nrm[(size_t)3 * v[k] + d] += c[d];
    }
    std::vector<double> xs((size_t)3 * n), ns((size_t)3 * n), ones(n, 1.0), zeros((size_t)3 * n, 0.0);
    for (int i = 0; i < n; ++i) {
        double *m = nrm.data() + 3 * (size_t)i;
        double l = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        for (int d = 0; d < 3; ++d) {
            xs[(size_t)d * n + i] = q[3 * (size_t)i + d];
            ns[(size_t)d * n + i] = l > 0 ? m[d] / l : 0.0;
        }
    }
    std::vector<double> hc((size_t)3 * std::max(nt, 1)), hu((size_t)3 * std::max(nt, 1));
    for (int j = 0; j < nt; ++j)
        for (int d = 0; d < 3; ++d) {
            hc[(size_t)d * nt + j] = tc[3 * (size_t)j + d];
            hu[(size_t)d * nt + j] = tu[3 * (size_t)j + d];
        }
    Slab s;
    double *dx, *dn, *dc, *du, *da, *d1, *dz, *part, *mom;
    s.reserve(&dx, xs.size());
    s.reserve(&dn, ns.size());
    s.reserve(&dc, hc.size());
    s.reserve(&du, hu.size());
    s.reserve(&da, std::max(nt, 1));
    s.reserve(&d1, n);
    s.reserve(&dz, zeros.size());
    using Op = OpMoments<3>;
    const PairSplit spt = make_split_ti(n, std::max(nt, 1), 64, 1024), sps = make_split_ti(n, n, 64, 1024);
    s.reserve(&part, (size_t)std::max(spt.splits, sps.splits) * Op::OUT * n);
    s.reserve(&mom, (size_t)Op::OUT * n * 6);
    s.allocate();
    LDDMM_HIP(hipMemcpyAsync(dx, xs.data(), xs.size() * 8, hipMemcpyHostToDevice, st));
    LDDMM_HIP(hipMemcpyAsync(dn, ns.data(), ns.size() * 8, hipMemcpyHostToDevice, st));
    if (nt > 0) {
        LDDMM_HIP(hipMemcpyAsync(dc, hc.data(), (size_t)3 * nt * 8, hipMemcpyHostToDevice, st));
        LDDMM_HIP(hipMemcpyAsync(du, hu.data(), (size_t)3 * nt * 8, hipMemcpyHostToDevice, st));
        LDDMM_HIP(hipMemcpyAsync(da, ta, (size_t)nt * 8, hipMemcpyHostToDevice, st));
    }
    LDDMM_HIP(hipMemcpyAsync(d1, ones.data(), (size_t)n * 8, hipMemcpyHostToDevice, st));
    LDDMM_HIP(hipMemcpyAsync(dz, zeros.data(), zeros.size() * 8, hipMemcpyHostToDevice, st));
    const double widths[3] = {kFeatWidths[0], kFeatWidths[1], kFeatWidths[2]};
    for (int k = 0; k < 3; ++k) {
        const double s2 = widths[k] * widths[k];
        for (int pass = 0; pass < 2; ++pass) {
            if (pass == 0 && nt == 0) {
                LDDMM_HIP(hipMemsetAsync(mom + (size_t)(2 * k) * Op::OUT * n, 0, (size_t)Op::OUT * n * 8, st));
                continue;
            }
            Op::Args A = pass == 0 ? Op::Args{dx, dn, n, dc, du, da, nt, 1.0 / s2, 1.0 / widths[k]}
                                   : Op::Args{dx, dn, n, dx, dz, d1, n, 1.0 / s2, 1.0 / widths[k]};
            const PairSplit &sp = pass == 0 ? spt : sps;
            launch_pairs<Op>(A, n, pass == 0 ? nt : n, sp, 0, part, st);
            dim3 g(ceil_div(n, 256), Op::OUT);
            k_sum_moment_splits<<<g, 256, 0, st>>>(part, sp.splits, Op::OUT, n, mom + (size_t)(2 * k + pass) * Op::OUT * n);
            LDDMM_LAUNCH_CHECK();
        }
    }
    std::vector<double> hm((size_t)Op::OUT * n * 6);
    LDDMM_HIP(hipMemcpyAsync(hm.data(), mom, hm.size() * 8, hipMemcpyDeviceToHost, st));
    LDDMM_HIP(hipStreamSynchronize(st));
    const double inv_sqrt_n = 1.0 / std::sqrt((double)n);
    for (int i = 0; i < n; ++i) {
        float *f = feat + (size_t)kFeat * i;
        int o = 0;
        for (int d = 0; d < 3; ++d) f[o++] = (float)q[3 * (size_t)i + d];
        for (int d = 0; d < 3; ++d) f[o++] = (float)ns[(size_t)d * n + i];
        for (int k = 0; k < 3; ++k) {
            // target mass, mean offset to the target mass, orientation moment
            const double *mt = hm.data() + (size_t)(2 * k) * Op::OUT * n;
            const double m0 = mt[i], den = m0 + 1e-3 * widths[k] * widths[k];
            f[o++] = (float)m0;
            for (int c = 0; c < 3; ++c) f[o++] = (float)(mt[(size_t)(1 + c) * n + i] * widths[k] / den);
            for (int c = 0; c < 3; ++c) f[o++] = (float)mt[(size_t)(4 + c) * n + i];
        }
        for (int k = 0; k < 3; ++k) f[o++] = (float)(hm[(size_t)(2 * k + 1) * Op::OUT * n + i] / n);
        f[o++] = (float)knl_width;
        f[o++] = (float)inv_sqrt_n;
    }
}

// ======================================================================
Mlp::Mlp(int hidden, int layers, unsigned long long seed, hipStream_t st) : st_(st), hidden_(hidden), layers_(layers)
{
    if (layers < 2 || hidden < 32 || hidden % 32) throw Error(kErrArg, "MLP needs >= 2 layers and hidden % 32 == 0");
    for (int l = 0; l < layers; ++l) {
        din_.push_back(l == 0 ? kFeat : hidden);
        dout_.push_back(l == layers - 1 ? kMlpOut : hidden);
        woff_.push_back(nparam_);
        nparam_ += din_[l] * dout_[l];
        boff_.push_back(nparam_);
        nparam_ += dout_[l];
    }
    slab_.reserve(&p_, nparam_);
    slab_.reserve(&g_, nparam_);
    slab_.reserve(&m_, nparam_);
    slab_.reserve(&v_, nparam_);
    slab_.reserve(&pb_, nparam_);
    slab_.allocate();
    // He-uniform weights from a splitmix64 stream, zero biases; the padded
    // output columns (>= 3) start at zero
    std::vector<float> h(nparam_, 0.0f);
    unsigned long long x = seed;
    auto next = [&]() {
        x += 0x9E3779B97F4A7C15ULL;
        unsigned long long z = x;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int l = 0; l < layers; ++l) {
        const double lim = std::sqrt(6.0 / din_[l]) * (l == layers - 1 ? 0.1 : 1.0);
        for (int i = 0; i < din_[l]; ++i)
            for (int j = 0; j < dout_[l]; ++j) {
                const double r = (2.0 * next() - 1.0) * lim;
                h[woff_[l] + (size_t)i * dout_[l] + j] = (l == layers - 1 && j >= 3) ? 0.0f : (float)r;
            }
    }
    set_params(h.data());
}

Mlp::~Mlp()
{
    for (void *p : {(void *)act_, (void *)dact_, (void *)part_, (void *)lossbuf_, (void *)actb_, (void *)dactb_,
                    (void *)xb_})
        if (p) (void)hipFree(p);
}

void Mlp::get_params(float *out) const
{
    LDDMM_HIP(hipMemcpyAsync(out, p_, (size_t)nparam_ * 4, hipMemcpyDeviceToHost, st_));
    LDDMM_HIP(hipStreamSynchronize(st_));
}

void Mlp::set_params(const float *in)
{
    LDDMM_HIP(hipMemcpyAsync(p_, in, (size_t)nparam_ * 4, hipMemcpyHostToDevice, st_));
    LDDMM_HIP(hipMemsetAsync(m_, 0, (size_t)nparam_ * 4, st_));
    LDDMM_HIP(hipMemsetAsync(v_, 0, (size_t)nparam_ * 4, st_));
    step_ = 0;
    sync_bf16_weights();
    LDDMM_HIP(hipStreamSynchronize(st_));
}

void Mlp::sync_bf16_weights()
{
    k_cast_bf16<<<nb(nparam_), 256, 0, st_>>>(p_, pb_, nparam_);
    LDDMM_LAUNCH_CHECK();
}

void Mlp::ensure_rows(int rows)
{
    const int cap = ((rows + kMlpRows - 1) / kMlpRows) * kMlpRows;
    if (cap <= cap_) return;
    for (void *p : {(void *)act_, (void *)dact_, (void *)part_, (void *)lossbuf_, (void *)actb_, (void *)dactb_,
                    (void *)xb_})
        if (p) (void)hipFree(p);
    act_ = dact_ = part_ = nullptr;
    lossbuf_ = nullptr;
    actb_ = dactb_ = xb_ = nullptr;
    cap_ = 0;   // stays 0 if an allocation below fails
    long long per_row = 0;
    for (int l = 0; l < layers_; ++l) per_row += dout_[l];
    auto alloc = [&](void **p, size_t bytes) {
        if (hipMalloc(p, std::max<size_t>(bytes, 256)) != hipSuccess) throw Error(kErrAlloc, "MLP activations");
    };
    alloc((void **)&act_, (size_t)cap * per_row * 4);
    alloc((void **)&dact_, (size_t)cap * per_row * 4);
    alloc((void **)&actb_, (size_t)cap * per_row * 2);
    alloc((void **)&dactb_, (size_t)cap * per_row * 2);
    alloc((void **)&xb_, (size_t)cap * kFeat * 2);
    alloc((void **)&lossbuf_, (size_t)nb((long long)cap * kMlpOut) * 8);
    // split-K partials of a weight gradient: at most 64 splits of hidden x hidden
    alloc((void **)&part_, (size_t)64 * std::max(hidden_, kFeat) * std::max(hidden_, kMlpOut) * 4);
    cap_ = cap;
}

void gemm_bf16(bool ta, bool tb, const __bf16 *A, int lda, const __bf16 *B, int ldb, float *C, int ldc, int M, int N, int K,
          int splits, float *part, hipStream_t st)
{
    if (M % 32 || N % 32 || K % 16) throw Error(kErrArg, "GEMM shapes must be multiples of 32 x 32 x 16");
    int kchunk = ((ceil_div(K, splits) + 15) / 16) * 16;
    splits = ceil_div(K, kchunk);
    dim3 grid(N / 32, ceil_div(M, 128), splits);
    float *out = splits > 1 ? part : C;
    const int ldo = splits > 1 ? N : ldc;
    const long long stride = (long long)M * N;
    if (!ta && !tb) k_gemm<false, false><<<grid, kBlock, 0, st>>>(A, lda, B, ldb, out, ldo, M, N, K, kchunk, stride);
    else if (ta && !tb) k_gemm<true, false><<<grid, kBlock, 0, st>>>(A, lda, B, ldb, out, ldo, M, N, K, kchunk, stride);
    else if (!ta && tb) k_gemm<false, true><<<grid, kBlock, 0, st>>>(A, lda, B, ldb, out, ldo, M, N, K, kchunk, stride);
    else k_gemm<true, true><<<grid, kBlock, 0, st>>>(A, lda, B, ldb, out, ldo, M, N, K, kchunk, stride);
    LDDMM_LAUNCH_CHECK();
    if (splits > 1) {
        if (ldc != N) throw Error(kErrArg, "split-K output must be dense");
        k_sum_parts<<<nb(stride), 256, 0, st>>>(part, splits, stride, C);
        LDDMM_LAUNCH_CHECK();
    }
}

void Mlp::forward(const float *x, int rows, float *y)
{
    ensure_rows(rows);
    const int cap = ((rows + kMlpRows - 1) / kMlpRows) * kMlpRows;
    k_pad_rows<<<nb((long long)cap * kFeat), 256, 0, st_>>>(x, rows, cap, xb_);
    LDDMM_LAUNCH_CHECK();
    const __bf16 *in = xb_;
    long long off = 0;
    for (int l = 0; l < layers_; ++l) {
        float *h = act_ + (size_t)cap * off;
        __bf16 *hb = actb_ + (size_t)cap * off;
        gemm_bf16(false, false, in, din_[l], pb_ + woff_[l], dout_[l], h, dout_[l], cap, dout_[l], din_[l], 1, part_, st_);
        k_bias_act<<<nb((long long)cap * dout_[l]), 256, 0, st_>>>(h, p_ + boff_[l], cap, dout_[l],
                                                                   l < layers_ - 1, hb);
        LDDMM_LAUNCH_CHECK();
        in = hb;
        off += dout_[l];
    }
    const float *out = act_ + (size_t)cap * (off - kMlpOut);
    if (y) {
        k_take_cols<<<nb((long long)rows * 3), 256, 0, st_>>>(out, rows, 3, y);
        LDDMM_LAUNCH_CHECK();
    }
}

double Mlp::train_step(const float *x, const float *y, int rows, float lr, float *grad_out)
{
    forward(x, rows, nullptr);
    const int cap = ((rows + kMlpRows - 1) / kMlpRows) * kMlpRows;
    std::vector<long long> off(layers_ + 1, 0);
    for (int l = 0; l < layers_; ++l) off[l + 1] = off[l] + dout_[l];
    // loss gradient at the output (mean over valid rows and the 3 used columns)
    const int L = layers_ - 1;
    float *dz = dact_ + (size_t)cap * off[L];
    __bf16 *dzb = dactb_ + (size_t)cap * off[L];
    const int nlb = nb((long long)cap * kMlpOut);
    k_loss<<<nlb, kBlock, 0, st_>>>(act_ + (size_t)cap * off[L], y, rows, cap, 3, 1.0f / (3.0f * rows), dz, dzb,
                                    (double *)lossbuf_);
    LDDMM_LAUNCH_CHECK();
    const int ksplit = std::max(1, std::min(64, cap / 2048));
    for (int l = L; l >= 0; --l) {
        const __bf16 *hin = l == 0 ? xb_ : actb_ + (size_t)cap * off[l - 1];
        float *gz = dact_ + (size_t)cap * off[l];
        __bf16 *gzb = dactb_ + (size_t)cap * off[l];
        // dW_l = H_{l-1}^T dZ_l  (din x dout, reduction over rows), db_l = column sums
        gemm_bf16(true, false, hin, din_[l], gzb, dout_[l], g_ + woff_[l], dout_[l], din_[l], dout_[l], cap, ksplit, part_,
             st_);
        k_colsum<<<dout_[l], kBlock, 0, st_>>>(gz, cap, dout_[l], g_ + boff_[l]);
        LDDMM_LAUNCH_CHECK();
        if (l == 0) break;
        // dH_{l-1} = dZ_l W_l^T, then through the ReLU of layer l-1
        float *dh = dact_ + (size_t)cap * off[l - 1];
        __bf16 *dhb = dactb_ + (size_t)cap * off[l - 1];
        gemm_bf16(false, true, gzb, dout_[l], pb_ + woff_[l], dout_[l], dh, din_[l], cap, din_[l], dout_[l], 1, part_, st_);
        k_relu_back<<<nb((long long)cap * din_[l]), 256, 0, st_>>>(dh, act_ + (size_t)cap * off[l - 1],
                                                                   (long long)cap * din_[l], dh, dhb);
        LDDMM_LAUNCH_CHECK();
    }
    std::vector<double> lp(nlb);
    LDDMM_HIP(hipMemcpyAsync(lp.data(), lossbuf_, (size_t)nlb * 8, hipMemcpyDeviceToHost, st_));
    if (grad_out) LDDMM_HIP(hipMemcpyAsync(grad_out, g_, (size_t)nparam_ * 4, hipMemcpyDeviceToHost, st_));
    if (lr > 0.0f) {
        ++step_;
        const float b1 = 0.9f, b2 = 0.999f;
        k_adam<<<nb(nparam_), 256, 0, st_>>>(p_, g_, m_, v_, nparam_, lr, b1, b2, 1e-8f,
                                             1.0f - std::pow(b1, (float)step_), 1.0f - std::pow(b2, (float)step_));
        LDDMM_LAUNCH_CHECK();
        sync_bf16_weights();
    }
    LDDMM_HIP(hipStreamSynchronize(st_));
    double loss = 0.0;
    for (double v : lp) loss += v;
    return loss / (3.0 * rows);
}

}  // namespace lddmm
