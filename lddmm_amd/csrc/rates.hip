// rates.hip -- measured instruction rates of the chip, the denominators of
// the transcendental and VALU utilisation that the benchmark reports
// (lddmm_measure_rates): v_exp_f32, FP32 FMA and FP64 FMA, each timed over
// a full grid of independent dependency chains.
#include "kernels.hpp"

namespace lddmm {

namespace {

constexpr int kChains = 8, kSteps = 2048, kRateBlocks = 256 * 16;

// 2^-x on 8 independent chains: x stays in [1/2, 1]; one v_exp_f32 (negated
// operand) per step
__global__ __launch_bounds__(256) void k_rate_exp(float *out, float seed)
{
    float x[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) x[c] = seed + 0.01f * c + 1e-7f * threadIdx.x;
    for (int s = 0; s < kSteps; ++s)
#pragma unroll
        for (int c = 0; c < kChains; ++c) x[c] = __builtin_amdgcn_exp2f(-x[c]);
    float t = 0.0f;
#pragma unroll
    for (int c = 0; c < kChains; ++c) t += x[c];
    if (t == 12345.0f) out[blockIdx.x] = t;   // keeps the chains live
}

template <typename real>
__global__ __launch_bounds__(256) void k_rate_fma(real *out, real a, real b)
{
    real x[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) x[c] = real(0.5) + real(0.01) * c + real(1e-7) * threadIdx.x;
    for (int s = 0; s < kSteps; ++s)
#pragma unroll
        for (int c = 0; c < kChains; ++c) x[c] = fma(x[c], a, b);
    real t = real(0);
#pragma unroll
    for (int c = 0; c < kChains; ++c) t += x[c];
    if (t == real(12345)) out[blockIdx.x] = t;
}

template <typename L>
double time_ms(L &&launch, hipStream_t st)
{
    hipEvent_t a, b;
    LDDMM_HIP(hipEventCreate(&a));
    LDDMM_HIP(hipEventCreate(&b));
    launch();   // warm-up
    LDDMM_HIP(hipEventRecord(a, st));
    for (int r = 0; r < 5; ++r) launch();
    LDDMM_HIP(hipEventRecord(b, st));
    LDDMM_HIP(hipEventSynchronize(b));
    float ms = 0.0f;
    LDDMM_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms / 5.0;
}

}  // namespace

void measure_rates(double *out, hipStream_t st)
{
    void *buf = nullptr;
    LDDMM_HIP(hipMalloc(&buf, kRateBlocks * sizeof(double)));
    const double ops = (double)kRateBlocks * 256 * kChains * kSteps;
    try {
        const double te = time_ms([&] { k_rate_exp<<<kRateBlocks, 256, 0, st>>>((float *)buf, 0.7f); }, st);
        LDDMM_LAUNCH_CHECK();
        const double tf = time_ms([&] { k_rate_fma<float><<<kRateBlocks, 256, 0, st>>>((float *)buf, 0.999f, 1e-4f); },
                                  st);
        LDDMM_LAUNCH_CHECK();
        const double td =
            time_ms([&] { k_rate_fma<double><<<kRateBlocks, 256, 0, st>>>((double *)buf, 0.999, 1e-4); }, st);
        LDDMM_LAUNCH_CHECK();
        out[0] = ops / (te * 1e-3);
        out[1] = ops / (tf * 1e-3);
        out[2] = ops / (td * 1e-3);
    } catch (...) {
        (void)hipFree(buf);
        throw;
    }
    LDDMM_HIP(hipFree(buf));
}

}  // namespace lddmm
