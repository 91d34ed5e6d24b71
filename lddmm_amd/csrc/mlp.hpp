// mlp.hpp -- the learned warm start: per-vertex point features and a small
// multilayer perceptron trained on the GPU with hand-written BF16 MFMA
// kernels (FP32 accumulation, FP32 master weights, Adam).  The network maps
// 32 features of a source vertex to a 3-vector momentum used as the initial
// L-BFGS iterate (every time slice of alpha set to it); the zero-momentum
// run stays the fallback and the yardstick.  No counterpart in the reference.
#pragma once

#include <vector>

#include "engine.hpp"

namespace lddmm {

constexpr int kFeat = 32;          // input features per vertex
constexpr int kMlpOut = 32;        // output columns (3 used, padded to one MFMA tile)
constexpr int kMlpRows = 128;      // row granularity of the GEMM kernels

constexpr double kFeatWidths[3] = {0.1, 0.25, 0.5};

// Features of every source vertex (host layout in, row-major N x 32 out):
// position (3), area-weighted unit normal (3); for each width s in
// kFeatWidths, with w_j = a_j exp(-|x - c_j|^2/s^2) over the target
// varifold's elements: the mass m0 = sum_j w_j, the mean offset
// sum_j w_j (c_j - x) / (m0 + 1e-3 s^2) (3) and the orientation moment
// sum_j w_j (u_j.n) u_j (3); the source's own vertex density
// (1/N) sum_j exp(-|x - q_j|^2/s^2) at the three widths (3); the
// deformation kernel width and 1/sqrt(N).
void point_features(int dim, const double *q, int n, const int *elm, int ne, const double *tc, const double *tu,
                    const double *ta, int nt, double knl_width, float *feat_out, hipStream_t st);

// C (M x N, row stride ldc) = op(A) op(B) on the BF16 MFMA kernel; A is
// M x K (K x M when ta), B is K x N (N x K when tb), both row-major with
// strides lda, ldb.  M, N multiples of 32, K of 16.  splits > 1 splits K
// over workgroups into `part` (splits x M x N floats, requires ldc == N),
// then sums the partials in a fixed order.
void gemm_bf16(bool ta, bool tb, const __bf16 *A, int lda, const __bf16 *B, int ldb, float *C, int ldc, int M, int N,
               int K, int splits, float *part, hipStream_t st);

class Mlp {
public:
    // layers >= 2: in -> hidden -> ... -> hidden -> out (kMlpOut padded)
    Mlp(int hidden, int layers, unsigned long long seed, hipStream_t st);
    ~Mlp();
    Mlp(const Mlp &) = delete;
    Mlp &operator=(const Mlp &) = delete;

    int param_count() const { return nparam_; }
    int hidden() const { return hidden_; }
    int layers() const { return layers_; }
    // host copies of the FP32 master parameters (W0, b0, W1, b1, ...)
    void get_params(float *out) const;
    void set_params(const float *in);
    // rows x kFeat features (device) -> rows x 3 outputs (device)
    void forward(const float *x, int rows, float *y);
    // one Adam step on the mean squared error over the first `outs`
    // columns; returns the loss before the step.  grad_out (host, optional)
    // receives the gradient in the parameter order.
    double train_step(const float *x, const float *y, int rows, float lr, float *grad_out = nullptr);

private:
    hipStream_t st_;
    int hidden_, layers_, nparam_ = 0, step_ = 0;
    std::vector<int> din_, dout_;                 // per layer
    std::vector<long long> woff_, boff_;          // offsets into the parameter vector
    Slab slab_;
    float *p_ = nullptr, *g_ = nullptr, *m_ = nullptr, *v_ = nullptr;   // params, grads, Adam moments
    __bf16 *pb_ = nullptr;                        // bf16 copy of the weights
    // activations for `cap_` rows
    int cap_ = 0;
    float *act_ = nullptr, *dact_ = nullptr, *part_ = nullptr, *lossbuf_ = nullptr;
    __bf16 *actb_ = nullptr, *dactb_ = nullptr, *xb_ = nullptr;
    void ensure_rows(int rows);
    void sync_bf16_weights();
};

}  // namespace lddmm
