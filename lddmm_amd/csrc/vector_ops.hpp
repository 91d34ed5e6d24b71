// vector_ops.hpp -- launchers of the L-BFGS vector updates (vector_kernels.hip).
#pragma once

#include "common.hpp"

namespace lddmm {

constexpr int kMaxDots = 4;

// d = -g
void vec_copy_neg(double *d, const double *g, long long n, hipStream_t st);
// rec[0] = sd/sy;  d -= rec[0] y
void vec_loop1(double *d, const double *y, const double *sd, const double *sy, double *rec, long long n,
               hipStream_t st);
// d *= sy/yy
void vec_scale_h0(double *d, const double *sy, const double *yy, long long n, hipStream_t st);
// d += (rec[0] - yd/sy) s, negated when `negate`
void vec_loop2(double *d, const double *s, const double *yd, const double *sy, const double *rec, bool negate,
               long long n, hipStream_t st);
void vec_scale(double *d, double a, long long n, hipStream_t st);
// xt = x + rel d
void vec_try(double *xt, const double *x, double rel, const double *d, long long n, hipStream_t st);
// s = xt - x, y = gt - g, x = xt, g = gt
void vec_accept(double *x, double *g, const double *xt, const double *gt, double *s, double *y, long long n,
                hipStream_t st);

}  // namespace lddmm
