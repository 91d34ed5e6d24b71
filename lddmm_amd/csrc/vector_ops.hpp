// vector_ops.hpp -- launchers of the L-BFGS vector work (vector_kernels.hip).
#pragma once

#include "common.hpp"

namespace lddmm {

constexpr int kMaxDots = 4;

// Products of ring slots with vectors: for combination c < ncombo (<= 3) and
// slot j < h, out[c*h + j] = <rings[c] + j*nv, vecs[c]>; scratch holds
// ring_dot_scratch(ncombo*h) doubles.  Fixed reduction shapes: bitwise
// repeatable.
size_t ring_dot_scratch(int nd);
void launch_ring_dots(int ncombo, const double *const *rings, const double *const *vecs, int h, long long nv,
                      double *scratch, double *out, hipStream_t st);
// Coefficients of the compact two-loop recursion (one block): from
// ab = [s_j.g (h) | y_j.g (h)] by slot and the m x m tables SY, YY of the h
// stored pairs (newest in slot `newest`); work: 3m doubles; coef: 1 + 2m
// (g, then the y slots, then the s slots).  h = 0 gives d = -g.
void launch_lbfgs_coef(const double *ab, const double *SY, const double *YY, int m, int h, int newest, double *work,
                       double *coef, hipStream_t st);
// d = coef[0] g + sum_{j<h} coef[1+j] Y_j + coef[1+m+j] S_j; out[0] = g.d,
// out[1] = d.d; scratch: 2*kDotBlocks doubles
void launch_lbfgs_dir(double *d, const double *g, const double *S, const double *Y, long long nv, const double *coef,
                      int h, int m, double *scratch, double *out, hipStream_t st);
// the newest pair's table entries from dots = [y_j.s_n | s_j.y_n | y_j.y_n], j < h
void launch_lbfgs_store(const double *dots, int h, int newest, int m, double *SY, double *YY, hipStream_t st);
// xt = x + rel d
void vec_try(double *xt, const double *x, double rel, const double *d, long long n, hipStream_t st);
// s = xt - x, y = gt - g, x = xt, g = gt
void vec_accept(double *x, double *g, const double *xt, const double *gt, double *s, double *y, long long n,
                hipStream_t st);

}  // namespace lddmm
