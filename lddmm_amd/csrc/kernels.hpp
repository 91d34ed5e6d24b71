// kernels.hpp -- host launchers of the HIP kernels (all asynchronous on the
// stream they are given).  Arrays are device pointers in the SoA layout of
// common.hpp.  `prec` selects kFP64 or kFP32 pair arithmetic.
#pragma once

#include "common.hpp"

namespace lddmm {

// ---- pair sums over points (pair_kernels.hip) ---------------------------
// Culling of FP32 Gaussian tile pairs whose kernel values are all below
// 2^-48 (pair_driver.hpp); on by default, read at every launch (a batch keys
// its captured graphs on it).  Off computes every pair; the results agree to
// a few FP64 ulps.
void set_pair_culling(bool on);
bool pair_culling();
// The culling level: kernel values below 2^-bits are skipped (24..126; 48 by
// default); pair_cull_bits() is the level in force, 0 when culling is off.
constexpr int kDefaultCullBits = 48;
void set_pair_cull_bits(int bits);
int pair_cull_bits();
// Culling by precomputed 64-source group bounds with the surviving groups
// packed densely (on by default); off keeps the tile-by-tile range test.
// pair_groups() is true when groups are on and culling is on.
void set_pair_groups(bool on);
bool pair_groups();
// The expanded form of the FP32 Gaussian pair terms (pair_driver.hpp); on
// by default, read at every launch like the culling switch.  Off gives the
// direct form (results differ by FP32 rounding).
void set_pair_expansion(bool on);
bool pair_expansion();
// Executed-pair counting (diagnostic): with three device counters set, the
// culling operators (FP32 Gaussian K.a, adjoint and 'G' varifold, slots 0, 1,
// 2) add the pairs they compute to theirs; null (the default) counts nothing.  A batch
// launches directly (no graph replay) while a counter is set.
void set_pair_counter(unsigned long long *dev_counter);
unsigned long long *pair_counter();
// Split plans for the point kernels: plan_kv for K.a and plan_bwd for the
// fused backward pass (two targets per lane in FP32), plan_single for the dq
// kernels (one target per lane).
PairSplit plan_kv(int ni, int nj, int prec, int target_blocks);
PairSplit plan_bwd(int ni, int nj, int prec, int order, int target_blocks);
PairSplit plan_single(int ni, int nj, int target_blocks);
// part[(s*D + c)*ni + i] = sum_{j in chunk s} k(qi_i, qj_j) a_j[c]
// Bounds (centroid, radius) of the aligned 64-point groups of an SoA point set,
// one double4 per group: the culling operators (FP32) skip whole groups out of
// range of a block's targets and pack the others densely.  group_count(n) groups.
int group_count(int n);
void launch_group_bounds(int dim, const double *x, int n, double4 *out, hipStream_t st);
// grp (optional, FP32 Gaussian only): group bounds of the sources qj
void launch_kv(int dim, int order, int prec, const double *qi, int ni, const double *qj, const double *a, int nj,
               double width, const PairSplit &sp, double *part, hipStream_t st, const double4 *grp = nullptr);
// out[c*n + i] = scale * sum_s part[(s*comps + c)*n + i]
void launch_sum_splits(int comps, int splits, int n, const double *part, double scale, double *out,
                       hipStream_t st);
// Forward Euler step: v = sum_s part, q_next = q + h v, eparts[b] = partial <a, v>
// over the block's targets, v written to vout when non-null.
void launch_fwd_reduce(int dim, int splits, int n, const double *part, const double *q, const double *a,
                       double h, double *qnext, double *vout, double *eparts, hipStream_t st);
int fwd_reduce_blocks(int n);
// Fused backward pair pass: part[(s*2D + c)*n + i] for c < D holds
// sum_j k_ij amp_j[c] (amp = a - p), for c >= D the dq sum of
// l^T K r with l = amp - a/2, r = a.  Plan: plan_bwd.
void launch_bwd_pairs(int dim, int order, int prec, const double *q, const double *a, const double *p, int n,
                      double width, const PairSplit &sp, double *part, hipStream_t st,
                      const double4 *grp = nullptr);
// g = h * sum_s gpart;  p -= h * sum_s dpart
void launch_bwd_reduce(int dim, int splits, int n, const double *part, double h, double *g, double *p,
                       hipStream_t st);
// Self dq: part = sum_j (l_i.r_j + l_j.r_i) c_ij (q_i - q_j).  Plan: plan_single.
void launch_dq_self(int dim, int order, int prec, const double *q, const double *l, const double *r, int n,
                    double width, const PairSplit &sp, double *part, hipStream_t st);
// Cross dq: part = sum_j (u_i.v_j) c_ij (qi_i - qj_j).  Plan: plan_single.
void launch_dq_cross(int dim, int order, int prec, const double *qi, const double *u, int ni, const double *qj,
                     const double *v, int nj, double width, const PairSplit &sp, double *part, hipStream_t st);
// Dense K[j*ni + i] = k(qi_i, qj_j), FP64
void launch_dense_kernel(int dim, int order, const double *qi, int ni, const double *qj, int nj, double width,
                         double *K, hipStream_t st);

// ---- varifold (varifold_kernels.hip) -------------------------------------
struct VfParams {
    char cen_type, dir_type;
    double cen_width, dir_width;
};
void launch_elm_geometry(int dim, const double *q, int nq, const int *elm, int ne, double *cen, double *dir,
                         double *vol, hipStream_t st);
constexpr int vf_comps(int dim) { return 2 * dim + 2; }
PairSplit plan_vf(int ni, int nj, int target_blocks);
// Pair pass of the deformed elements i against the deformed set (splits
// [0, spd.splits)) and the target set (splits [spd.splits, +spt.splits)):
// partial sums [(s*(2D+2) + c)*ne + i].
void launch_vf_pairs(int dim, int prec, const VfParams &vp, const double *dcen, const double *ddir, const double *dvol,
                     int ne, const double *tcen, const double *tdir, const double *tvol, int nt, const PairSplit &spd,
                     const PairSplit &spt, double *part, hipStream_t st, const double4 *grp_d = nullptr,
                     const double4 *grp_t = nullptr);
// Per element: energy partials per block into eparts, dc and dd; the
// first `sd` splits are deformed-deformed, the rest deformed-target.
void launch_vf_finalize(int dim, const VfParams &vp, int sd, int splits, int ne, const double *part,
                        const double *ddir, const double *dvol, double *dc, double *dd, double *eparts,
                        hipStream_t st);
int vf_finalize_blocks(int ne);
// Value-only pair sum sum_j kc ku a_j of a set with itself (target-target), FP64
void launch_vf_value_pairs(int dim, const VfParams &vp, const double *cen, const double *dir, const double *vol,
                           int n, const PairSplit &sp, double *part, hipStream_t st);
// per-block partials of sum_i a_i * sum_s part[s][i]
void launch_vf_value_finalize(int splits, int n, const double *part, const double *vol, double *eparts,
                              hipStream_t st);
int vf_value_finalize_blocks(int n);
// Vertex gather: out = scale * dq with dq from dc, dd through the adjacency table.
void launch_vf_gather(int dim, const double *q, int nq, const int *ifo, int ifo_rows, const double *dc,
                      const double *dd, int ne, double scale, double *out, hipStream_t st);

// ---- reductions and vector ops (vector_kernels.hip) ------------------------
// out[0] = sum of x[0..n) in a fixed order (two launches, scratch of kDotBlocks)
void launch_sum(const double *x, long long n, double *scratch, double *out, hipStream_t st);
// out[k] = <x_k, y_k> for k < nd pairs; scratch holds nd*kDotBlocks doubles
void launch_dots(int nd, const double *const *xs, const double *const *ys, long long n, double *scratch,
                 double *out, hipStream_t st);
// Energy: out = 0.5*h*sum(fwd parts) + w*(sum(vf parts) + tt[0])
// l2Dist end point: p = -w (q - tl), parts[b] = per-block |q - tl|^2 / 2
int l2_match_blocks(int n);
void launch_l2_match(int dim, const double *q, const double *tl, int n, double w, double *p, double *parts,
                     hipStream_t st);
void launch_energy(const double *fwd_parts, int nf, const double *vf_parts, int nv, const double *tt, double h,
                   double w, double *out, hipStream_t st);
// Vertex permutation of `rows` stacked length-n rows (perm[slot] = vertex):
// gather dst[r*n + s] = src[r*n + perm[s]]; scatter dst[r*n + perm[s]] = src[r*n + s]
void launch_permute(const double *src, double *dst, const int *perm, int n, int rows, bool gather, hipStream_t st);

// ---- measured instruction rates (rates.hip): out[0] v_exp_f32 per second,
// out[1] FP32 FMAs per second, out[2] FP64 FMAs per second, whole chip
void measure_rates(double *out, hipStream_t st);

// ---- fixed scratch of the entry points that allocate nothing (noalloc.hip) ----
// 2 Mi doubles (16 MiB) per GPU, held from module load to process end
constexpr size_t kNoAllocScratchDoubles = size_t(1) << 21;
// device address of the current device's scratch (looked up once per device)
double *noalloc_scratch();

}  // namespace lddmm
