// batch.hpp -- many independent registrations of uneven sizes evaluated
// together: one launch per Euler step (and per varifold pass) covers every
// active problem through a table of (problem, target tile, source chunk)
// work items, so small and large problems share the chip.  Each problem
// keeps its own L-BFGS state (Cpp/src/LBFGS.cpp:25-112 per problem); the
// line searches advance in lock-step rounds, one batched evaluation per
// round.
#pragma once

#include <memory>
#include <vector>

#include "engine.hpp"

namespace lddmm {

// Per-problem descriptor on the device (all offsets in elements).
struct BatchDesc {
    long long vo;     // first vertex of the problem in the concatenated vertex arrays
    long long eo;     // first deformed element
    long long to;     // first target element
    long long po;     // point-kernel partial sums (doubles)
    long long vpo;    // varifold partial sums (doubles)
    long long tpo;    // target-target partial sums (doubles)
    long long gvo;    // group bounds of q_t: T slices of ceil(n / 64) (double4)
    long long geo;    // group bounds of the deformed element centres
    long long gto;    // group bounds of the target centres
    int n, ne, nt;
    int chunk;        // point-kernel source chunk
    int splits;       // ceil(n / chunk)
    int vchunk;       // varifold source chunk
    int vsd, vst;     // varifold splits over deformed / target sources
    int tsplits;      // target-target splits
    int pad;
    double iw, iw2;   // 1/sigma, 1/sigma^2 of the deformation kernel
};

// Host description of a batch: per-problem sizes and widths, shared
// kernel types, time discretisation and matching weight.
struct BatchSpec {
    int dim = 3, T = 2, order = -1, prec = kFP64;
    double h = 0.0, w = 0.0;
    VfParams vp{'G', 'B', 1.0, 0.0};
    std::vector<int> n, ne, nt;
    std::vector<double> width;
    int chunk = 0;    // point-kernel source chunk; 0 = autotune
    int P() const { return (int)n.size(); }
};

struct BatchLbfgsOptions {
    LbfgsOptions base;   // itr_max, mem, tol, wolfe1, wolfe2, rel_tol, refine, target (shared)
    std::vector<double> targets;   // per-problem target energies (optional)
    int prec = -1;                 // pair arithmetic of the run; -1 = the batch's
    int progress = 0;              // > 0: a progress line on stderr every `progress` rounds
};

struct BatchProblemResult {
    double f = 0.0, gnorm = 0.0;
    int iters = 0, evals = 0, status = 0, fp32_iters = 0, reached_target = 0;
};

struct BatchRunResult {
    std::vector<BatchProblemResult> prob;
    int rounds = 0;
    double seconds = 0.0;
};

class Batch {
public:
    // Host inputs in the host layout, concatenated problem after problem:
    // q0 D x N_p blocks, elm V x E_p blocks (0-based, local to the problem),
    // target centres / directions D x Et_p blocks and volumes Et_p.
    Batch(const BatchSpec &spec, const double *q0, const int *elm, const double *tc, const double *ta,
          const double *tu, hipStream_t st);
    ~Batch();
    Batch(const Batch &) = delete;
    Batch &operator=(const Batch &) = delete;

    int P() const { return spec_.P(); }
    long long varnum_total() const { return vtot_; }
    long long var_offset(int p) const { return (long long)spec_.dim * (spec_.T - 1) * vo_[p]; }
    long long varnum(int p) const { return (long long)spec_.dim * (spec_.T - 1) * spec_.n[p]; }
    const BatchSpec &spec() const { return spec_; }
    int chunk() const { return chunk_; }
    // targets per lane of the FP32 K.alpha and adjoint launches (autotuned)
    int kv_lanes() const { return rkv_; }
    int bwd_lanes() const { return rbwd_; }
    // fixes the source chunk (rounded up to a multiple of 64; the partial
    // buffer holds chunks down to min(256, the chunk at creation)) and the
    // FP32 targets per lane (K.alpha 1, 2 or 4; adjoint 1 or 2); the stream
    // must be idle
    void set_tuning(int chunk, int kv_lanes, int bwd_lanes);
    double pairs_per_eval(int p) const;

    // f_dev[p], grd (device, concatenated per-problem [T-1][D][N_p] blocks)
    // for the problems with prec_of[p] >= 0 (their pair arithmetic); grd =
    // nullptr evaluates the energies only (no adjoint sweep)
    void eval(const double *alp, double *grd, double *f_dev, const std::vector<int> &prec_of);
    // host layout (caller's vertex order) <-> device layout (internal
    // spatial order of every problem, as in Objective)
    std::vector<double> to_device_layout(const double *host_alp) const;
    void to_host_layout(const std::vector<double> &dev, double *host_alp) const;
    // the same for one vector per vertex (D x N_p blocks, no time axis)
    std::vector<double> vec_to_device_layout(const double *host_vec) const;
    // x = the momentum stack whose every time slice is s[p] u_p: u the
    // device-layout D x N_p blocks, s one multiple per problem (both device)
    void const_momentum(const double *u, const double *s, double *x);
    const double *target_terms() const { return tt_; }

    BatchRunResult run(double *x, const BatchLbfgsOptions &o);
    // evaluations replay captured hipGraphs (default on)
    void set_graphs(bool on);
    bool graphs() const { return use_graphs_; }

private:
    BatchSpec spec_;
    hipStream_t st_;
    int chunk_ = 1024, vchunk_ = 1024, rows_ = 1;
    int rkv_ = 2, rbwd_ = 2;   // FP32 targets per lane: K.alpha (1, 2, 4), adjoint (1, 2; 4 = matrix cores)
    long long ntot_ = 0, etot_ = 0, ttot_ = 0, vtot_ = 0;
    size_t part_cap_ = 0;   // doubles in part_
    std::vector<long long> vo_, eo_, to_;
    std::vector<int> vperm_;   // per problem: internal slot -> caller's vertex (local indices)
    std::vector<BatchDesc> desc_h_;
    Slab slab_;
    BatchDesc *desc_ = nullptr;
    double *q0_ = nullptr, *tc_ = nullptr, *tu_ = nullptr, *ta_ = nullptr;
    int *elm_ = nullptr, *ifo_ = nullptr;
    double *nde_ = nullptr, *padj_ = nullptr, *part_ = nullptr, *vpart_ = nullptr;
    double *ecen_ = nullptr, *edir_ = nullptr, *evol_ = nullptr, *dc_ = nullptr, *dd_ = nullptr;
    double *efwd_ = nullptr, *evf_ = nullptr, *tt_ = nullptr;
    double4 *grpv_ = nullptr, *grpe_ = nullptr, *grpt_ = nullptr;   // group bounds (FP32 culling)
    float2 *grpm_ = nullptr;   // group maxima of alpha_t and alpha_t - p (matrix-core adjoint)
    // work tables for the current active set
    struct Tables;
    std::unique_ptr<Tables> tab_;
    std::vector<int> tab_key_;
    struct GraphEntry {
        const double *alp;
        double *grd, *f;
        int cull, expand;   // pair_cull_bits(), pair_expansion() at capture
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs_;
    bool use_graphs_ = true;
    void clear_graphs();
    void launch_eval(const double *alp, double *grd, double *f_dev);
    void build_descs();
    void upload_descs();
    void build_tables(const std::vector<int> &prec_of);
    void compute_target_terms();
    void autotune();
    // the matrix-core adjoint (four targets per lane's tiles of 256) applies
    bool mx_ok() const { return spec_.dim == 3 && spec_.order == -1; }
};

}  // namespace lddmm
