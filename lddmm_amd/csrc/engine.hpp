// engine.hpp -- host side of the LDDMM library: the device objective
// (Cpp/src/objgrd.cpp:26-90), the L-BFGS driver with device-resident state
// (Cpp/src/LBFGS.cpp:25-112) and kernel-time accounting.
#pragma once

#include <algorithm>
#include <array>
#include <cfloat>
#include <mutex>
#include <utility>
#include <vector>

#include "kernels.hpp"

namespace lddmm {

// ---------------------------------------------------------------- profiling
enum ProfId : int {
    kProfFwdPairs = 0,
    kProfFwdReduce,
    kProfBwdPairs,
    kProfBwdReduce,
    kProfVfGeometry,
    kProfVfPairs,
    kProfVfFinalize,
    kProfVfGather,
    kProfVfTargetTarget,
    kProfEnergy,
    kProfLbfgsVector,
    kProfLbfgsDots,
    kProfOperator,
    kProfCount
};
const char *prof_name(int id);

// Records a start/stop event pair around a launch group on its stream when
// enabled; the time of every group is summed at read-out.
class Profiler {
public:
    static Profiler &get();
    void enable(bool on);
    bool enabled() const { return on_; }
    void record(int id, hipStream_t st, bool start);
    void read(double *ms, long long *count);   // synchronises pending events
    void reset();

private:
    struct Pending {
        int id;
        hipEvent_t a, b;
    };
    bool on_ = false;
    std::mutex mu_;
    std::vector<Pending> pend_;
    std::vector<hipEvent_t> pool_;
    std::vector<std::pair<int, hipEvent_t>> open_;
    std::array<double, kProfCount> ms_{};
    std::array<long long, kProfCount> cnt_{};
    hipEvent_t take();
    void collect_locked();
};

struct ProfScope {
    int id;
    hipStream_t st;
    bool on;
    ProfScope(int i, hipStream_t s) : id(i), st(s), on(Profiler::get().enabled())
    {
        if (on) Profiler::get().record(id, st, true);
    }
    ~ProfScope()
    {
        if (on) Profiler::get().record(id, st, false);
    }
};

// ---------------------------------------------------------------- device memory
// One device slab per owner, carved into aligned arrays (the role of
// assignFcnStructMemory / assignOptStructMemory, Cpp/src/assignStructMemory.cpp).
class Slab {
public:
    Slab() = default;
    Slab(const Slab &) = delete;
    Slab &operator=(const Slab &) = delete;
    ~Slab();
    template <typename T>
    void reserve(T **slot, size_t count)
    {
        size_t off = (bytes_ + 255) & ~size_t(255);
        bytes_ = off + count * sizeof(T);
        reqs_.push_back({(void **)slot, off});
    }
    void allocate();
    size_t bytes() const { return bytes_; }

private:
    struct Req {
        void **slot;
        size_t off;
    };
    std::vector<Req> reqs_;
    size_t bytes_ = 0;
    char *base_ = nullptr;
};

// ---------------------------------------------------------------- problem
// Device-resident problem description in the SoA (CUDA) layout.
struct DeviceProblem {
    int dim = 3, n = 0, T = 2, order = -1;
    double width = 1.0, h = 0.0, w = 0.0;
    VfParams vp{'G', 'B', 1.0, 0.0};
    int ne = 0, nt = 0, ifo_rows = 1;
    int prec = kFP64;             // pair arithmetic: kFP64 or kFP32
    const double *q0 = nullptr;   // D x N
    const int *elm = nullptr;     // Ed x V
    const int *ifo = nullptr;     // N x rows
    const double *tc = nullptr, *tu = nullptr, *ta = nullptr;  // Et x D, Et x D, Et
    // landmark targets N x D: when set, the matching term is l2Dist
    // (MATLAB/src/l2Dist.m:1-5) and the element / varifold fields are unused
    const double *tl = nullptr;
    long long varnum() const { return (long long)dim * n * (T - 1); }
};

void validate(const DeviceProblem &p);

// ---------------------------------------------------------------- spatial order
// Order of n points (coord(i, c) = coordinate c of point i) such that every
// aligned block of 64 slots, and of 128 slots, holds the points of one
// subtree of a median bisection: a node of more than 64 points is split at
// the median of its widest coordinate into a first part whose size is a
// multiple of g (g = 128 above 256 points, else 64).  Block offsets are
// multiples of g at every level, so no aligned block straddles two
// subtrees, and the pair kernels' target tiles (64 or 128 consecutive
// slots) are spatially compact.  Returns perm[slot] = point.
template <class Coord>
std::vector<int> spatial_order(int n, int D, Coord coord)
{
    std::vector<int> idx(n > 0 ? n : 0);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::vector<std::pair<int, int>> todo;
    if (n > 64) todo.push_back({0, n});
    while (!todo.empty()) {
        const int lo = todo.back().first, hi = todo.back().second;
        todo.pop_back();
        const int len = hi - lo;
        double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
        for (int k = lo; k < hi; ++k)
            for (int c = 0; c < D; ++c) {
                const double v = coord(idx[k], c);
                mn[c] = std::min(mn[c], v);
                mx[c] = std::max(mx[c], v);
            }
        int ax = 0;
        for (int c = 1; c < D; ++c)
            if (mx[c] - mn[c] > mx[ax] - mn[ax]) ax = c;
        const int g = len > 256 ? 128 : 64;
        const int left = g * ((len + 2 * g - 1) / (2 * g));
        std::nth_element(idx.begin() + lo, idx.begin() + lo + left, idx.begin() + hi, [&](int a, int b) {
            const double va = coord(a, ax), vb = coord(b, ax);
            return va < vb || (va == vb && a < b);
        });
        if (left > 64) todo.push_back({lo, lo + left});
        if (hi - lo - left > 64) todo.push_back({lo + left, hi});
    }
    return idx;
}

// E(alpha) and dE/dalpha on the device.  The flow stack, adjoint and all
// partial sums live in one slab; the constant target-target varifold term
// is computed once at construction.  Internally the problem is held in a
// spatial order (vertices, deformed elements and target elements each by
// spatial_order of their initial positions, the element and adjacency
// tables renumbered to match); eval() and trajectory_to() take and give
// the caller's vertex order.
class Objective {
public:
    Objective(const DeviceProblem &p, hipStream_t st, int target_blocks = 0);
    ~Objective();
    // f_dev[0] = E(alp); grd = dE/dalp.  alp and grd are (T-1) SoA slices
    // in the caller's vertex order.
    void eval(const double *alp, double *grd, double *f_dev);
    // the same with alp and grd in the internal vertex order
    void eval_internal(const double *alp, double *grd, double *f_dev);
    // stacks of `slices` D x N SoA slices between the two vertex orders
    void to_internal(const double *user, double *internal, int slices);
    void to_user(const double *internal, double *user, int slices);
    // T SoA slices of the last evaluation's flow, in the caller's order
    void trajectory_to(double *dst);
    double target_target() const { return tt_host_; }
    const DeviceProblem &problem() const { return p_; }
    // switch the pair arithmetic (kFP64 / kFP32) of later evaluations
    void set_precision(int prec);
    int precision() const { return p_.prec; }
    // pair interactions of one evaluation: forward, fused backward, varifold
    double pairs_per_eval() const;
    double pairs_target_target() const { return (double)p_.nt * p_.nt; }

private:
    DeviceProblem p_;
    hipStream_t st_;
    PairSplit sp_kv_, sp_bwd_, sp_vfd_, sp_vft_;
    Slab slab_;
    double *nde_ = nullptr, *p_adj_ = nullptr, *part_ = nullptr;
    double *ecen_ = nullptr, *edir_ = nullptr, *evol_ = nullptr, *dc_ = nullptr, *dd_ = nullptr;
    double *ep_fwd_ = nullptr, *ep_vf_ = nullptr, *tt_ = nullptr, *scratch_ = nullptr;
    // the problem in the internal order, and the vertex permutation
    double *q0_ = nullptr, *tc_ = nullptr, *tu_ = nullptr, *ta_ = nullptr, *tl_ = nullptr;
    int *elm_ = nullptr, *ifo_ = nullptr, *vperm_ = nullptr;
    double *alp_i_ = nullptr, *grd_i_ = nullptr;
    // FP32 culling: group bounds of q_t (T slices), of the deformed element
    // centres and of the target centres
    double4 *grp_q_ = nullptr, *grp_e_ = nullptr, *grp_t_ = nullptr;
    int nb_fwd_ = 0, nb_vf_ = 0, nb_l2_ = 0, target_blocks_ = 0;
    double tt_host_ = 0.0;
    void reorder();
    void compute_target_target(int target_blocks);
};

// ---------------------------------------------------------------- L-BFGS
struct LbfgsOptions {
    int itr_max = 200, mem = 5;
    double tol = 0.0, wolfe1 = 0.0, wolfe2 = 0.5;
    bool verbose = false;
    double rel_tol = 0.0;   // > 0: also stop when (f_prev - f) / |f| < rel_tol
    // FP32 objectives: on a stall (line-search failure, or relative decrease
    // below refine_rel_tol) switch the objective to FP64, re-evaluate and go on
    bool fp64_refine = true;
    double refine_rel_tol = 1e-7;
    bool has_target = false;   // stop once f <= target
    double target = 0.0;
};

struct LbfgsResult {
    double f = 0.0, gnorm = 0.0, seconds = 0.0;
    int iters = 0, evals = 0, status = 0, fp32_iters = 0, reached_target = 0;
    std::vector<double> trace_f, trace_g, trace_step;
    std::vector<int> trace_evals;
};

// x (device, nv) holds the start on entry and the last accepted iterate on
// return.
LbfgsResult run_lbfgs(Objective &obj, double *x, const LbfgsOptions &o, hipStream_t st);

}  // namespace lddmm
