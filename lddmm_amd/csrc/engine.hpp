// engine.hpp -- host side of the LDDMM library: the device objective
// (Cpp/src/objgrd.cpp:26-90), the L-BFGS driver with device-resident state
// (Cpp/src/LBFGS.cpp:25-112) and kernel-time accounting.
#pragma once

#include <array>
#include <mutex>
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

// E(alpha) and dE/dalpha on the device.  The flow stack, adjoint and all
// partial sums live in one slab; the constant target-target varifold term
// is computed once at construction.
class Objective {
public:
    Objective(const DeviceProblem &p, hipStream_t st, int target_blocks = 0);
    ~Objective();
    // f_dev[0] = E(alp); grd = dE/dalp.  alp and grd are (T-1) SoA slices.
    void eval(const double *alp, double *grd, double *f_dev);
    const double *trajectory() const { return nde_; }   // T SoA slices of the last eval
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
    int nb_fwd_ = 0, nb_vf_ = 0, nb_l2_ = 0, target_blocks_ = 0;
    double tt_host_ = 0.0;
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
