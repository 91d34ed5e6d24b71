// engine.cpp -- device objective, device-resident L-BFGS driver, profiler.
#include "engine.hpp"

#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>

#include "vector_ops.hpp"

namespace lddmm {

// ======================================================================
// Profiler
// ======================================================================
const char *prof_name(int id)
{
    static const char *names[kProfCount] = {
        "fwd_pairs",  "fwd_reduce",  "bwd_pairs", "bwd_reduce",   "vf_geometry", "vf_pairs",  "vf_finalize",
        "vf_gather",  "vf_target_target", "energy", "lbfgs_vector", "lbfgs_dots", "operator",
    };
    return (id >= 0 && id < kProfCount) ? names[id] : "?";
}

Profiler &Profiler::get()
{
    static Profiler p;
    return p;
}

void Profiler::enable(bool on)
{
    std::lock_guard<std::mutex> g(mu_);
    on_ = on;
}

hipEvent_t Profiler::take()
{
    if (!pool_.empty()) {
        hipEvent_t e = pool_.back();
        pool_.pop_back();
        return e;
    }
    hipEvent_t e;
    LDDMM_HIP(hipEventCreate(&e));
    return e;
}

void Profiler::record(int id, hipStream_t st, bool start)
{
    std::lock_guard<std::mutex> g(mu_);
    hipEvent_t e = take();
    LDDMM_HIP(hipEventRecord(e, st));
    if (start) {
        open_.push_back({id, e});
        return;
    }
    for (size_t k = open_.size(); k-- > 0;) {
        if (open_[k].first == id) {
            pend_.push_back({id, open_[k].second, e});
            open_.erase(open_.begin() + (long)k);
            break;
        }
    }
    if (pend_.size() > 4096) collect_locked();
}

void Profiler::collect_locked()
{
    for (auto &p : pend_) {
        LDDMM_HIP(hipEventSynchronize(p.b));
        float ms = 0.f;
        LDDMM_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        ms_[p.id] += ms;
        cnt_[p.id] += 1;
        pool_.push_back(p.a);
        pool_.push_back(p.b);
    }
    pend_.clear();
}

void Profiler::read(double *ms, long long *count)
{
    std::lock_guard<std::mutex> g(mu_);
    collect_locked();
    for (int k = 0; k < kProfCount; ++k) {
        if (ms) ms[k] = ms_[k];
        if (count) count[k] = cnt_[k];
    }
}

void Profiler::reset()
{
    std::lock_guard<std::mutex> g(mu_);
    collect_locked();
    ms_.fill(0.0);
    cnt_.fill(0);
}

// ======================================================================
// Slab
// ======================================================================
Slab::~Slab()
{
    if (base_) (void)hipFree(base_);
}

void Slab::allocate()
{
    if (bytes_ == 0) return;
    if (hipMalloc((void **)&base_, bytes_) != hipSuccess) {
        base_ = nullptr;
        throw Error(kErrAlloc, "device allocation of " + std::to_string(bytes_) + " bytes failed");
    }
    for (auto &r : reqs_) *r.slot = base_ + r.off;
}

// ======================================================================
// Objective
// ======================================================================
void validate(const DeviceProblem &p)
{
    if (p.dim != 2 && p.dim != 3) throw Error(kErrArg, "only 2D and 3D are supported");
    if (p.n <= 0) throw Error(kErrArg, "ndeNum must be positive");
    if (p.T < 2) throw Error(kErrArg, "timeNum must be at least 2");
    if (p.order < -1 || p.order > 4) throw Error(kErrArg, "knlOrder must be -1 or 0..4");
    if (!(p.width > 0.0)) throw Error(kErrArg, "knlWidth must be positive");
    if ((long long)p.dim * p.n * (p.T - 1) >= (1LL << 40)) throw Error(kErrArg, "problem too large");
    if (p.tl) {
        if (p.ne != 0 || p.nt != 0) throw Error(kErrArg, "l2 landmark matching takes no elements or target varifold");
        return;
    }
    if (p.ne <= 0 || p.nt < 0) throw Error(kErrArg, "element counts must be positive");
    if (p.ifo_rows < 1) throw Error(kErrArg, "dfmElmIfoNum must be at least 1");
    if (p.vp.cen_type != 'G' && p.vp.cen_type != 'C') throw Error(kErrArg, "cenKnlType must be 'G' or 'C'");
    if (p.vp.dir_type != 'B' && p.vp.dir_type != 'L' && p.vp.dir_type != 'O' && p.vp.dir_type != 'U')
        throw Error(kErrArg, "dirKnlType must be 'B', 'L', 'O' or 'U'");
    if (!(p.vp.cen_width > 0.0)) throw Error(kErrArg, "cenKnlWidth must be positive");
    if ((p.vp.dir_type == 'O' || p.vp.dir_type == 'U') && !(p.vp.dir_width > 0.0))
        throw Error(kErrArg, "dirKnlWidth must be positive for 'O' and 'U'");
    if ((long long)p.dim * p.n * (p.T - 1) >= (1LL << 40)) throw Error(kErrArg, "problem too large");
}

static int default_target_blocks()
{
    // ~16 workgroups of 256 threads per CU on the 256-CU chip: at 4-8
    // resident workgroups per CU, two or more rounds of equal blocks leave a
    // smaller idle tail than one round (measured on the 10k-vertex
    // sphere->cube evaluation: FP64 -8%, FP32 -4% against 2048; more splits
    // cost more in the split reductions)
    return 4096;
}

Objective::Objective(const DeviceProblem &p, hipStream_t st, int target_blocks) : p_(p), st_(st)
{
    validate(p_);
    const int tb = target_blocks > 0 ? target_blocks : default_target_blocks();
    target_blocks_ = tb;
    const int D = p_.dim;
    sp_kv_ = plan_kv(p_.n, p_.n, p_.prec, tb);
    sp_bwd_ = plan_bwd(p_.n, p_.n, p_.prec, p_.order, tb);
    const int other = p_.prec == kFP32 ? kFP64 : kFP32;
    const PairSplit kv_other = plan_kv(p_.n, p_.n, other, tb), bwd_other = plan_bwd(p_.n, p_.n, other, p_.order, tb);
    // the varifold sums (one launch each for the deformed and the target
    // sources) take twice the point kernels' workgroup count: their finer
    // source chunks also cull more of the far pairs in FP32
    const double fd = p_.ne + p_.nt > 0 ? (double)p_.ne / (double)(p_.ne + p_.nt) : 1.0;
    int tbd = (int)std::max(1.0, 2 * tb * fd), tbt = std::max(1, 2 * tb - tbd);
    sp_vfd_ = plan_vf(p_.ne, p_.ne, tbd);
    sp_vft_ = plan_vf(p_.ne, std::max(p_.nt, 1), tbt);
    PairSplit sp_tt = plan_vf(std::max(p_.nt, 1), std::max(p_.nt, 1), tb);

    size_t part = (size_t)std::max(sp_kv_.splits, kv_other.splits) * D * p_.n;
    part = std::max(part, (size_t)std::max(sp_bwd_.splits, bwd_other.splits) * 2 * D * p_.n);
    part = std::max(part, (size_t)(sp_vfd_.splits + sp_vft_.splits) * vf_comps(D) * p_.ne);
    part = std::max(part, (size_t)sp_tt.splits * std::max(p_.nt, 1));
    nb_fwd_ = fwd_reduce_blocks(p_.n);
    nb_vf_ = vf_finalize_blocks(p_.ne);
    nb_l2_ = p_.tl ? l2_match_blocks(p_.n) : 0;
    const int nb_tt = vf_value_finalize_blocks(std::max(p_.nt, 1));

    slab_.reserve(&nde_, (size_t)D * p_.n * p_.T);
    slab_.reserve(&p_adj_, (size_t)D * p_.n);
    slab_.reserve(&part_, part);
    slab_.reserve(&ecen_, (size_t)D * p_.ne);
    slab_.reserve(&edir_, (size_t)D * p_.ne);
    slab_.reserve(&evol_, (size_t)p_.ne);
    slab_.reserve(&dc_, (size_t)D * p_.ne);
    slab_.reserve(&dd_, (size_t)D * p_.ne);
    slab_.reserve(&ep_fwd_, (size_t)(p_.T - 1) * nb_fwd_);
    slab_.reserve(&ep_vf_, (size_t)std::max(std::max(nb_vf_, nb_tt), nb_l2_));
    slab_.reserve(&tt_, 1);
    slab_.reserve(&scratch_, (size_t)kDotBlocks);
    // the problem in the internal order
    slab_.reserve(&q0_, (size_t)D * p_.n);
    slab_.reserve(&vperm_, (size_t)p_.n);
    slab_.reserve(&alp_i_, (size_t)p_.varnum());
    slab_.reserve(&grd_i_, (size_t)p_.varnum());
    slab_.reserve(&grp_q_, (size_t)p_.T * group_count(p_.n));
    slab_.reserve(&grp_e_, (size_t)std::max(1, group_count(p_.ne)));
    slab_.reserve(&grp_t_, (size_t)std::max(1, group_count(p_.nt)));
    if (p_.tl) {
        slab_.reserve(&tl_, (size_t)D * p_.n);
    } else {
        slab_.reserve(&elm_, (size_t)D * p_.ne);
        slab_.reserve(&ifo_, (size_t)p_.ifo_rows * p_.n);
        if (p_.nt > 0) {
            slab_.reserve(&tc_, (size_t)D * p_.nt);
            slab_.reserve(&tu_, (size_t)D * p_.nt);
            slab_.reserve(&ta_, (size_t)p_.nt);
        }
    }
    slab_.allocate();
    reorder();
    compute_target_target(tb);
}

Objective::~Objective() = default;

// Reads the problem back, orders vertices, deformed elements and target
// elements spatially (spatial_order of the initial vertex positions, the
// initial element centres and the target centres), renumbers the element
// and adjacency tables, checks every index on the way, and points p_ at
// the reordered copies.  Every sum of the objective is over the same
// terms; only their order changes.
void Objective::reorder()
{
    const int D = p_.dim, n = p_.n, ne = p_.ne, nt = p_.nt, rows = p_.ifo_rows;
    auto down = [&](auto *dst, const auto *src, size_t count) {
        if (count) LDDMM_HIP(hipMemcpyAsync(dst, src, count * sizeof(*src), hipMemcpyDeviceToHost, st_));
    };
    auto up = [&](auto *dst, const auto *src, size_t count) {
        if (count) LDDMM_HIP(hipMemcpyAsync(dst, src, count * sizeof(*src), hipMemcpyHostToDevice, st_));
    };
    std::vector<double> q((size_t)D * n), tl, tc, tu, ta;
    std::vector<int> elm, ifo;
    down(q.data(), p_.q0, q.size());
    if (p_.tl) {
        tl.resize((size_t)D * n);
        down(tl.data(), p_.tl, tl.size());
    } else {
        elm.resize((size_t)D * ne);
        ifo.resize((size_t)rows * n);
        down(elm.data(), p_.elm, elm.size());
        down(ifo.data(), p_.ifo, ifo.size());
        if (nt > 0) {
            tc.resize((size_t)D * nt);
            tu.resize((size_t)D * nt);
            ta.resize(nt);
            down(tc.data(), p_.tc, tc.size());
            down(tu.data(), p_.tu, tu.size());
            down(ta.data(), p_.ta, ta.size());
        }
    }
    LDDMM_HIP(hipStreamSynchronize(st_));

    const std::vector<int> vp = spatial_order(n, D, [&](int i, int c) { return q[(size_t)c * n + i]; });
    std::vector<int> vinv(n);
    for (int s = 0; s < n; ++s) vinv[vp[s]] = s;
    std::vector<double> qs((size_t)D * n);
    for (int c = 0; c < D; ++c)
        for (int s = 0; s < n; ++s) qs[(size_t)c * n + s] = q[(size_t)c * n + vp[s]];
    up(q0_, qs.data(), qs.size());
    up(vperm_, vp.data(), vp.size());
    p_.q0 = q0_;

    std::vector<double> tls, tcs, tus, tas;
    std::vector<int> elms, ifos;
    if (p_.tl) {
        tls.resize(tl.size());
        for (int c = 0; c < D; ++c)
            for (int s = 0; s < n; ++s) tls[(size_t)c * n + s] = tl[(size_t)c * n + vp[s]];
        up(tl_, tls.data(), tls.size());
        p_.tl = tl_;
    } else {
        for (int v : elm)
            if (v < 0 || v >= n) throw Error(kErrArg, "dfmElmVtxMat holds a vertex index outside [0, ndeNum)");
        std::vector<double> cen((size_t)D * ne, 0.0);
        for (int e = 0; e < ne; ++e)
            for (int k = 0; k < D; ++k) {
                const int v = elm[(size_t)k * ne + e];
                for (int c = 0; c < D; ++c) cen[(size_t)c * ne + e] += q[(size_t)c * n + v] / D;
            }
        const std::vector<int> ep = spatial_order(ne, D, [&](int i, int c) { return cen[(size_t)c * ne + i]; });
        std::vector<int> einv(ne);
        for (int s = 0; s < ne; ++s) einv[ep[s]] = s;
        elms.resize(elm.size());
        for (int k = 0; k < D; ++k)
            for (int s = 0; s < ne; ++s) elms[(size_t)k * ne + s] = vinv[elm[(size_t)k * ne + ep[s]]];
        // adjacency rows: count, then per element (element, sign) in 2D or
        // (element, vertex, vertex) in 3D (Cpp/src/getElmIfo.m)
        ifos.assign(ifo.size(), 0);
        for (int s = 0; s < n; ++s) {
            const int u = vp[s], cnt = ifo[u];
            if (cnt < 0 || 1 + (long long)cnt * D > rows)
                throw Error(kErrArg, "dfmElmIfoMat count exceeds its row count");
            ifos[s] = cnt;
            for (int a = 0; a < cnt; ++a) {
                const size_t b = 1 + (size_t)a * D;
                const int e = ifo[b * n + u];
                if (e < 0 || e >= ne) throw Error(kErrArg, "dfmElmIfoMat holds an element index outside [0, dfmElmNum)");
                ifos[b * n + s] = einv[e];
                if (D == 2) {
                    ifos[(b + 1) * n + s] = ifo[(b + 1) * n + u];
                } else {
                    for (int z = 1; z <= 2; ++z) {
                        const int v = ifo[(b + z) * n + u];
                        if (v < 0 || v >= n) throw Error(kErrArg, "dfmElmIfoMat holds a vertex index outside [0, ndeNum)");
                        ifos[(b + z) * n + s] = vinv[v];
                    }
                }
            }
        }
        up(elm_, elms.data(), elms.size());
        up(ifo_, ifos.data(), ifos.size());
        p_.elm = elm_;
        p_.ifo = ifo_;
        if (nt > 0) {
            const std::vector<int> tp = spatial_order(nt, D, [&](int i, int c) { return tc[(size_t)c * nt + i]; });
            tcs.resize(tc.size());
            tus.resize(tu.size());
            tas.resize(nt);
            for (int s = 0; s < nt; ++s) {
                for (int c = 0; c < D; ++c) {
                    tcs[(size_t)c * nt + s] = tc[(size_t)c * nt + tp[s]];
                    tus[(size_t)c * nt + s] = tu[(size_t)c * nt + tp[s]];
                }
                tas[s] = ta[tp[s]];
            }
            up(tc_, tcs.data(), tcs.size());
            up(tu_, tus.data(), tus.size());
            up(ta_, tas.data(), tas.size());
            p_.tc = tc_;
            p_.tu = tu_;
            p_.ta = ta_;
        }
    }
    LDDMM_HIP(hipStreamSynchronize(st_));   // host copies go out of scope
}

void Objective::to_internal(const double *user, double *internal, int slices)
{
    launch_permute(user, internal, vperm_, p_.n, slices * p_.dim, true, st_);
}

void Objective::to_user(const double *internal, double *user, int slices)
{
    launch_permute(internal, user, vperm_, p_.n, slices * p_.dim, false, st_);
}

void Objective::trajectory_to(double *dst) { to_user(nde_, dst, p_.T); }

void Objective::eval(const double *alp, double *grd, double *f_dev)
{
    to_internal(alp, alp_i_, p_.T - 1);
    eval_internal(alp_i_, grd_i_, f_dev);
    to_user(grd_i_, grd, p_.T - 1);
}

void Objective::set_precision(int prec)
{
    if (prec != kFP64 && prec != kFP32) throw Error(kErrArg, "unknown precision policy");
    p_.prec = prec;
    // the partial-sum buffer was sized for the plans of both policies
    sp_kv_ = plan_kv(p_.n, p_.n, prec, target_blocks_);
    sp_bwd_ = plan_bwd(p_.n, p_.n, prec, p_.order, target_blocks_);
}

void Objective::compute_target_target(int target_blocks)
{
    if (p_.nt <= 0) {
        LDDMM_HIP(hipMemsetAsync(tt_, 0, sizeof(double), st_));
        LDDMM_HIP(hipStreamSynchronize(st_));
        tt_host_ = 0.0;
        return;
    }
    launch_group_bounds(p_.dim, p_.tc, p_.nt, grp_t_, st_);   // fixed targets: once
    PairSplit sp = plan_vf(p_.nt, p_.nt, target_blocks);
    {
        ProfScope ps(kProfVfTargetTarget, st_);
        launch_vf_value_pairs(p_.dim, p_.vp, p_.tc, p_.tu, p_.ta, p_.nt, sp, part_, st_);
        launch_vf_value_finalize(sp.splits, p_.nt, part_, p_.ta, ep_vf_, st_);
        launch_sum(ep_vf_, vf_value_finalize_blocks(p_.nt), scratch_, tt_, st_);
    }
    LDDMM_HIP(hipMemcpyAsync(&tt_host_, tt_, sizeof(double), hipMemcpyDeviceToHost, st_));
    LDDMM_HIP(hipStreamSynchronize(st_));
}

double Objective::pairs_per_eval() const
{
    const double n = p_.n, T1 = p_.T - 1;
    return 2.0 * T1 * n * n + (double)p_.ne * p_.ne + (double)p_.ne * p_.nt;   // l2 matching: ne = nt = 0
}

void Objective::eval_internal(const double *alp, double *grd, double *f_dev)
{
    const int D = p_.dim, n = p_.n, T = p_.T;
    const size_t dn = (size_t)D * n;

    // forward flow q_{t+1} = q_t + h K(q_t) alpha_t, energy partials <alpha_t, v_t>
    // (Cpp/src/objgrd.cpp:41-59)
    LDDMM_HIP(hipMemcpyAsync(nde_, p_.q0, dn * sizeof(double), hipMemcpyDeviceToDevice, st_));
    // FP32 Gaussian: group bounds of each q_t for the culling of the pair
    // kernels (forward here, the adjoint below)
    const bool groups = p_.prec == kFP32 && p_.order == -1 && pair_groups();
    const int ng = group_count(n);
    for (int t = 0; t < T - 1; ++t) {
        const double *qt = nde_ + (size_t)t * dn;
        const double *at = alp + (size_t)t * dn;
        {
            ProfScope ps(kProfFwdPairs, st_);
            if (groups) launch_group_bounds(D, qt, n, grp_q_ + (size_t)t * ng, st_);
            launch_kv(D, p_.order, p_.prec, qt, n, qt, at, n, p_.width, sp_kv_, part_, st_,
                      groups ? grp_q_ + (size_t)t * ng : nullptr);
        }
        {
            ProfScope ps(kProfFwdReduce, st_);
            launch_fwd_reduce(D, sp_kv_.splits, n, part_, qt, at, p_.h, nde_ + (size_t)(t + 1) * dn, nullptr,
                              ep_fwd_ + (size_t)t * nb_fwd_, st_);
        }
    }

    const double *qT = nde_ + (size_t)(T - 1) * dn;
    if (p_.tl) {
        // l2Dist at q_{T-1} (MATLAB/src/objgrd.m:29-34 with disHdl = l2Dist)
        {
            ProfScope ps(kProfVfFinalize, st_);
            launch_l2_match(D, qT, p_.tl, n, p_.w, p_adj_, ep_vf_, st_);
        }
        {
            ProfScope ps(kProfEnergy, st_);
            launch_energy(ep_fwd_, (T - 1) * nb_fwd_, ep_vf_, nb_l2_, tt_, p_.h, p_.w, f_dev, st_);
        }
    } else {
        // varifold at q_{T-1}; adjoint start p = -w grad V (Cpp/src/objgrd.cpp:61-67)
        {
            ProfScope ps(kProfVfGeometry, st_);
            launch_elm_geometry(D, qT, n, p_.elm, p_.ne, ecen_, edir_, evol_, st_);
        }
        {
            ProfScope ps(kProfVfPairs, st_);
            const bool vgroups = p_.prec == kFP32 && p_.vp.cen_type == 'G' && pair_groups();
            if (vgroups) launch_group_bounds(D, ecen_, p_.ne, grp_e_, st_);
            launch_vf_pairs(D, p_.prec, p_.vp, ecen_, edir_, evol_, p_.ne, p_.tc, p_.tu, p_.ta, p_.nt, sp_vfd_,
                            sp_vft_, part_, st_, vgroups ? grp_e_ : nullptr,
                            vgroups && p_.nt > 0 ? grp_t_ : nullptr);
        }
        {
            ProfScope ps(kProfVfFinalize, st_);
            launch_vf_finalize(D, p_.vp, sp_vfd_.splits, sp_vfd_.splits + (p_.nt > 0 ? sp_vft_.splits : 0), p_.ne,
                               part_, edir_, evol_, dc_, dd_, ep_vf_, st_);
        }
        {
            ProfScope ps(kProfVfGather, st_);
            launch_vf_gather(D, qT, n, p_.ifo, p_.ifo_rows, dc_, dd_, p_.ne, -p_.w, p_adj_, st_);
        }
        {
            ProfScope ps(kProfEnergy, st_);
            launch_energy(ep_fwd_, (T - 1) * nb_fwd_, ep_vf_, nb_vf_, tt_, p_.h, p_.w, f_dev, st_);
        }
    }

    // backward adjoint sweep (Cpp/src/objgrd.cpp:69-87)
    for (int t = T - 2; t >= 0; --t) {
        const double *qt = nde_ + (size_t)t * dn;
        const double *at = alp + (size_t)t * dn;
        {
            ProfScope ps(kProfBwdPairs, st_);
            launch_bwd_pairs(D, p_.order, p_.prec, qt, at, p_adj_, n, p_.width, sp_bwd_, part_, st_,
                             groups ? grp_q_ + (size_t)t * ng : nullptr);
        }
        {
            ProfScope ps(kProfBwdReduce, st_);
            launch_bwd_reduce(D, sp_bwd_.splits, n, part_, p_.h, grd + (size_t)t * dn, p_adj_, st_);
        }
    }
}

// ======================================================================
// L-BFGS with device-resident vectors and scalars
// ======================================================================
namespace {

// device scalar slots; kF..kDD are read back together after a trial
enum : int { kF = 0, kGG, kGD, kSL, kDD, kNumScal };

struct LbfgsState {
    Slab slab;
    double *g = nullptr, *d = nullptr, *xt = nullptr, *gt = nullptr, *S = nullptr, *Y = nullptr;
    double *scal = nullptr, *SY = nullptr, *YY = nullptr, *rdots = nullptr, *coef = nullptr, *work = nullptr;
    double *scratch = nullptr, *rscratch = nullptr;
    double *host = nullptr;   // pinned mirror of scal
    ~LbfgsState()
    {
        if (host) (void)hipHostFree(host);
    }
};

void dots(LbfgsState &s, int nd, const double *const *xs, const double *const *ys, long long n, double *out,
          hipStream_t st)
{
    ProfScope ps(kProfLbfgsDots, st);
    launch_dots(nd, xs, ys, n, s.scratch, out, st);
}

void fetch(LbfgsState &s, int first, int count, hipStream_t st)
{
    LDDMM_HIP(hipMemcpyAsync(s.host + first, s.scal + first, sizeof(double) * count, hipMemcpyDeviceToHost, st));
    LDDMM_HIP(hipStreamSynchronize(st));
}

}  // namespace

// The L-BFGS of Cpp/src/LBFGS.cpp:25-112 with the line search of
// Cpp/src/lineSearch.cpp:26-110.  The direction is the compact form of the
// two-loop recursion (vector_kernels.hip): per accepted pair the tables
// s_i.y_j and y_i.y_j gain one row and column (3h products in one launch),
// and a direction is 2h products in one launch, the recursion coefficients
// in one block and one fused pass that also forms g.d and d.d -- a constant
// number of launches whatever the memory.  The first trial point (step 1)
// is evaluated before the host reads the slope and |d|, so every trial costs
// one host read; the host discards it when the direction is not a descent
// direction.  The ring has min(memNum, itrMax) slots: a run of itrMax
// iterations never stores more pairs, so the iterates are those of memNum.
LbfgsResult run_lbfgs(Objective &obj, double *x_user, const LbfgsOptions &o, hipStream_t st)
{
    auto t0 = std::chrono::steady_clock::now();
    if (o.mem < 1) throw Error(kErrArg, "memNum must be at least 1");
    if (o.itr_max < 0) throw Error(kErrArg, "itrMax must be non-negative");
    const long long nv = obj.problem().varnum();
    const int m = std::max(1, std::min(o.mem, o.itr_max));
    LbfgsState s;
    // the iterate and every L-BFGS vector live in the objective's internal
    // vertex order; x_user is read at the start and written at the end
    double *x = nullptr;
    s.slab.reserve(&x, nv);
    s.slab.reserve(&s.g, nv);
    s.slab.reserve(&s.d, nv);
    s.slab.reserve(&s.xt, nv);
    s.slab.reserve(&s.gt, nv);
    s.slab.reserve(&s.S, (size_t)nv * m);
    s.slab.reserve(&s.Y, (size_t)nv * m);
    s.slab.reserve(&s.scal, kNumScal);
    s.slab.reserve(&s.SY, (size_t)m * m);
    s.slab.reserve(&s.YY, (size_t)m * m);
    s.slab.reserve(&s.rdots, (size_t)3 * m);
    s.slab.reserve(&s.coef, (size_t)1 + 2 * m);
    s.slab.reserve(&s.work, (size_t)3 * m);
    s.slab.reserve(&s.scratch, (size_t)kMaxDots * kDotBlocks);
    s.slab.reserve(&s.rscratch, ring_dot_scratch(3 * m));
    s.slab.allocate();
    LDDMM_HIP(hipHostMalloc((void **)&s.host, sizeof(double) * kNumScal, hipHostMallocDefault));
    obj.to_internal(x_user, x, obj.problem().T - 1);

    LbfgsResult r;
    auto push_trace = [&](double f, double g, double step, int cnt) {
        r.trace_f.push_back(f);
        r.trace_g.push_back(g);
        r.trace_step.push_back(step);
        r.trace_evals.push_back(cnt);
    };

    // f, g at the start (Cpp/src/LBFGS.cpp:50-52)
    obj.eval_internal(x, s.g, s.scal + kF);
    {
        const double *xs[1] = {s.g}, *ys[1] = {s.g};
        dots(s, 1, xs, ys, nv, s.scal + kGG, st);
    }
    fetch(s, kF, 2, st);
    double fnow = s.host[kF];
    double glen = std::sqrt(s.host[kGG]);
    int evals = 1;
    if (o.verbose) {
        printf("%5s   %13s  %13s  %13s  %9s\n", "iter", "f", "|grad f|", "step length", "fcn eval");
        char sep[65] = {0};
        memset(sep, '-', 62);
        printf("%s\n", sep);
        printf("%5d:  %13.6e  %13.6e\n", 0, fnow, glen);
        fflush(stdout);
    }
    push_trace(fnow, glen, 0.0, 1);

    // FP32 objectives finish in FP64: re-evaluate f, g at the current
    // iterate and keep the curvature pairs (no counterpart in the reference,
    // whose arithmetic is FP64 throughout)
    auto refine = [&](const char *why) {
        obj.set_precision(kFP64);
        obj.eval_internal(x, s.g, s.scal + kF);
        const double *xs[1] = {s.g}, *ys[1] = {s.g};
        dots(s, 1, xs, ys, nv, s.scal + kGG, st);
        fetch(s, kF, 2, st);
        fnow = s.host[kF];
        glen = std::sqrt(s.host[kGG]);
        ++evals;
        if (o.verbose) {
            printf("       FP32 phase ended (%s): continuing in FP64, f = %13.6e  |grad f| = %13.6e\n", why, fnow,
                   glen);
            fflush(stdout);
        }
    };
    const bool can_refine = o.fp64_refine && obj.precision() == kFP32;
    // one trial point x + rel d: f, |g|^2 and g.d into kF, kGG, kGD
    auto trial = [&](double rel) {
        {
            ProfScope ps(kProfLbfgsVector, st);
            vec_try(s.xt, x, rel, s.d, nv, st);
        }
        obj.eval_internal(s.xt, s.gt, s.scal + kF);
        const double *xs[2] = {s.gt, s.gt}, *ys[2] = {s.gt, s.d};
        dots(s, 2, xs, ys, nv, s.scal + kGG, st);
    };

    int new_idx = -1, status = 0, stored = 0;
    for (int it = 1; it <= o.itr_max; ++it) {
        if (glen < o.tol) break;
        const bool in_fp32 = obj.precision() == kFP32;

        // direction (Cpp/src/getDirection.cpp:9-66, compact form), g.d and d.d
        const int his = std::min(stored, m);
        {
            if (his > 0) {
                const double *rings[2] = {s.S, s.Y}, *vecs[2] = {s.g, s.g};
                ProfScope ps(kProfLbfgsDots, st);
                launch_ring_dots(2, rings, vecs, his, nv, s.rscratch, s.rdots, st);
            }
            ProfScope ps(kProfLbfgsVector, st);
            launch_lbfgs_coef(s.rdots, s.SY, s.YY, m, his, new_idx < 0 ? 0 : new_idx, s.work, s.coef, st);
            launch_lbfgs_dir(s.d, s.g, s.S, s.Y, nv, s.coef, his, m, s.scratch, s.scal + kSL, st);
        }
        // the first trial (step 1) before the slope is read
        trial(1.0);
        fetch(s, kF, 5, st);
        const double slope = s.host[kSL];
        // a NaN slope is not a descent direction either
        const bool descent = slope <= -DBL_MIN;
        if (!descent && in_fp32 && can_refine) {
            refine("not a descent direction");
            if (o.has_target && fnow <= o.target) {
                r.reached_target = 1;
                break;
            }
            --it;
            continue;
        }
        if (!descent) {
            printf("\nNot a descent direction.\nQuit LBFGS.\n\n");
            fflush(stdout);
            status = 1;
            break;
        }

        // weak-Wolfe line search (Cpp/src/lineSearch.cpp:26-110)
        const double dlen = std::sqrt(s.host[kDD]);
        const int bis_max = (int)std::fmax(50.0, std::log2(1e5 * dlen));
        const int epd_max = (int)std::fmax(10.0, std::log2(1e5 / dlen));
        double lo = 0.0, hi = DBL_MAX, rel = 1.0, fnew = 0.0, gnew = 0.0, step = 0.0;
        int bis = 0, epd = 0, cnt = 0, le = -1;
        bool pending = true;   // the step-1 trial is evaluated and read
        while (bis <= bis_max && epd <= epd_max) {
            if (!pending) {
                trial(rel);
                fetch(s, kF, 3, st);
            }
            pending = false;
            ++cnt;
            fnew = s.host[kF];
            gnew = std::sqrt(s.host[kGG]);
            const double slope_t = s.host[kGD];
            if (gnew < o.tol) {
                step = rel * dlen;
                le = 0;
                break;
            }
            // a NaN energy or slope shrinks the step like a failed Armijo test
            if (fnew >= fnow + o.wolfe1 * rel * slope || std::isnan(fnew) || std::isnan(slope_t)) hi = rel;
            else if (slope_t <= o.wolfe2 * slope) lo = rel;
            else {
                step = rel * dlen;
                le = 0;
                break;
            }
            if (hi == DBL_MAX) {
                rel *= 2.0;
                ++epd;
            } else {
                rel = 0.5 * (lo + hi);
                ++bis;
            }
        }
        evals += cnt;
        if (le != 0 && in_fp32 && can_refine) {
            refine("line search failed");
            if (o.has_target && fnow <= o.target) {
                r.reached_target = 1;
                break;
            }
            --it;
            continue;
        }
        if (le != 0) {
            if (hi == DBL_MAX) {
                printf("\nLine search failed to bracket a point satisfying weak Wolfe conditions.\n"
                       "Function may be unbounded below.\nQuit LBFGS.\n\n");
                status = 2;
            } else {
                printf("\nLine search failed to satisfy weak Wolfe conditions, although\n"
                       "a point satisfying conditions was bracketed in [%e, %e].\nQuit LBFGS.\n\n",
                       lo, hi);
                status = 3;
            }
            fflush(stdout);
            break;
        }

        // accept: curvature pair into the ring buffer (Cpp/src/LBFGS.cpp:94-105)
        // and its row and column of the tables
        new_idx = (new_idx == m - 1) ? 0 : new_idx + 1;
        stored = std::min(stored + 1, m);
        double *sv = s.S + (size_t)new_idx * nv, *yv = s.Y + (size_t)new_idx * nv;
        {
            ProfScope ps(kProfLbfgsVector, st);
            vec_accept(x, s.g, s.xt, s.gt, sv, yv, nv, st);
        }
        {
            const double *rings[3] = {s.Y, s.S, s.Y}, *vecs[3] = {sv, yv, yv};
            ProfScope ps(kProfLbfgsDots, st);
            launch_ring_dots(3, rings, vecs, stored, nv, s.rscratch, s.rdots, st);
            launch_lbfgs_store(s.rdots, stored, new_idx, m, s.SY, s.YY, st);
        }
        const double fprev = fnow;
        fnow = fnew;
        glen = gnew;
        r.iters = it;
        if (o.verbose) {
            printf("%5d:  %13.6e  %13.6e  %13.6e  %9d\n", it, fnow, glen, step, cnt);
            fflush(stdout);
        }
        push_trace(fnow, glen, step, cnt);
        if (in_fp32) ++r.fp32_iters;
        if (o.has_target && fnow <= o.target) {
            // a target met in FP32 arithmetic is confirmed in FP64 before the run stops
            if (in_fp32 && can_refine) refine("target met in FP32, confirming in FP64");
            if (fnow <= o.target) {
                r.reached_target = 1;
                break;
            }
            continue;
        }
        const bool small = (fprev - fnow) < o.refine_rel_tol * std::fabs(fnow);
        if (in_fp32 && can_refine && small) {
            refine("relative decrease below threshold");
            if (o.has_target && fnow <= o.target) {
                r.reached_target = 1;
                break;
            }
            continue;
        }
        if (o.rel_tol > 0.0 && (fprev - fnow) < o.rel_tol * std::fabs(fnow)) break;
    }
    obj.to_user(x, x_user, obj.problem().T - 1);
    LDDMM_HIP(hipStreamSynchronize(st));
    r.f = fnow;
    r.gnorm = glen;
    r.evals = evals;
    r.status = status;
    r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return r;
}

}  // namespace lddmm
