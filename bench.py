"""Benchmark of the LDDMM registration hot path on one MI355X.

`value` (the default `--config sphere_cube`): 3D surface matching, source an
icosphere of radius 1 subdivided 5 times (10,242 vertices, 20,480
triangles), target the surface of the cube [-0.8, 0.8]^3 with 58 x 58 cells
per face (40,368 triangles), Matern nu = 3/2 deformation kernel of width
0.3, varifold Gaussian width 0.15 x Binet, 20 Euler steps, matching weight
100, all FP64.  A step is one evaluation of the objective and its gradient
(forward flow, varifold, adjoint sweep) with every input already resident in
HBM; the metric is kernel pair-interactions per second: (T-1) N^2 forward +
(T-1) N^2 fused backward + Ed^2 + Ed Et varifold pairs per evaluation (the
constant target-target term is computed once per problem and excluded).
The same evaluation under the FP32 policy is reported beside it, with its
errors against FP64, and the roofline of the dominant kernel from HIP
events on the stream it runs on.

`batch` (the registrations/s half of the metric): the full batched
configuration, 256 problems of the uneven generator (vertex counts
log-uniform in [500, 20000], seed 4, T = 21, L-BFGS memory 5).  The plain
FP64 method registers them one after another (zero start, stop at a
relative decrease below 1e-6 or 300 iterations) and sets each problem's
energy; the batched FP32 policy from zero momentum and the batched FP32
policy from the learned warm start (network trained offline by
lddmm_amd.train_warmstart) register all of them until each reaches its
plain energy within 1e-6, confirmed in FP64.  `--batch 32 --batch-nmax 5000`
gives a short run.

`cpu_baseline`: the reference's own Cpp/src kernels (compiled into
oracle/_ref; the FP64 C restatement in oracle/ where that is absent) timed on
a bounded sample on one host core and scaled to one full evaluation; the
batch section carries the same for its 256 registrations.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import lddmm_amd  # noqa: E402
from lddmm_amd import geometry as G  # noqa: E402

T_START = time.perf_counter()

FP64_PEAK = 78.6e12   # MI355X FP64 vector FLOP/s

# Algorithmic FLOPs per pair interaction (FMA = 2, add/sub/mul = 1;
# transcendentals (exp, sqrt) counted separately), minimal algebraic form,
# frozen before measuring.  Keys: (kernel, dim, knl_order or vf kernel pair).
FLOPS_PER_PAIR = {
    ("fwd_pairs", 3, -1): 15, ("fwd_pairs", 3, 1): 17, ("fwd_pairs", 2, -1): 10,
    ("bwd_pairs", 3, -1): 33, ("bwd_pairs", 3, 1): 35, ("bwd_pairs", 2, -1): 22,
    ("vf_pairs", 3, "GB"): 32, ("vf_pairs", 2, "GB"): 23,
    # Matern order 3 (nu = 7/2): cubic polynomial in d for k (3 FMAs instead of the add of order 1),
    # quadratic for the gradient factor (2 FMAs)
    ("fwd_pairs", 3, 3): 22, ("bwd_pairs", 3, 3): 44,
    # 'C' (Cauchy) position kernel: 1 + r^2/s^2 (an FMA), a reciprocal (0) instead of the exp, and
    # one more multiply for the squared kernel of its gradient
    ("vf_pairs", 3, "CB"): 35,
}
# hardware transcendental instructions per pair of the FP32 kernels (v_exp_f32,
# plus v_sqrt_f32 for the Matern distance)
TRANS_PER_PAIR = {("fwd_pairs", -1): 1, ("fwd_pairs", 1): 2, ("bwd_pairs", -1): 1, ("bwd_pairs", 1): 2,
                  ("vf_pairs", "GB"): 1}

# Per-launch counter figures of the pair kernels on the sphere->cube workload, from
# profiles/pair_kernels_sphere_cube.md (rocprofv3 PMC passes on this build):
# traffic = HBM bytes per launch (FETCH_SIZE x 1.9960 + WRITE_SIZE, calibrated on a
# vector kernel of known traffic, k_try), valu_busy = AMD's VALUBusy, valu_issue = share of SIMD
# issue slots at 4 cycles per wave64 VALU instruction.
ROCPROF_SPHERE_CUBE = {
    ("fwd_pairs", "f64"): dict(traffic=8.07e6, valu_busy=0.81, valu_issue=0.77),
    ("bwd_pairs", "f64"): dict(traffic=1.728e7, valu_busy=0.80, valu_issue=0.78),
    ("vf_pairs", "f64"): dict(traffic=2.671e7, valu_busy=0.83, valu_issue=0.86),
    ("fwd_pairs", "f32"): dict(traffic=1.237e7, valu_busy=0.89, valu_issue=0.81),
    ("bwd_pairs", "f32"): dict(traffic=1.727e7, valu_busy=1.02, valu_issue=0.97),
    ("vf_pairs", "f32"): dict(traffic=2.678e7, valu_busy=0.69, valu_issue=0.67),
}


def workload(name):
    if name == "sphere_cube":
        return G.sphere_cube(subdivisions=5, cells=58), dict(itr=300)
    if name == "circle_square":
        return G.circle_square(400), dict(itr=200)
    if name == "sphere2cube_example":
        # the reference's own example (Cpp/examples/sphere2cube/run_sphere2cube.m:1-31): its
        # sphereData.mat / cubeData.mat arrays (tests/golden/sphere2cube.npz), Matern order 3
        # of width 0.25, 'C' 0.1 x 'B' varifold, weight 100, T = 21; L-BFGS to |g| < 1e-2
        from lddmm_amd import io
        z = np.load(os.path.join(ROOT, "tests", "golden", "sphere2cube.npz"))
        w = io.problem_from_arrays({k: z[k] for k in ("iniNdeMat", "elmVtxMat", "tgtCenPosMat", "tgtUniDirMat",
                                                      "tgtElmVolVec")})
        w.update(cen_type="C", cen_width=0.1, dir_type="B", dir_width=0.0, dis_wgt=100.0, knl_order=3,
                 knl_width=0.25, time_stp=0.05, time_num=21)
        return w, dict(itr=100000, tol=1e-2, mem=10000)
    raise SystemExit("unknown config " + name)


def make_objective(w):
    n = w["q0"].shape[0]
    ifo = lddmm_amd.getElmIfo(w["elm"], n, one_based=False)
    obj = lddmm_amd.Objective(w["q0"], w["elm"], ifo, w["tgt_cen"], w["tgt_vol"], w["tgt_dir"], w["cen_type"],
                              w["cen_width"], w["dir_type"], w["dir_width"], w["dis_wgt"], w["knl_order"],
                              w["knl_width"], w["time_stp"], w["time_num"])
    return obj, ifo


def cpu_baseline(w, ifo, pairs_per_eval):
    """The reference's CPU path, one core, on a bounded sample: one forward
    K.alpha product and one adjoint step (K.amp and dq) at the full N, and
    the varifold with gradient of the full deformed surface against a 1/16
    slice of the target; the per-pair cost of each part is scaled to one
    full evaluation (the reference's own evaluation does 3 (T-1) N^2 +
    Ed^2 + Ed Et + Et^2 pair evaluations; the rate is quoted in the new
    path's pair count so the two are comparable per evaluation).

    kind "reference": the reference's own Cpp/src multiplyKernel, dqKernel
    and varifold, compiled from its sources into oracle/_ref (its objgrd
    driver needs MATLAB's BLAS header and is not built); kind "port" where
    oracle/_ref was not built: the same operations in the FP64 C
    restatement (oracle/)."""
    import oracle
    q = w["q0"]
    n, d = q.shape
    T1 = w["time_num"] - 1
    a = 0.01 * G.normal(7, n * d).reshape(n, d)
    L = oracle.ref(d)
    v, g = np.empty_like(q), np.empty_like(q)
    t0 = time.perf_counter()
    if L is not None:
        L.ref_multiply_kernel(v, q, a, w["knl_order"], w["knl_width"], n, d)
    else:
        v = oracle.multiply_kernel(q, a, w["knl_order"], w["knl_width"])
    t_mult = time.perf_counter() - t0
    t0 = time.perf_counter()
    if L is not None:
        L.ref_dq_kernel(g, q, a, v, w["knl_order"], w["knl_width"], n, d)
    else:
        oracle.dq_kernel(q, a, v, w["knl_order"], w["knl_width"])
    t_dq = time.perf_counter() - t0
    et = len(w["tgt_vol"])
    sl = slice(0, max(1, et // 16))
    t0 = time.perf_counter()
    oracle.varifold(q, w["elm"], ifo, w["tgt_cen"][sl], w["tgt_dir"][sl], w["tgt_vol"][sl], w["cen_type"],
                    w["cen_width"], w["dir_type"], w["dir_width"],
                    **(dict(L=L, prefix="ref") if L is not None else {}))
    t_vf = time.perf_counter() - t0
    ed = w["elm"].shape[0]
    ets = sl.stop
    vf_pairs_sample = ed * ed + ed * ets + ets * ets
    vf_pairs_full = ed * ed + ed * et + et * et     # the reference recomputes the target-target term
    t_eval = T1 * (2 * t_mult + t_dq) + t_vf * vf_pairs_full / vf_pairs_sample
    sample_s = t_mult + t_dq + t_vf
    model = _cpu_model()
    what = ("the reference's Cpp/src multiplyKernel, dqKernel and varifold (oracle/_ref)" if L is not None
            else "oracle/lddmm_oracle.c")
    return dict(value=pairs_per_eval / t_eval, unit="pair-interactions/s", cores=1,
                kind="reference" if L is not None else "port",
                host=dict(cpu=model, logical_cpus=os.cpu_count()),
                sample=("%s, 1 core: K.alpha + dq at N=%d and the varifold with gradient vs %d of %d target "
                        "triangles (%.1f s of CPU work), scaled to one full evaluation = %.1f s"
                        % (what, n, ets, et, sample_s, t_eval)),
                seconds_per_eval=t_eval)


def c4_cpu_baseline(probs, evals, pairs, sample=2):
    """The reference's CPU path on the batched configuration, one core,
    extrapolated from a bounded sample: the reference's own multiplyKernel,
    dqKernel and varifold (oracle/_ref; the oracle where it was not built)
    timed on the `sample` problems nearest the median size (one forward
    product, one adjoint step and the varifold with gradient against a slice
    of the target), giving a per-pair cost that is applied to every
    problem's pair count times the evaluations its plain run took."""
    import oracle
    k = G.BATCH_KERNELS
    T1 = k["time_num"] - 1
    order = np.argsort([len(p["q0"]) for p in probs])
    pick = order[len(order) // 2: len(order) // 2 + sample]
    L = oracle.ref(3)
    t_pts = t_vf = 0.0
    n_pts = n_vf = 0
    for i in pick:
        p = probs[i]
        q = p["q0"]
        n = len(q)
        a = 0.01 * G.normal(7, q.size).reshape(q.shape)
        v, g = np.empty_like(q), np.empty_like(q)
        t0 = time.perf_counter()
        if L is not None:
            L.ref_multiply_kernel(v, q, a, k["knl_order"], p["knl_width"], n, 3)
            L.ref_dq_kernel(g, q, a, v, k["knl_order"], p["knl_width"], n, 3)
            L.ref_multiply_kernel(v, q, a, k["knl_order"], p["knl_width"], n, 3)
        else:
            v = oracle.multiply_kernel(q, a, k["knl_order"], p["knl_width"])
            oracle.dq_kernel(q, a, v, k["knl_order"], p["knl_width"])
            oracle.multiply_kernel(q, a, k["knl_order"], p["knl_width"])
        t_pts += time.perf_counter() - t0
        n_pts += 3 * n * n
        ifo = lddmm_amd.getElmIfo(p["elm"], n, one_based=False)
        et = len(p["tgt_vol"])
        sl = slice(0, max(1, et // 8))
        t0 = time.perf_counter()
        oracle.varifold(q, p["elm"], ifo, p["tgt_cen"][sl], p["tgt_dir"][sl], p["tgt_vol"][sl], k["cen_type"],
                        k["cen_width"], k["dir_type"], k["dir_width"], **(dict(L=L, prefix="ref") if L else {}))
        t_vf += time.perf_counter() - t0
        ed = len(p["elm"])
        n_vf += ed * ed + ed * sl.stop + sl.stop * sl.stop
    per_pt, per_vf = t_pts / n_pts, t_vf / n_vf
    total = 0.0
    for p, e in zip(probs, evals):
        n, ed, et = len(p["q0"]), len(p["elm"]), len(p["tgt_vol"])
        total += e * (3 * T1 * n * n * per_pt + (ed * ed + ed * et + et * et) * per_vf)
    return dict(value=len(probs) / total, unit="registrations/s", cores=1,
                kind="reference" if L is not None else "port",
                sample="the reference's Cpp/src multiplyKernel, dqKernel and varifold (oracle/_ref), 1 core, on "
                       "%d median-size problems (%.1f s of CPU work): %.3g s per point pair, %.3g s per varifold "
                       "pair, applied to every problem's pairs x the evaluations of its plain run = %.0f s for "
                       "the %d registrations" % (len(pick), t_pts + t_vf, per_pt, per_vf, total, len(probs)),
                seconds=total)


def warm_start_leg(probs, targets, e_ref, B, B64, t_setup, itr, weights):
    """The learned warm start on a batch whose plain energies are known,
    with the network trained offline by lddmm_amd.train_warmstart (its
    parameters and training record in `weights`).  Timed: the features
    (including the start's varifold gradient), inference, the per-problem
    start scale (choose_start: the multiple of the prediction, zero included,
    with the lowest objective, from energy-only evaluations of the batch on
    the device), the batched FP32 registration to the plain energies from
    that start (at most 1.25 times the plain iteration limit) and the
    fallback, a zero-momentum batched FP32 run for every problem the
    warm-started run did not bring to its target.  B and B64 are the batch in
    FP32 and FP64 (the latter re-evaluates results)."""
    from lddmm_amd import warmstart as W
    k = G.BATCH_KERNELS
    T = k["time_num"]
    ws = W.WarmStart.load(weights)
    t0 = time.perf_counter()
    base = ws.start_vectors(probs)
    t1 = time.perf_counter()
    scale, _, _ = W.choose_start(B, base)
    t2 = time.perf_counter()
    aw, iw, rw = B.run_scaled(base, scale, int(1.25 * itr), 0.0, 0.0, 0.5, 5, precision="fp32", targets=targets)
    t3 = time.perf_counter()
    missed = [i for i, r in enumerate(iw) if not r["reached_target"]]
    # where the warm-started run left each missed problem: its multiple and
    # the relative gap of its energy (batch precision) to the target
    missed_at = [dict(problem=i, n=int(B.n[i]), start_multiple=float(scale[i]),
                      rel_gap=float((iw[i]["fcn"] - targets[i]) / abs(targets[i]))) for i in missed]
    its_fb = []
    if missed:
        Bf = lddmm_amd.Batch([probs[i] for i in missed], k["cen_type"], k["cen_width"], k["dir_type"],
                             k["dir_width"], k["dis_wgt"], k["knl_order"], k["time_stp"], T, precision="fp32")
        af, ifb, _ = Bf.run(3 * itr, 0.0, 0.0, 0.5, 5, precision="fp32", targets=targets[missed])
        Bf.close()
        for j, i in enumerate(missed):
            aw[i] = af[j]
        its_fb = [r["iterations"] for r in ifb]
    t4 = time.perf_counter()
    t_warm = t4 - t0 + t_setup
    t_inf = t1 - t0
    ws.close()
    f, _ = B64.objgrd(aw)
    rel_w = (f - e_ref) / np.abs(e_ref)
    count = len(probs)
    m = ws.meta
    return dict(seconds=t_warm, registrations_per_s=count / t_warm, rounds=rw, inference_seconds=t_inf,
                phase_seconds=dict(setup=t_setup, features_and_inference=t_inf, choose_start=t2 - t1,
                                   warm_run=t3 - t2, fallback=t4 - t3),
                start_multiples={str(c): int(np.sum(scale == c)) for c in np.unique(scale)},
                median_iterations=float(np.median([r["iterations"] for r in iw])),
                warm_start_used=int(np.sum(scale > 0)), reached_target=count - len(missed),
                fallback_problems=len(missed), fallback_iterations=its_fb, fallback_from=missed_at,
                within_1e6=int(np.sum(rel_w <= 1e-6 + 1e-8)), max_rel_energy_vs_plain=float(rel_w.max()),
                training=dict(problems=int(m.get("problems", 0)), rows=int(m.get("rows", 0)),
                              solver="plain FP64 L-BFGS (batched evaluation), other seeds than the benchmark's",
                              solve_seconds=m.get("solve_seconds"), fit_seconds=m.get("fit_seconds"),
                              adam_steps=int(m.get("steps", 0)), loss_first=m.get("loss_first"),
                              loss_last=m.get("loss_last"), weights=os.path.relpath(weights, ROOT)))


WARM_WEIGHTS = os.path.join(ROOT, "lddmm_amd", "data", "warmstart_c4.npz")


def launch_shapes_report(count, nmin, nmax, seed=4):
    """FP32 K.alpha and adjoint pair-kernel time per evaluation of the
    uneven batch (HIP events around each launch group, direct launches) at
    the autotuned launch shape and at every shape of the grid source chunk
    {1024, 2048, 4096} x targets per lane (K.alpha, adjoint) {(1, 1), (2, 2),
    (2, 1), (4, 2), (1, 4), (2, 4), (4, 4)} (adjoint 4: the matrix-core
    adjoint of 3D Gaussian batches), at momenta 0.02 x N(0, 1)."""
    probs = G.batch_problems(count, seed, nmin, nmax)
    k = G.BATCH_KERNELS
    T = k["time_num"]
    t0 = time.perf_counter()
    B = lddmm_amd.Batch(probs, k["cen_type"], k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"],
                        k["knl_order"], k["time_stp"], T, precision="fp32", chunk=0)
    t_setup = time.perf_counter() - t0
    tuned = B.tuning()
    B.set_graphs(False)
    alps = [0.02 * G.normal(100 + i, B.sizes[i]).reshape(T - 1, B.n[i], 3) for i in range(B.P)]
    rows = []
    for sh in [tuned] + [(c, kv, bw) for c in (1024, 2048, 4096)
                         for kv, bw in ((1, 1), (2, 2), (2, 1), (4, 2), (1, 4), (2, 4), (4, 4))]:
        B.set_tuning(*sh)
        B.objgrd(alps)
        lddmm_amd.profile_reset()
        lddmm_amd.profile_enable(True)
        for _ in range(2):
            B.objgrd(alps)
        pr = lddmm_amd.profile_read()
        lddmm_amd.profile_enable(False)
        fw, bw = pr["fwd_pairs"][0] / 2, pr["bwd_pairs"][0] / 2
        rows.append(dict(chunk=sh[0], kv_lanes=sh[1], bwd_lanes=sh[2], fwd_pairs_ms=fw, bwd_pairs_ms=bw,
                         sum_ms=fw + bw))
        print("launch shape %s: K.alpha %.2f ms, adjoint %.2f ms per evaluation" % (sh, fw, bw), file=sys.stderr,
              flush=True)
    B.close()
    return dict(problems=count, n_range=[nmin, nmax], setup_seconds=t_setup,
                tuned=dict(chunk=tuned[0], kv_lanes=tuned[1], bwd_lanes=tuned[2]), shapes=rows)


def batch_bench(count, nmin, nmax, itr, seed=4, warm=True, batched_fp64=False, budget_s=None, t_start=None):
    """The batched-registration configuration: `count` problems of the
    uneven generator (vertex counts log-uniform in [nmin, nmax]; 256 in
    [500, 20000] is the full configuration).  The plain FP64 method
    registers them one after another (zero start, stop at a relative energy
    decrease below 1e-6 or `itr` iterations) and sets each problem's energy;
    the batched FP32 policy then registers all of them until each reaches
    its plain energy x (1 + 1e-6) (at most 3 `itr` iterations, finishing in
    FP64 on a stall), and so does the batched FP32 policy from the learned
    warm start.  Every result is re-evaluated in FP64 against the plain
    energy.  Batched times include the batch's set-up, as plain times
    include each run's.  The warm-start leg is skipped (and says so) when
    `budget_s` seconds since `t_start` have passed before it starts."""
    probs = G.batch_problems(count, seed, nmin, nmax)
    k = G.BATCH_KERNELS
    T = k["time_num"]
    seq, a_plain = [], []
    t0 = time.perf_counter()
    for i, p in enumerate(probs):
        ifo = lddmm_amd.getElmIfo(p["elm"], len(p["q0"]), one_based=False)
        a, info = lddmm_amd.run(p["q0"], None, p["elm"], ifo, p["tgt_cen"], p["tgt_vol"], p["tgt_dir"],
                                k["cen_type"], k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"],
                                k["knl_order"], p["knl_width"], k["time_stp"], k["time_num"], itr, 0.0, 0.0, 0.5, 5,
                                rel_energy_tol=1e-6)
        seq.append(info)
        a_plain.append(a)
        if i % 16 == 15:
            print("plain FP64: %d of %d problems, %.1f s" % (i + 1, count, time.perf_counter() - t0),
                  file=sys.stderr, flush=True)
    t_seq = time.perf_counter() - t0
    e_ref = np.array([i["fcn"] for i in seq])
    targets = e_ref * (1.0 + 1e-6)

    def make(ps, prec, chunk=0):
        return lddmm_amd.Batch(ps, k["cen_type"], k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"],
                               k["knl_order"], k["time_stp"], T, precision=prec, chunk=chunk)

    # the batch's set-up (spatial orders, uploads, target terms, chunk autotune)
    # counts toward every batched time, as each plain run's set-up counts toward its
    t0 = time.perf_counter()
    B = make(probs, "fp32")
    chunk, pairs = B.info()
    tuning = B.tuning()
    t_setup = time.perf_counter() - t0
    a32, i32, r32 = B.run(3 * itr, 0.0, 0.0, 0.5, 5, precision="fp32", targets=targets, progress=100)
    t32 = i32[0]["seconds"] + t_setup
    print("batched FP32: %.1f s" % t32, file=sys.stderr, flush=True)
    B64 = make(probs, "fp64", chunk)
    f64, _ = B64.objgrd(a32)
    rel = (f64 - e_ref) / np.abs(e_ref)
    f_plain, _ = B64.objgrd(a_plain)   # consistency of the two FP64 evaluations
    out = dict(problems=count, n_range=[int(min(B.n)), int(max(B.n))], chunk=chunk,
               fp32_targets_per_lane=dict(kv=tuning[1], bwd=tuning[2]),
               pairs_per_eval_total=float(pairs.sum()), batch_setup_seconds=t_setup,
               plain_fp64_sequential=dict(seconds=t_seq, registrations_per_s=count / t_seq,
                                          median_iterations=float(np.median([i["iterations"] for i in seq])),
                                          evaluations=int(sum(i["evaluations"] for i in seq))),
               batched_fp32=dict(seconds=t32, registrations_per_s=count / t32, rounds=r32,
                                 median_iterations=float(np.median([r["iterations"] for r in i32])),
                                 max_iterations=int(max(r["iterations"] for r in i32)),
                                 reached_target=int(sum(r["reached_target"] for r in i32)),
                                 within_1e6=int(np.sum(rel <= 1e-6 + 1e-8)),
                                 max_rel_energy_vs_plain=float(rel.max()),
                                 missed=[dict(problem=i, n=int(B.n[i]), plain_energy=float(e_ref[i]),
                                              fp32_energy_fp64=float(f64[i]), iterations=int(i32[i]["iterations"]),
                                              status=int(i32[i]["status"]))
                                         for i in range(count) if not i32[i]["reached_target"]]),
               plain_energy_check=dict(max_rel_diff_batched_fp64=float(np.max(np.abs(f_plain - e_ref) /
                                                                             np.abs(e_ref)))),
               speedup_batched_fp32_vs_plain_fp64=t_seq / t32)
    if batched_fp64:
        a64, i64, r64 = B.run(itr, 0.0, 0.0, 0.5, 5, precision="fp64", rel_energy_tol=1e-6)
        t64 = i64[0]["seconds"] + t_setup
        out["batched_fp64"] = dict(seconds=t64, registrations_per_s=count / t64, rounds=r64,
                                   median_iterations=float(np.median([i["iterations"] for i in i64])))
    if warm:
        spent = time.perf_counter() - t_start if t_start is not None else 0.0
        if not os.path.exists(WARM_WEIGHTS):
            out["batched_fp32_warm_start"] = dict(skipped="no trained weights at " +
                                                  os.path.relpath(WARM_WEIGHTS, ROOT))
        elif budget_s is not None and spent > budget_s:
            out["batched_fp32_warm_start"] = dict(skipped="time budget: %.0f s spent before the leg" % spent)
        else:
            wleg = warm_start_leg(probs, targets, e_ref, B, B64, t_setup, itr, WARM_WEIGHTS)
            wleg["median_iterations_plain"] = float(np.median([i["iterations"] for i in seq]))
            out["batched_fp32_warm_start"] = wleg
            out["speedup_batched_fp32_warm_vs_plain_fp64"] = t_seq / wleg["seconds"]
            out["warm_vs_zero_start"] = t32 / wleg["seconds"]
    out["cpu_baseline"] = c4_cpu_baseline(probs, [i["evaluations"] for i in seq], pairs)
    out["speedup_plain_fp64_gpu_vs_cpu"] = out["cpu_baseline"]["seconds"] / t_seq
    B.close()
    B64.close()
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baselines_report(threads=16, rows=1500):
    """CPU baselines of the single-problem configurations, beside the GPU on
    the same host.

    circle->square (400 landmarks, 2D, 200 L-BFGS iterations): the plain FP64
    registration on the GPU, and end to end on one core the oracle's full run
    (the FP64 C restatement of Cpp/src objgrd, LBFGS, getDirection and
    lineSearch; kind "port": the reference's own driver is not built), with
    final energies and evaluation counts.

    The single large problem (seeded 100,002-vertex superquadric, Gaussian
    kernel of width 0.2): one matrix-free forward Euler step v = K(q) alpha
    (1.0e10 point pairs).  GPU: the multiplyKernel entry point in FP64 (host
    arrays in and out).  CPU: the reference's own multiplyKernel (oracle/_ref,
    its cross form; the oracle's where oracle/_ref was not built) on `rows`
    target rows against all sources, on one core and on `threads` host threads
    at once (each its own rows; the reference has no threading, so the
    all-core figure is this build's parallelisation of the reference's
    kernel), extrapolated to every row."""
    import concurrent.futures as cf
    import oracle
    out = dict(host=dict(cpu=_cpu_model(), logical_cpus=os.cpu_count()))

    # circle->square, 200 iterations, GPU and the oracle's full run on one core
    w, extra = workload("circle_square")
    n = w["q0"].shape[0]
    ifo = lddmm_amd.getElmIfo(w["elm"], n, one_based=False)
    common = (w["q0"], None, w["elm"], ifo, w["tgt_cen"], w["tgt_vol"], w["tgt_dir"], w["cen_type"], w["cen_width"],
              w["dir_type"], w["dir_width"], w["dis_wgt"], w["knl_order"], w["knl_width"], w["time_stp"],
              w["time_num"])
    lddmm_amd.run(*common, 5, 0.0, 0.0, 0.5, 5)   # first-call set-up
    t0 = time.perf_counter()
    _, ig = lddmm_amd.run(*common, extra["itr"], 0.0, 0.0, 0.5, 5)
    t_gpu = time.perf_counter() - t0
    P = oracle.OracleProblem(w["q0"], w["elm"], w["tgt_cen"], w["tgt_dir"], w["tgt_vol"], w["cen_type"],
                             w["cen_width"], w["dir_type"], w["dir_width"], w["dis_wgt"], w["knl_order"],
                             w["knl_width"], w["time_stp"], w["time_num"])
    t0 = time.perf_counter()
    _, f_cpu, tr = P.run(itr_max=extra["itr"], tol=0.0, wolfe1=0.0, wolfe2=0.5, mem=5)
    t_cpu = time.perf_counter() - t0
    out["circle_square"] = dict(
        iterations=extra["itr"], gpu_seconds=t_gpu, gpu_energy=ig["fcn"], gpu_evaluations=ig["evaluations"],
        cpu=dict(value=t_cpu, unit="s per registration", cores=1, kind="port",
                 sample="the oracle's full %d-iteration FP64 run, end to end" % extra["itr"]),
        cpu_energy=f_cpu, cpu_evaluations=int(tr["total_evals"]),
        rel_energy_gpu_vs_cpu=abs(ig["fcn"] - f_cpu) / abs(f_cpu), gpu_speedup=t_cpu / t_gpu)
    print("circle->square: GPU %.2f s, CPU %.1f s" % (t_gpu, t_cpu), file=sys.stderr, flush=True)

    # the large problem: one forward Euler step K(q) alpha
    w = G.superquadric_pair(100, 31, knl_width=0.2)
    q = np.ascontiguousarray(w["q0"], dtype=np.float64)
    n, d = q.shape
    a = 0.01 * G.normal(7, q.size).reshape(q.shape)
    lddmm_amd.multiplyKernel(q, a, w["knl_order"], w["knl_width"])   # set-up
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        lddmm_amd.multiplyKernel(q, a, w["knl_order"], w["knl_width"])
    t_gpu = (time.perf_counter() - t0) / reps
    L = oracle.ref(3)
    nthr = max(1, min(int(threads), os.cpu_count() or 1))

    def slice_rows(k):
        qi = np.ascontiguousarray(q[k * rows:(k + 1) * rows])
        v = np.empty_like(qi)
        if L is not None:
            L.ref_multiply_kernel_cross(v, qi, q, a, w["knl_order"], w["knl_width"], len(qi), n, d)
        else:
            v = oracle.multiply_kernel_cross(qi, q, a, w["knl_order"], w["knl_width"])
        return v

    t0 = time.perf_counter()
    slice_rows(0)
    t_one = time.perf_counter() - t0
    with cf.ThreadPoolExecutor(nthr) as ex:
        t0 = time.perf_counter()
        list(ex.map(slice_rows, range(nthr)))
        t_all = time.perf_counter() - t0
    pairs = float(n) * n
    step_one = t_one * pairs / (rows * n)
    step_all = t_all * pairs / (nthr * rows * n)
    what = ("the reference's Cpp/src multiplyKernel (oracle/_ref)" if L is not None
            else "oracle/lddmm_oracle.c multiply_kernel_cross")
    out["large_single_step"] = dict(
        n=n, pairs=pairs, gpu_seconds=t_gpu, gpu_pairs_per_s=pairs / t_gpu,
        cpu_one_core=dict(value=pairs / step_one, unit="pair-interactions/s", cores=1,
                          kind="reference" if L is not None else "port",
                          sample="%s, %d target rows x %d sources (%.1f s), scaled to the full step = %.1f s"
                                 % (what, rows, n, t_one, step_one), seconds_per_step=step_one),
        cpu_all_cores=dict(value=pairs / step_all, unit="pair-interactions/s", cores=nthr,
                           kind="reference" if L is not None else "port",
                           sample="%d host threads at once, %d rows each (%.1f s wall), scaled to the full "
                                  "step = %.1f s" % (nthr, rows, t_all, step_all), seconds_per_step=step_all),
        gpu_speedup_vs_one_core=step_one / t_gpu, gpu_speedup_vs_all_cores=step_all / t_gpu)
    print("large step: GPU %.3f s, CPU 1 core %.0f s, %d cores %.1f s" % (t_gpu, step_one, nthr, step_all),
          file=sys.stderr, flush=True)
    return out


def culling_report(count=32, seed=4, nmin=500, nmax=5000, reps=5):
    """Culling of far FP32 Gaussian tile pairs (kernel values below 2^-48):
    mean time per launch group of the pair kernels with culling off and on,
    on the sphere->cube FP32 evaluation and on the uneven batch (all problems
    active, seeded momentum), from the library's HIP-event accounting."""
    def groups(fn):
        out, res = {}, {}
        for on in (False, True):
            lddmm_amd.set_pair_culling(on)
            res[on] = fn()
            lddmm_amd.profile_reset()
            lddmm_amd.profile_enable(True)
            for _ in range(reps):
                fn()
            lddmm_amd.profile_enable(False)
            r = lddmm_amd.profile_read()
            out["on" if on else "off"] = {k: r[k][0] / r[k][1] for k in ("fwd_pairs", "bwd_pairs", "vf_pairs")
                                          if r.get(k, (0, 0))[1]}
        lddmm_amd.set_pair_culling(True)
        # results with culling on against every pair computed (an objective
        # returns one energy and gradient, a batch one per problem)
        def split(r):
            f, g = r
            return np.atleast_1d(np.asarray(f, dtype=float)), (list(g) if isinstance(g, list) else [g])
        (f0, g0), (f1, g1) = split(res[False]), split(res[True])
        out["max_rel_energy_diff"] = float(np.max(np.abs(f1 - f0) / np.abs(f0)))
        out["max_rel_grad_diff"] = max(float(np.max(np.abs(a - b)) / np.max(np.abs(b))) for a, b in zip(g1, g0))
        return out
    w, _ = workload("sphere_cube")
    n, T = len(w["q0"]), w["time_num"]
    obj, _ = make_objective(w)
    obj.set_precision("fp32")
    alp = 0.02 * G.normal(1, (T - 1) * n * 3)
    c2 = groups(lambda: obj(alp))
    probs = G.batch_problems(count, seed, nmin, nmax)
    k = G.BATCH_KERNELS
    B = lddmm_amd.Batch(probs, k["cen_type"], k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"],
                        k["knl_order"], k["time_stp"], k["time_num"], precision="fp32")
    alps = [0.02 * G.normal(10 + i, B.sizes[i]).reshape(k["time_num"] - 1, B.n[i], 3) for i in range(B.P)]
    batch = groups(lambda: B.objgrd(alps))
    return dict(report="culling below 2^-48, ms per launch group (fwd/bwd: one Euler step; vf: DD + DT of one "
                       "evaluation)", sphere_cube_fp32=c2, batch_fp32=dict(problems=count, seed=seed, **batch))


def precision_report(freq=100, seed=31, itr=10):
    """FP32 policy vs FP64 on the single large problem: a seeded superquadric
    source (geodesic frequency 100: 100,002 vertices, 200,000 triangles) and an
    independent superquadric target (200,000 triangles), Gaussian kernel of
    width 0.2, varifold Gaussian 0.1 x Binet, T-1 = 20.  Reports (1) one
    evaluation in both policies at the same momentum (the FP64 run's result
    after `itr` iterations): relative energy, gradient (max norm) and final
    vertex error; (2) `itr`-iteration registrations from zero momentum in each
    policy (FP32 without the FP64 finish), the FP32 result re-evaluated in
    FP64, and the distance between the two final shapes."""
    w = G.superquadric_pair(freq, seed, knl_width=0.2)
    obj, ifo = make_objective(w)
    n, d = w["q0"].shape
    common = (w["q0"], None, w["elm"], ifo, w["tgt_cen"], w["tgt_vol"], w["tgt_dir"], w["cen_type"],
              w["cen_width"], w["dir_type"], w["dir_width"], w["dis_wgt"], w["knl_order"], w["knl_width"],
              w["time_stp"], w["time_num"])
    a64, i64 = lddmm_amd.run(*common, itr, 0.0, 0.0, 0.5, 5, precision="fp64")
    a32, i32 = lddmm_amd.run(*common, itr, 0.0, 0.0, 0.5, 5, precision="fp32", fp64_refine=False)
    scale = float(np.max(np.abs(w["q0"])))
    t0 = time.perf_counter()
    f64, g64 = obj(a64)
    t_e64 = time.perf_counter() - t0
    q64 = obj.trajectory()[-1]
    obj.set_precision("fp32")
    t0 = time.perf_counter()
    f32, g32 = obj(a64)
    t_e32 = time.perf_counter() - t0
    q32 = obj.trajectory()[-1]
    obj.set_precision("fp64")
    e32, _ = obj(a32)
    qr32 = obj.trajectory()[-1]
    _ = obj(a64)
    qr64 = obj.trajectory()[-1]
    disp = float(np.sqrt(np.mean(np.sum((q64 - w["q0"]) ** 2, axis=1))))
    return dict(
        workload="superquadric pair, N=%d vertices, Ed=%d, Et=%d triangles, Gaussian w=0.2, varifold G0.1xBinet, "
                 "T-1=%d, seed %d" % (n, w["elm"].shape[0], len(w["tgt_vol"]), w["time_num"] - 1, seed),
        pairs_per_eval=obj.pairs_per_eval(),
        same_momentum=dict(energy_fp64=f64, energy_fp32=f32, energy_rel_err=abs(f32 - f64) / abs(f64),
                           grad_rel_err=float(np.max(np.abs(g32 - g64)) / np.max(np.abs(g64))),
                           final_vertex_max_err=float(np.max(np.abs(q32 - q64))),
                           final_vertex_max_err_rel_to_extent=float(np.max(np.abs(q32 - q64))) / scale,
                           rms_displacement=disp, eval_seconds_fp64=t_e64, eval_seconds_fp32=t_e32),
        registration=dict(iterations=itr, fp64=dict(energy=i64["fcn"], seconds=i64["seconds"],
                                                    evaluations=i64["evaluations"]),
                          fp32=dict(energy=i32["fcn"], energy_fp64_eval=e32, seconds=i32["seconds"],
                                    evaluations=i32["evaluations"]),
                          energy_rel_diff=(e32 - i64["fcn"]) / abs(i64["fcn"]),
                          final_shape_max_dist=float(np.max(np.linalg.norm(qr32 - qr64, axis=1))),
                          final_shape_rms_dist=float(np.sqrt(np.mean(np.sum((qr32 - qr64) ** 2, axis=1))))))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="sphere_cube")
    ap.add_argument("--registration", action="store_true",
                    help="also one full L-BFGS registration of the workload (FP64 and the FP32 policy)")
    ap.add_argument("--no-registration", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--reg-iters", type=int, default=0, help="L-BFGS iterations of the registration (0 = config)")
    ap.add_argument("--batch", type=int, default=256,
                    help="problems of the batched-registration leg (256: the full configuration; 0 = skip)")
    ap.add_argument("--batch-nmin", type=int, default=500)
    ap.add_argument("--batch-nmax", type=int, default=20000)
    ap.add_argument("--batch-chunk", type=int, default=512, help="source chunk of --batch-profile")
    ap.add_argument("--batch-lanes", default="2,2",
                    help="FP32 targets per lane of --batch-profile, K.alpha,adjoint (adjoint 4: matrix cores)")
    ap.add_argument("--batch-evals", type=int, default=0,
                    help="--batch-profile: this many objective evaluations at fixed momenta instead of a run")
    ap.add_argument("--batched-fp64", action="store_true", help="also the batched FP64 run of the batch leg")
    ap.add_argument("--no-warm-start", action="store_true", help="skip the learned warm start of the batch leg")
    ap.add_argument("--budget", type=float, default=470.0,
                    help="seconds after which the warm-start leg is skipped rather than started")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision-report", action="store_true",
                    help="only the FP32-vs-FP64 report on the 100k-vertex superquadric pair")
    ap.add_argument("--batch-profile", action="store_true",
                    help="only the batched FP32 registration of the uneven batch (for kernel traces)")
    ap.add_argument("--culling-report", action="store_true",
                    help="only the pair-kernel times with culling off and on")
    ap.add_argument("--launch-shapes", action="store_true",
                    help="only the batched FP32 pair-kernel times of the uneven batch at each launch shape")
    ap.add_argument("--cpu-baselines", action="store_true",
                    help="only the CPU baselines of circle->square (full run) and of one step of the large problem")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (problems would shard by "
              "independent registrations, or a single large problem by target rows)", file=sys.stderr)
        sys.exit(2)

    if args.precision_report:
        print(json.dumps(precision_report()), flush=True)
        return
    if args.culling_report:
        print(json.dumps(culling_report()), flush=True)
        return
    if args.cpu_baselines:
        print(json.dumps(cpu_baselines_report()), flush=True)
        return
    if args.launch_shapes:
        print(json.dumps(launch_shapes_report(args.batch, args.batch_nmin, args.batch_nmax)), flush=True)
        return
    if args.batch_profile:
        probs = G.batch_problems(args.batch, 4, args.batch_nmin, args.batch_nmax)
        k = G.BATCH_KERNELS
        # a fixed source chunk (default 512, the autotuner's choice for the
        # 32-problem batch when nothing is profiling): under PMC counters its
        # timing launches can pick differently from pass to pass
        B = lddmm_amd.Batch(probs, k["cen_type"], k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"],
                            k["knl_order"], k["time_stp"], k["time_num"], precision="fp32", chunk=args.batch_chunk)
        kv_lanes, bwd_lanes = (int(x) for x in args.batch_lanes.split(","))
        B.set_tuning(args.batch_chunk, kv_lanes, bwd_lanes)
        chunk, pairs = B.info()
        # direct launches (as whenever rocprofv3 drives the process): its kernel
        # tracing has crashed the host process during hipGraph capture here;
        # graphs do not change kernel times
        B.set_graphs(False)
        if args.batch_evals > 0:
            # fixed momenta 0.02 N(0, 1) (as --launch-shapes), one untimed evaluation first
            T = k["time_num"]
            alps = [0.02 * G.normal(100 + i, B.sizes[i]).reshape(T - 1, B.n[i], 3) for i in range(B.P)]
            B.objgrd(alps)
            lddmm_amd.profile_reset()
            lddmm_amd.profile_enable(True)
            for _ in range(args.batch_evals):
                B.objgrd(alps)
            pr = lddmm_amd.profile_read()
            lddmm_amd.profile_enable(False)
            print(json.dumps(dict(problems=args.batch, n_range=[int(min(B.n)), int(max(B.n))], chunk=chunk,
                                  lanes=[kv_lanes, bwd_lanes], evaluations=args.batch_evals,
                                  fwd_pairs_ms=pr["fwd_pairs"][0] / args.batch_evals,
                                  bwd_pairs_ms=pr["bwd_pairs"][0] / args.batch_evals)), flush=True)
            return
        _, info, rounds = B.run(300, 0.0, 0.0, 0.5, 5, precision="fp32", rel_energy_tol=1e-6)
        print(json.dumps(dict(problems=args.batch, chunk=chunk, lanes=[kv_lanes, bwd_lanes],
                              pairs_per_eval_total=float(pairs.sum()),
                              seconds=info[0]["seconds"], rounds=rounds,
                              evaluations=int(sum(i["evaluations"] for i in info)))), flush=True)
        return

    w, extra = workload(args.config)
    obj, ifo = make_objective(w)
    n, d = w["q0"].shape
    T = w["time_num"]
    nv = obj.varnum
    # a seeded, non-trivial momentum in the device layout (T-1, D, N)
    alp = (0.02 * G.normal(11, nv)).reshape(T - 1, d, n)
    d_alp = lddmm_amd.DeviceArray(alp)
    d_grd = lddmm_amd.DeviceArray(nbytes=nv * 8)
    d_f = lddmm_amd.DeviceArray(nbytes=8)
    pairs = obj.pairs_per_eval()

    for _ in range(args.warmup):
        obj.eval_device(d_f.ptr, d_grd.ptr, d_alp.ptr)
    obj.synchronize()
    lddmm_amd.profile_reset()
    lddmm_amd.profile_enable(True)
    lddmm_amd.lib().lddmm_device_synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        obj.eval_device(d_f.ptr, d_grd.ptr, d_alp.ptr)
    obj.synchronize()
    lddmm_amd.lib().lddmm_device_synchronize()
    elapsed = time.perf_counter() - t0
    lddmm_amd.profile_enable(False)
    prof = lddmm_amd.profile_read()
    ms_step = 1e3 * elapsed / args.steps
    value = pairs * args.steps / elapsed

    # dominant kernel and its roofline (algorithmic FLOPs / HIP-event time)
    order = w["knl_order"]
    vfkey = w["cen_type"] + w["dir_type"]
    per_launch = {
        "fwd_pairs": (n * n, FLOPS_PER_PAIR.get(("fwd_pairs", d, order))),
        "bwd_pairs": (n * n, FLOPS_PER_PAIR.get(("bwd_pairs", d, order))),
        "vf_pairs": (w["elm"].shape[0] * (w["elm"].shape[0] + len(w["tgt_vol"])),
                     FLOPS_PER_PAIR.get(("vf_pairs", d, vfkey))),
    }
    kernels = {}
    for k, (np_, fl) in per_launch.items():
        ms, cnt = prof.get(k, (0.0, 0))
        if cnt:
            avg = ms / cnt * 1e-3
            kernels[k] = dict(avg_ms=1e3 * avg, launches=cnt, pairs_per_launch=np_,
                              pairs_per_s=np_ / avg, flops_per_pair=fl,
                              tflops=(np_ * fl / avg / 1e12) if fl else None)
    dom = max(kernels, key=lambda k: prof[k][0])
    kd = kernels[dom]
    rp = ROCPROF_SPHERE_CUBE.get((dom, "f64")) if args.config == "sphere_cube" else None
    roofline = dict(bound="valu", kernel=dom, achieved=kd["tflops"], peak=FP64_PEAK / 1e12, unit="TFLOP/s",
                    frac=(kd["tflops"] / (FP64_PEAK / 1e12)) if kd["tflops"] else None,
                    traffic=rp["traffic"] if rp else None,
                    pairs_per_s=kd["pairs_per_s"], avg_launch_ms=kd["avg_ms"],
                    flops_per_pair=kd["flops_per_pair"],
                    valu_busy_rocprof=rp["valu_busy"] if rp else None,
                    valu_issue_rocprof=rp["valu_issue"] if rp else None,
                    note="VALU-issue bound: exp/sqrt count zero flops and FP64 exp/sqrt are instruction "
                         "sequences, so algorithmic TFLOP/s sits below the issue-slot figures; the "
                         "rocprof fields come from profiles/pair_kernels_sphere_cube.md")

    out = {
        "metric": "kernel pair-interactions/s (objective+gradient evaluation, FP64)",
        "value": value,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded icosphere -> cube, seeded momentum)",
        "config": {"workload": args.config + (": icosphere-5 (N=10242, Ed=20480) -> cube 58x58/face (Et=40368), "
                                              "Matern nu=3/2 w=0.3, varifold G0.15xBinet, T-1=20"
                                              if args.config == "sphere_cube" else ""),
                   "pairs_per_eval": pairs, "evals_per_s": 1e3 / ms_step},
        "roofline": roofline,
        "kernels": kernels,
    }

    # FP32 policy on the same evaluation: rate and error against FP64
    f64 = d_f.get()[0]
    g64 = d_grd.get()
    obj.set_precision("fp32")
    for _ in range(args.warmup):
        obj.eval_device(d_f.ptr, d_grd.ptr, d_alp.ptr)
    obj.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        obj.eval_device(d_f.ptr, d_grd.ptr, d_alp.ptr)
    obj.synchronize()
    el32 = time.perf_counter() - t0
    f32, g32 = d_f.get()[0], d_grd.get()
    # kernel breakdown of the FP32 policy (separate, event-timed pass)
    lddmm_amd.profile_reset()
    lddmm_amd.profile_enable(True)
    for _ in range(max(2, args.steps // 4)):
        obj.eval_device(d_f.ptr, d_grd.ptr, d_alp.ptr)
    obj.synchronize()
    lddmm_amd.profile_enable(False)
    p32 = lddmm_amd.profile_read()
    # pairs the culling kernels compute (the rest are skipped below 2^-48):
    # one counted evaluation; rates and roofline use computed pairs only
    lddmm_amd.pair_count_begin()
    obj.eval_device(d_f.ptr, d_grd.ptr, d_alp.ptr)
    counted = lddmm_amd.pair_count_end()
    launches = {"fwd_pairs": T - 1, "bwd_pairs": T - 1, "vf_pairs": 1}
    k32 = {}
    for k, (np_, _) in per_launch.items():
        ms, cnt = p32.get(k, (0.0, 0))
        if cnt:
            ex = counted[k] / launches[k] if counted[k] else np_   # not culled: every pair computed
            k32[k] = dict(avg_ms=ms / cnt, pairs_per_s=np_ / (ms / cnt * 1e-3), executed_pairs_per_launch=ex,
                          executed_fraction=ex / np_, executed_pairs_per_s=ex / (ms / cnt * 1e-3))
    dom32 = max(k32, key=lambda k: k32[k]["avg_ms"] * p32[k][1]) if k32 else None
    rp32 = ROCPROF_SPHERE_CUBE.get((dom32, "f32")) if args.config == "sphere_cube" else None
    roof32 = None
    if dom32 and per_launch[dom32][1]:
        t32 = k32[dom32]["executed_pairs_per_launch"] * per_launch[dom32][1] / (k32[dom32]["avg_ms"] * 1e-3) / 1e12
        roof32 = dict(bound="valu", kernel=dom32, achieved=t32, peak=FP64_PEAK / 1e12, unit="TFLOP/s",
                      frac=t32 / (FP64_PEAK / 1e12), traffic=rp32["traffic"] if rp32 else None,
                      valu_busy_rocprof=rp32["valu_busy"] if rp32 else None,
                      valu_issue_rocprof=rp32["valu_issue"] if rp32 else None,
                      frac_vs_packed_fp32_peak=t32 / 157.3,
                      note="computed pairs only (executed_pairs_per_launch); FP32 VALU peak without packed "
                           "FP32 = 78.6 TFLOP/s (this build compiles all HIP code without packed FP32 "
                           "instructions; 157.3 TFLOP/s needs them)")
    # transcendental utilisation of the FP32 pair kernels: computed pairs x
    # exp/sqrt per pair against the measured v_exp_f32 rate of this chip
    rates = lddmm_amd.measure_rates()
    out["measured_rates"] = dict(rates, note="lddmm_measure_rates: v_exp_f32 and FMA throughput of 4,096 blocks x "
                                            "256 threads x 8 independent chains (rates.hip)")
    for k, kd32 in k32.items():
        tp = TRANS_PER_PAIR.get((k, order if k != "vf_pairs" else vfkey))
        if tp:
            kd32["trans_per_pair"] = tp
            kd32["trans_utilisation"] = kd32["executed_pairs_per_s"] * tp / rates["exp2_f32"]
    if roof32:
        roof32["valu_peak_measured_tflops"] = 2 * rates["fma_f32"] / 1e12
        roof32["frac_vs_measured_peak"] = roof32["achieved"] / (2 * rates["fma_f32"] / 1e12)
    roofline["valu_peak_measured_tflops"] = 2 * rates["fma_f64"] / 1e12
    roofline["frac_vs_measured_peak"] = (roofline["achieved"] or 0.0) / (2 * rates["fma_f64"] / 1e12)
    out["fp32_policy"] = dict(value=pairs * args.steps / el32, ms_per_step=1e3 * el32 / args.steps,
                              speedup_vs_fp64=elapsed / el32, energy_rel_err=abs(f32 - f64) / abs(f64),
                              grad_rel_err=float(np.max(np.abs(g32 - g64)) / np.max(np.abs(g64))), kernels=k32,
                              roofline=roof32)
    obj.set_precision("fp64")

    if args.registration:
        # Protocol of the registrations/s metric: the plain FP64 run (the
        # config's iteration count) sets E_ref; the FP32 policy runs until
        # its energy is within 1e-6 of E_ref (finishing in FP64 if its FP32
        # phase stalls), and its result is re-evaluated in FP64.
        itr = args.reg_iters or extra["itr"]
        common = (w["q0"], None, w["elm"], ifo, w["tgt_cen"], w["tgt_vol"], w["tgt_dir"], w["cen_type"],
                  w["cen_width"], w["dir_type"], w["dir_width"], w["dis_wgt"], w["knl_order"], w["knl_width"],
                  w["time_stp"], T)
        regs = {}
        tol, mem = extra.get("tol", 0.0), extra.get("mem", 5)
        a64, i64 = lddmm_amd.run(*common, itr, tol, 0.0, 0.5, mem, precision="fp64")
        e_ref = i64["fcn"]
        a32, i32 = lddmm_amd.run(*common, 3 * itr, 0.0, 0.0, 0.5, mem, precision="fp32",
                                 target_energy=e_ref * (1.0 + 1e-6))
        e32, _ = obj(a32)   # FP32-policy result re-evaluated in FP64
        for k, info in (("fp64", i64), ("fp32", i32)):
            regs[k] = dict(iterations=info["iterations"], evaluations=info["evaluations"], status=info["status"],
                           final_energy=info["fcn"], grad_norm=info["grad_norm"], seconds=info["seconds"],
                           registrations_per_s=1.0 / info["seconds"])
        regs["fp32"].update(fp32_iterations=i32["fp32_iterations"], reached_target=i32["reached_target"],
                            final_energy_fp64_eval=e32, energy_rel_to_fp64_run=(e32 - e_ref) / abs(e_ref))
        regs["fp32_speedup"] = i64["seconds"] / i32["seconds"]
        out["registration"] = regs

    if args.batch > 0:
        out["batch"] = batch_bench(args.batch, args.batch_nmin, args.batch_nmax, 300, warm=not args.no_warm_start,
                                   batched_fp64=args.batched_fp64, budget_s=args.budget, t_start=T_START)

    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, ifo, pairs)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
