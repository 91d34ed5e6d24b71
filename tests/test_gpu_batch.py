"""Batched uneven registrations on the GPU: every problem of a batch must
give what the single-problem path (itself pinned to the oracle) gives."""
import numpy as np
import pytest

import oracle
from conftest import rel_err
from lddmm_amd import geometry as G

pytestmark = pytest.mark.gpu


def small_batch():
    probs = G.batch_problems(5, seed=3, nmin=12, nmax=700)
    probs.append(G.batch_problems(1, seed=9, nmin=2500, nmax=2600)[0])
    return probs


def kernels(T=6):
    k = dict(G.BATCH_KERNELS)
    k["time_num"], k["time_stp"] = T, 1.0 / (T - 1)
    return k


def single(gpu, p, k, precision="fp64"):
    ifo = oracle.get_elm_ifo(p["elm"], len(p["q0"]))
    return gpu.Objective(p["q0"], p["elm"], ifo, p["tgt_cen"], p["tgt_vol"], p["tgt_dir"], k["cen_type"],
                         k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"], k["knl_order"], p["knl_width"],
                         k["time_stp"], k["time_num"], precision=precision), ifo


@pytest.mark.parametrize("chunk", [0, 64, 512])
def test_batched_objgrd_matches_single_problems(gpu, chunk):
    probs, k = small_batch(), kernels()
    print("batch objgrd chunk=%d sizes=%s" % (chunk, [len(p["q0"]) for p in probs]), flush=True)
    B = gpu.Batch(probs, k["cen_type"], k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"], k["knl_order"],
                  k["time_stp"], k["time_num"], chunk=chunk)
    ch, pairs = B.info()
    assert ch in ((256, 512, 1024, 2048, 4096) if chunk == 0 else (chunk,))
    alps = [0.05 * G.normal(40 + i, B.sizes[i]).reshape(k["time_num"] - 1, B.n[i], 3) for i in range(B.P)]
    f, g = B.objgrd(alps)
    for i, p in enumerate(probs):
        obj, _ = single(gpu, p, k)
        fs, gs = obj(alps[i])
        assert abs(f[i] - fs) <= 1e-12 * abs(fs), (i, f[i], fs)
        assert rel_err(g[i].reshape(-1), gs) <= 1e-11, (i, rel_err(g[i].reshape(-1), gs))
        assert pairs[i] == obj.pairs_per_eval()
        obj.close()
    B.close()


def test_batched_fp32_objgrd_error(gpu):
    probs, k = small_batch(), kernels()
    B = gpu.Batch(probs, k["cen_type"], k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"], k["knl_order"],
                  k["time_stp"], k["time_num"], precision="fp32", chunk=256)
    alps = [0.05 * G.normal(60 + i, B.sizes[i]).reshape(k["time_num"] - 1, B.n[i], 3) for i in range(B.P)]
    f, g = B.objgrd(alps)
    for i, p in enumerate(probs):
        obj, _ = single(gpu, p, k)
        fs, gs = obj(alps[i])
        e_rel, g_rel = abs(f[i] - fs) / abs(fs), rel_err(g[i].reshape(-1), gs)
        print("batched fp32 problem %d (N=%d): energy %.2e gradient %.2e" % (i, B.n[i], e_rel, g_rel), flush=True)
        assert e_rel <= 1e-6 and g_rel <= 1e-4
        obj.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_every_launch_shape_gives_the_same_objgrd(gpu, precision):
    """The autotuned launch shape (source chunk, FP32 targets per lane of the
    K.alpha and adjoint launches) is one of the candidates, and every shape
    gives the same energies and gradients.  In FP64 the chunk changes only
    the order in which split sums are added.  In FP32 the targets per lane
    also set the target tile, whose centre anchors the FP32 offsets and
    whose bounds decide the grouped culling, so shapes differ at FP32
    rounding: each stays within the FP32 bounds of the FP64 values."""
    probs, k = small_batch(), kernels()
    B = gpu.Batch(probs, k["cen_type"], k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"], k["knl_order"],
                  k["time_stp"], k["time_num"], precision=precision, chunk=0)
    ch, kv, bw = B.tuning()
    print("%s autotuned shape: chunk %d, K.alpha lanes %d, adjoint lanes %d" % (precision, ch, kv, bw), flush=True)
    assert ch in (256, 512, 1024, 2048, 4096) and kv in (1, 2, 4) and bw in (1, 2, 4)
    alps = [0.05 * G.normal(70 + i, B.sizes[i]).reshape(k["time_num"] - 1, B.n[i], 3) for i in range(B.P)]
    f0, g0 = B.objgrd(alps)
    ref = []
    if precision == "fp32":
        for p, a in zip(probs, alps):
            obj, _ = single(gpu, p, k)
            ref.append(obj(a))
            obj.close()
    # adjoint lanes 4: the matrix-core adjoint of 3D Gaussian batches (FP32),
    # with its vector fallback for the tiles of the small problems
    for shape in [(256, 1, 1), (1024, 2, 2), (512, 4, 1), (2048, 4, 2), (1024, 2, 4), (4096, 1, 4), (64, 1, 2)]:
        if shape[0] < 256:
            with pytest.raises(gpu.LddmmError):   # the partial buffer holds chunks down to 256
                B.set_tuning(*shape)
            continue
        B.set_tuning(*shape)
        assert B.tuning() == shape
        f, g = B.objgrd(alps)
        for i in range(B.P):
            e_rel, g_rel = abs(f[i] - f0[i]) / abs(f0[i]), rel_err(g[i].reshape(-1), g0[i].reshape(-1))
            if precision == "fp64":
                assert e_rel <= 1e-11 and g_rel <= 1e-10, (shape, i, e_rel, g_rel)
            else:
                assert e_rel <= 1e-7 and g_rel <= 1e-5, (shape, i, e_rel, g_rel)
                fd, gd = ref[i]
                assert abs(f[i] - fd) <= 1e-6 * abs(fd) and rel_err(g[i].reshape(-1), gd) <= 1e-4, (shape, i)
    for bad in [(256, 3, 1), (256, 2, 3), (0, 1, 1)]:
        with pytest.raises(gpu.LddmmError):
            B.set_tuning(*bad)
    B.close()


@pytest.mark.parametrize("mem,itr", [(5, 8), (1, 6), (12, 16)])
def test_batched_run_matches_single_runs(gpu, mem, itr):
    """The batched L-BFGS (compact form of the two-loop recursion, dot
    products in groups of 16 per launch) against the single-problem two-loop
    runs: memory 1, the default 5, and 12 with the ring wrapped (2m = 24 and
    3m = 36 dot products per pass)."""
    probs, k = small_batch(), kernels()
    B = gpu.Batch(probs, k["cen_type"], k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"], k["knl_order"],
                  k["time_stp"], k["time_num"], chunk=256)
    alps, infos, rounds = B.run(itr, 0.0, 0.0, 0.5, mem)
    print("batched run, memory %d: %d rounds, iterations %s" % (mem, rounds, [i["iterations"] for i in infos]),
          flush=True)
    for i, p in enumerate(probs):
        obj, ifo = single(gpu, p, k)
        a1, i1 = gpu.run(p["q0"], None, p["elm"], ifo, p["tgt_cen"], p["tgt_vol"], p["tgt_dir"], k["cen_type"],
                         k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"], k["knl_order"], p["knl_width"],
                         k["time_stp"], k["time_num"], itr, 0.0, 0.0, 0.5, mem)
        assert infos[i]["status"] == i1["status"] and infos[i]["iterations"] == i1["iterations"]
        assert abs(infos[i]["fcn"] - i1["fcn"]) <= 1e-8 * abs(i1["fcn"]), (i, infos[i]["fcn"], i1["fcn"])
        fb, _ = obj(alps[i])
        assert abs(fb - infos[i]["fcn"]) <= 1e-12 * abs(fb)
        obj.close()


def test_batched_run_relative_decrease_stop_and_fp32(gpu):
    probs, k = small_batch(), kernels()
    B = gpu.Batch(probs, k["cen_type"], k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"], k["knl_order"],
                  k["time_stp"], k["time_num"], chunk=256)
    a64, i64, r64 = B.run(300, 0.0, 0.0, 0.5, 5, rel_energy_tol=1e-6)
    a32, i32, r32 = B.run(300, 0.0, 0.0, 0.5, 5, precision="fp32", rel_energy_tol=1e-6)
    print("fp64 rounds %d its %s; fp32 rounds %d its %s fp32-its %s" %
          (r64, [i["iterations"] for i in i64], r32, [i["iterations"] for i in i32],
           [i["fp32_iterations"] for i in i32]), flush=True)
    f64, _ = B.objgrd(a64)
    f32, _ = B.objgrd(a32)
    stopped = 0
    for i in range(B.P):
        assert i64[i]["status"] == 0 and i32[i]["status"] == 0
        assert i64[i]["iterations"] <= 300 and i32[i]["iterations"] <= 300
        assert abs(f64[i] - i64[i]["fcn"]) <= 1e-12 * abs(f64[i])
        if i64[i]["iterations"] < 300 and i32[i]["iterations"] < 300:
            # both stopped on a relative decrease below 1e-6: energies agree to the scale of that stop
            stopped += 1
            assert abs(f32[i] - f64[i]) <= 2e-3 * abs(f64[i]), (i, f32[i], f64[i])
    assert stopped >= 1


def test_batched_fp32_reaches_fp64_targets_confirmed_in_fp64(gpu):
    """Each problem's FP32-policy run stops only once its energy, evaluated in
    FP64, is at or below the FP64 run's final energy x (1 + 1e-6)."""
    probs, k = small_batch(), kernels()
    B = gpu.Batch(probs, k["cen_type"], k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"], k["knl_order"],
                  k["time_stp"], k["time_num"], chunk=256)
    _, i64, _ = B.run(60, 0.0, 0.0, 0.5, 5)
    targets = np.array([r["fcn"] for r in i64]) * (1 + 1e-6)
    a32, i32, r32 = B.run(600, 0.0, 0.0, 0.5, 5, precision="fp32", targets=targets)
    f, _ = B.objgrd(a32)
    print("fp32 to fp64 targets: rounds %d, iterations %s, fp32 iterations %s" %
          (r32, [r["iterations"] for r in i32], [r["fp32_iterations"] for r in i32]), flush=True)
    for i in range(B.P):
        assert i32[i]["reached_target"], i
        assert f[i] <= targets[i] * (1 + 1e-13), (i, f[i], targets[i])
        assert abs(f[i] - i32[i]["fcn"]) <= 1e-12 * abs(f[i])
    B.close()


def test_graph_replays_match_direct_launches(gpu):
    """Evaluations replayed from captured hipGraphs equal the directly
    launched ones bitwise, replay after replay, with eager work and
    synchronisation in between and with the work tables rebuilt."""
    probs, k = small_batch(), kernels()
    B = gpu.Batch(probs, k["cen_type"], k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"], k["knl_order"],
                  k["time_stp"], k["time_num"], chunk=256)
    alps = [0.05 * G.normal(80 + i, B.sizes[i]).reshape(k["time_num"] - 1, B.n[i], 3) for i in range(B.P)]
    other = [0.5 * a for a in alps]
    B.set_graphs(False)
    f_ref, g_ref = B.objgrd(alps)
    f_oth, g_oth = B.objgrd(other)
    B.set_graphs(True)
    for rep in range(4):
        f, g = B.objgrd(alps if rep % 2 == 0 else other)
        fr, gr = (f_ref, g_ref) if rep % 2 == 0 else (f_oth, g_oth)
        assert np.array_equal(f, fr), rep
        assert all(np.array_equal(a, b) for a, b in zip(g, gr)), rep
        # eager work between replays: a single-problem evaluation on its own stream
        obj, _ = single(gpu, probs[rep % len(probs)], k)
        obj(alps[rep % len(probs)])
        obj.close()
    # a run (which rebuilds the tables as problems finish) is unchanged by graphs
    B.set_graphs(False)
    a1, i1, r1 = B.run(12, 0.0, 0.0, 0.5, 5, precision="fp32")
    B.set_graphs(True)
    a2, i2, r2 = B.run(12, 0.0, 0.0, 0.5, 5, precision="fp32")
    print("graph replays: run rounds %d / %d" % (r1, r2), flush=True)
    assert r1 == r2
    for x, y, p, q in zip(a1, a2, i1, i2):
        assert np.array_equal(x, y)
        assert p["fcn"] == q["fcn"] and p["iterations"] == q["iterations"]
    B.close()


def oracle_problem(p, k):
    ifo = oracle.get_elm_ifo(p["elm"], len(p["q0"]))
    return oracle.OracleProblem(p["q0"], p["elm"], p["tgt_cen"], p["tgt_dir"], p["tgt_vol"], k["cen_type"],
                                k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"], k["knl_order"],
                                p["knl_width"], k["time_stp"], k["time_num"], ifo=ifo)


def test_batched_objgrd_matches_oracle(gpu):
    """Every problem of a batch (sizes 12 to 2,600 in one launch per kernel)
    against the CPU oracle's objgrd (Cpp/src/objgrd.cpp:26-90) directly:
    energy and gradient within 1e-10, at zero and at seeded momenta."""
    probs, k = small_batch(), kernels()
    B = gpu.Batch(probs, k["cen_type"], k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"], k["knl_order"],
                  k["time_stp"], k["time_num"])
    for scale in (0.0, 0.05):
        alps = [scale * G.normal(140 + i, B.sizes[i]).reshape(k["time_num"] - 1, B.n[i], 3) for i in range(B.P)]
        f, g = B.objgrd(alps)
        for i, p in enumerate(probs):
            ef, eg = oracle_problem(p, k).objgrd(alps[i].reshape(-1))
            print("batch vs oracle: scale %.2f problem %d (N=%d): energy %.2e gradient %.2e" %
                  (scale, i, B.n[i], abs(f[i] - ef) / abs(ef), rel_err(g[i].reshape(-1), eg)), flush=True)
            assert abs(f[i] - ef) <= 1e-10 * abs(ef), (scale, i)
            assert rel_err(g[i].reshape(-1), eg) <= 1e-10, (scale, i)
    B.close()


@pytest.mark.parametrize("precision,shape", [("fp64", None), ("fp32", (1024, 2, 4))])
def test_objective_scan_and_scaled_start(gpu, precision, shape):
    """Momenta constant in time, built on the device from one (N, 3) array
    per problem: lddmm_batch_objective_scan's energies (no adjoint sweep) at
    multiples 0, 0.5, 1, 2 equal lddmm_batch_objgrd's at the same stacks, and
    in FP64 the oracle's objgrd (Cpp/src/objgrd.cpp:26-90) within 1e-10 for the
    problems of up to 700 points; lddmm_batch_run_scaled runs exactly as
    lddmm_batch_run from those stacks (a zero multiple is the zero start)."""
    probs, k = small_batch(), kernels()
    T1 = k["time_num"] - 1
    B = gpu.Batch(probs, k["cen_type"], k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"], k["knl_order"],
                  k["time_stp"], k["time_num"], precision=precision, chunk=256)
    if shape is not None:
        B.set_tuning(*shape)
    base = [0.05 * G.normal(180 + i, 3 * B.n[i]).reshape(B.n[i], 3) for i in range(B.P)]
    scales = [0.0, 0.5, 1.0, 2.0]
    fs = B.objective_scan(base, scales)
    print("objective scan (%s, shape %s): %s" % (precision, shape, np.round(fs[:, :3], 6).tolist()), flush=True)
    for j, c in enumerate(scales):
        stacks = [np.repeat((c * u)[None], T1, axis=0) for u in base]
        f, _ = B.objgrd(stacks)
        assert np.all(np.abs(fs[j] - f) <= 1e-13 * np.abs(f)), (c, fs[j], f)
        if precision == "fp64" and c in (0.0, 1.0):
            for i, p in enumerate(probs):
                if B.n[i] <= 700:
                    ef, _ = oracle_problem(p, k).objgrd(stacks[i].reshape(-1))
                    assert abs(fs[j][i] - ef) <= 1e-10 * abs(ef), (c, i)
    start = np.array([0.0, 1.0, 0.5, 2.0, 1.0, 0.5])
    a1, i1, r1 = B.run_scaled(base, start, 6, 0.0, 0.0, 0.5, 5, precision=precision)
    a2, i2, r2 = B.run(6, 0.0, 0.0, 0.5, 5, alp0=[np.repeat((c * u)[None], T1, axis=0) for c, u in zip(start, base)],
                       precision=precision)
    assert r1 == r2
    for i in range(B.P):
        assert (i1[i]["status"], i1[i]["iterations"], i1[i]["evaluations"]) == \
            (i2[i]["status"], i2[i]["iterations"], i2[i]["evaluations"]), i
        assert i1[i]["fcn"] == i2[i]["fcn"] and np.array_equal(a1[i], a2[i]), i
    with pytest.raises(ValueError):
        B.run_scaled(base, start[:-1], 2, 0.0, 0.0, 0.5, 5)
    B.close()


@pytest.mark.parametrize("precision,shape,bad_p", [("fp64", None, 2), ("fp32", None, 2), ("fp32", (1024, 2, 4), 5)])
def test_nan_in_one_problem_leaves_the_others_unchanged(gpu, precision, shape, bad_p):
    """A NaN in one problem's initial momentum: its objective and gradient
    are NaN, so its first direction has no negative slope and it stops with
    status 1 (not a descent direction, Cpp/src/lineSearch.cpp:39-46) at its
    start; every other problem runs exactly as in a batch without the NaN.
    The third case puts the NaN into the 2,562-point problem under the
    matrix-core adjoint (adjoint lanes 4), whose column scales come from
    per-group maxima that skip NaN."""
    probs, k = small_batch(), kernels()
    B = gpu.Batch(probs, k["cen_type"], k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"], k["knl_order"],
                  k["time_stp"], k["time_num"], chunk=256)
    if shape is not None:
        B.set_tuning(*shape)
    zero = [np.zeros((k["time_num"] - 1, B.n[i], 3)) for i in range(B.P)]
    a_ref, i_ref, _ = B.run(8, 0.0, 0.0, 0.5, 5, alp0=zero, precision=precision)
    bad = [z.copy() for z in zero]
    bad[bad_p][1, 3, 0] = np.nan
    a_nan, i_nan, _ = B.run(8, 0.0, 0.0, 0.5, 5, alp0=bad, precision=precision)
    print("NaN problem %d (%s, shape %s): status %d, iterations %d" %
          (bad_p, precision, shape, i_nan[bad_p]["status"], i_nan[bad_p]["iterations"]), flush=True)
    assert i_nan[bad_p]["status"] == 1 and i_nan[bad_p]["iterations"] == 0
    assert np.isnan(a_nan[bad_p][1, 3, 0]) and np.count_nonzero(np.isnan(a_nan[bad_p])) == 1   # the start is returned
    for i in range(B.P):
        if i == bad_p:
            continue
        assert i_nan[i]["status"] == i_ref[i]["status"] == 0
        assert i_nan[i]["iterations"] == i_ref[i]["iterations"] and i_nan[i]["fcn"] == i_ref[i]["fcn"]
        assert np.array_equal(a_nan[i], a_ref[i])
    B.close()


def test_batched_forced_line_search_failure_matches_oracle(gpu):
    """Wolfe constants (0.9, 0.1) shared by the batch: on a locally quadratic
    objective no step meets both, so a line search brackets and bisects to
    bisMax and returns code 3 (Cpp/src/lineSearch.cpp:52-57,104-109), at the
    first iteration or after a few accepted ones.  Every problem gives the
    status, iterations and evaluations of its single-problem run, and of the
    oracle's run for the problems of up to 700 points."""
    probs, k = small_batch(), kernels()
    B = gpu.Batch(probs, k["cen_type"], k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"], k["knl_order"],
                  k["time_stp"], k["time_num"], chunk=256)
    alps, infos, rounds = B.run(5, 0.0, 0.9, 0.1, 5)
    print("forced failure: statuses %s, iterations %s, evaluations %s, %d rounds" %
          ([r["status"] for r in infos], [r["iterations"] for r in infos], [r["evaluations"] for r in infos],
           rounds), flush=True)
    for i, p in enumerate(probs):
        ifo = oracle.get_elm_ifo(p["elm"], len(p["q0"]))
        _, i1 = gpu.run(p["q0"], None, p["elm"], ifo, p["tgt_cen"], p["tgt_vol"], p["tgt_dir"], k["cen_type"],
                        k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"], k["knl_order"], p["knl_width"],
                        k["time_stp"], k["time_num"], 5, 0.0, 0.9, 0.1, 5)
        assert infos[i]["status"] == i1["status"] == 3
        assert (infos[i]["iterations"], infos[i]["evaluations"]) == (i1["iterations"], i1["evaluations"])
        if B.n[i] <= 700:
            _, _, tr = oracle_problem(p, k).run(itr_max=5, tol=0.0, wolfe1=0.9, wolfe2=0.1, mem=5)
            assert tr["status"] == 3 and (infos[i]["iterations"], infos[i]["evaluations"]) == \
                (tr["iters"], tr["total_evals"]), (i, tr["iters"], tr["total_evals"])
    B.close()


def test_fp64_confirmation_does_not_exceed_the_iteration_cap(gpu):
    """With a relative-decrease threshold of 1 every FP32 iteration hands
    over to FP64 (the confirmation step of the FP32 policy); at the cap
    itrMax = 1 the problems stop there instead of running a second
    iteration, in the batch as in the single-problem driver."""
    probs, k = small_batch(), kernels()
    B = gpu.Batch(probs, k["cen_type"], k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"], k["knl_order"],
                  k["time_stp"], k["time_num"], chunk=256)
    for itr in (1, 2):
        _, infos, _ = B.run(itr, 0.0, 0.0, 0.5, 5, precision="fp32", refine_rel_tol=1.0)
        print("cap %d: iterations %s" % (itr, [r["iterations"] for r in infos]), flush=True)
        assert all(r["iterations"] <= itr for r in infos)
        p = probs[0]
        ifo = oracle.get_elm_ifo(p["elm"], len(p["q0"]))
        _, i1 = gpu.run(p["q0"], None, p["elm"], ifo, p["tgt_cen"], p["tgt_vol"], p["tgt_dir"], k["cen_type"],
                        k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"], k["knl_order"], p["knl_width"],
                        k["time_stp"], k["time_num"], itr, 0.0, 0.0, 0.5, 5, precision="fp32", refine_rel_tol=1.0)
        assert i1["iterations"] <= itr
    B.close()
