"""GPU L-BFGS registration against the oracle's L-BFGS (Cpp/src/LBFGS.cpp,
getDirection.cpp, lineSearch.cpp restated in oracle/lddmm_oracle.c).

Single evaluations agree to ~1e-13, but a line search amplifies rounding
differences over many iterations, so traces are compared iteration by
iteration with a tolerance that grows along the run, and final energies
of full runs within 1e-6 relative."""
import numpy as np
import pytest

import oracle
from lddmm_amd import geometry as G

pytestmark = pytest.mark.gpu


def circle_square_args(n=400, T=21):
    w = G.circle_square(n)
    w["time_num"] = T
    w["time_stp"] = 1.0 / (T - 1)
    return w


def run_both(gpu, w, itr, tol=0.0):
    n = w["q0"].shape[0]
    ifo = oracle.get_elm_ifo(w["elm"], n)
    alp, info = gpu.run(w["q0"], None, w["elm"], ifo, w["tgt_cen"], w["tgt_vol"], w["tgt_dir"], w["cen_type"],
                        w["cen_width"], w["dir_type"], w["dir_width"], w["dis_wgt"], w["knl_order"], w["knl_width"],
                        w["time_stp"], w["time_num"], itr, tol, 0.0, 0.5, 5)
    P = oracle.OracleProblem(w["q0"], w["elm"], w["tgt_cen"], w["tgt_dir"], w["tgt_vol"], w["cen_type"],
                             w["cen_width"], w["dir_type"], w["dir_width"], w["dis_wgt"], w["knl_order"],
                             w["knl_width"], w["time_stp"], w["time_num"], ifo=ifo)
    ealp, ef, tr = P.run(itr_max=itr, tol=tol, wolfe1=0.0, wolfe2=0.5, mem=5)
    return alp, info, ealp, ef, tr, P


def test_circle_to_square_trace_matches_oracle(gpu):
    print("lbfgs circle_square 400 points, 25 iterations", flush=True)
    w = circle_square_args()
    alp, info, ealp, ef, tr, P = run_both(gpu, w, 25)
    assert info["status"] == 0 and tr["status"] == 0
    assert info["iterations"] == tr["iters"] == 25
    # energies at the first iterations agree to rounding level, and stay close
    f_gpu, _ = P.objgrd(alp.reshape(-1) if alp.ndim == 1 else alp)
    assert abs(f_gpu - info["fcn"]) <= 1e-10 * abs(f_gpu)
    assert abs(info["fcn"] - ef) <= 1e-6 * abs(ef), (info["fcn"], ef)


def test_short_run_is_identical_to_oracle_run(gpu):
    print("lbfgs circle_square 100 points, 6 iterations", flush=True)
    w = circle_square_args(100, T=6)
    alp, info, ealp, ef, tr, P = run_both(gpu, w, 6)
    assert info["evaluations"] == tr["total_evals"]
    assert abs(info["fcn"] - ef) <= 1e-9 * abs(ef)
    assert np.max(np.abs(alp.reshape(-1) - ealp)) <= 1e-7 * np.max(np.abs(ealp))


@pytest.mark.parametrize("kind", ["sphere_cube_matern", "circle_square_matern2d"])
def test_short_matern_runs_are_identical_to_oracle_runs(gpu, kind):
    """6-iteration registrations with the C2 kernel family (3D Matern
    nu = 3/2, 'G' x 'B' varifold) and with a 2D Matern kernel: the same
    evaluations, energies and momenta as the oracle's L-BFGS."""
    if kind == "sphere_cube_matern":
        w = G.sphere_cube(subdivisions=2, cells=8)
    else:
        w = circle_square_args(120, T=6)
        w["knl_order"] = 2
    w["time_num"], w["time_stp"] = 6, 0.2
    print("lbfgs %s, 6 iterations" % kind, flush=True)
    alp, info, ealp, ef, tr, P = run_both(gpu, w, 6)
    assert info["status"] == 0 and tr["status"] == 0
    assert info["evaluations"] == tr["total_evals"]
    assert abs(info["fcn"] - ef) <= 1e-9 * abs(ef)
    assert np.max(np.abs(alp.reshape(-1) - ealp)) <= 1e-7 * np.max(np.abs(ealp))


def test_tolerance_stop_and_line_search_status(gpu):
    print("lbfgs tolerance stop", flush=True)
    w = circle_square_args(100, T=6)
    alp, info, *_ = run_both(gpu, w, 500, tol=1e-1)
    assert info["status"] in (0, 1, 2, 3)
    assert info["grad_norm"] < 1e-1 or info["status"] != 0 or info["iterations"] == 500


def test_mex_style_entry_points_agree(gpu):
    """lddmm_run_3d (host layout) and lddmm_run_dev (device layout) give the
    same bits on the same problem."""
    import ctypes as C
    L = gpu.lib()
    w = G.sphere_cube(subdivisions=1, cells=4)
    w["time_num"], w["time_stp"] = 4, 1.0 / 3
    q0, elm = w["q0"], w["elm"].astype(np.int32)
    n, d = q0.shape
    ifo = oracle.get_elm_ifo(elm, n)
    T = w["time_num"]
    nv = d * n * (T - 1)
    a0 = np.zeros(nv)
    out_h = np.empty(nv)
    tc, tu, ta = w["tgt_cen"], w["tgt_dir"], w["tgt_vol"]
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    rc = L.lddmm_run_3d(dp(q0), dp(a0), ip(elm), ip(ifo), dp(tc), dp(ta), dp(tu), b"G", 0.15, b"B", 0.0, 100.0,
                        1, 0.3, w["time_stp"], T, 8, 0.0, 0.0, 0.5, 5, 0, n, elm.shape[0], ifo.shape[1], len(ta),
                        dp(out_h))
    assert rc == 0, L.lddmm_last_error()

    # device layout copies: N x D column-major = (D, N) C-order, etc.
    def dev(arr):
        arr = np.ascontiguousarray(arr)
        p = C.c_void_p()
        assert L.lddmm_device_malloc(C.byref(p), arr.nbytes) == 0
        assert L.lddmm_memcpy_htod(p, arr.ctypes.data, arr.nbytes) == 0
        return p

    bufs = [dev(q0.T), dev(np.zeros(nv)), dev(elm.T.astype(np.int32)), dev(ifo.T.astype(np.int32)), dev(tc.T),
            dev(ta), dev(tu.T)]
    out_d = C.c_void_p()
    assert L.lddmm_device_malloc(C.byref(out_d), nv * 8) == 0
    rc = L.lddmm_run_dev(3, *bufs, b"G", 0.15, b"B", 0.0, 100.0, 1, 0.3, w["time_stp"], T, 8, 0.0, 0.0, 0.5, 5, 0,
                         n, elm.shape[0], ifo.shape[1], len(ta), out_d, None, None, None)
    assert rc == 0, L.lddmm_last_error()
    res = np.empty(nv)
    assert L.lddmm_memcpy_dtoh(res.ctypes.data, out_d, nv * 8) == 0
    for b in bufs + [out_d]:
        L.lddmm_device_free(b)
    # device layout (T-1, D, N) -> host layout (T-1, N, D)
    res_h = res.reshape(T - 1, d, n).transpose(0, 2, 1).reshape(-1)
    assert np.array_equal(res_h, out_h)


def test_matlab_style_driver(gpu):
    """LDDMM() mirrors Cpp/src/LDDMM.m: 1-based elements, dict structs,
    returns the trajectory of the optimal momentum."""
    print("LDDMM() driver", flush=True)
    w = G.circle_square(80)
    disObj = dict(dfmElmVtxMat=w["elm"] + 1, tgtCenPosMat=w["tgt_cen"], tgtElmVolVec=w["tgt_vol"],
                  tgtUniDirMat=w["tgt_dir"], cenKnlType="G", cenKnlWidth=0.1, dirKnlType="B", dirKnlWidth=0.0,
                  disWgt=100.0)
    optObj = dict(MaxIterations=5, OptimalityTolerance=0.0, Wolfe1=0.0, Wolfe2=0.5, Columns=5, Verbose=True)
    nde, vlc, alp = gpu.LDDMM(w["q0"], None, disObj, -1, 0.25, 1.0, 0.05, optObj)
    assert nde.shape == (21, 80, 2) and vlc.shape == (20, 80, 2) and alp.shape == (20, 80, 2)
    assert np.array_equal(nde[0], w["q0"])
    assert np.allclose(nde[1], nde[0] + 0.05 * vlc[0], rtol=0, atol=1e-15)


def test_fp32_policy_reaches_fp64_energy(gpu):
    """The registration protocol of the metric: the plain FP64 run (200
    iterations) sets E_ref; the FP32 policy (finishing in FP64 if its FP32
    phase stalls) runs until f <= E_ref (1 + 1e-6); its result, re-evaluated
    in FP64, must be within 1e-6 of E_ref."""
    print("fp32 vs fp64 registration, circle_square 400", flush=True)
    w = circle_square_args()
    n = w["q0"].shape[0]
    ifo = oracle.get_elm_ifo(w["elm"], n)
    common = (w["q0"], None, w["elm"], ifo, w["tgt_cen"], w["tgt_vol"], w["tgt_dir"], w["cen_type"],
              w["cen_width"], w["dir_type"], w["dir_width"], w["dis_wgt"], w["knl_order"], w["knl_width"],
              w["time_stp"], w["time_num"])
    a64, i64 = gpu.run(*common, 200, 0.0, 0.0, 0.5, 5, precision="fp64")
    e_ref = i64["fcn"]
    a32, i32 = gpu.run(*common, 2000, 0.0, 0.0, 0.5, 5, precision="fp32", target_energy=e_ref * (1 + 1e-6))
    obj = gpu.Objective(w["q0"], w["elm"], ifo, w["tgt_cen"], w["tgt_vol"], w["tgt_dir"], w["cen_type"],
                        w["cen_width"], w["dir_type"], w["dir_width"], w["dis_wgt"], w["knl_order"], w["knl_width"],
                        w["time_stp"], w["time_num"])
    f32, _ = obj(a32)
    print("fp64: %d it E=%.12e (%.3f s); fp32: %d it (%d in FP32) E=%.12e (%.3f s), rel %.3e, reached %s" %
          (i64["iterations"], e_ref, i64["seconds"], i32["iterations"], i32["fp32_iterations"], f32,
           i32["seconds"], (f32 - e_ref) / abs(e_ref), i32["reached_target"]), flush=True)
    assert i32["reached_target"]
    # a target met in FP32 is confirmed by an FP64 evaluation before the run stops
    assert f32 <= e_ref * (1 + 1e-6) * (1 + 1e-13)


def test_fp32_phase_hands_over_to_fp64_on_stall(gpu):
    """Forced stall: a refine threshold of 1e-2 ends the FP32 phase at once;
    the run continues in FP64 and its energy keeps decreasing."""
    w = circle_square_args(100, T=6)
    n = w["q0"].shape[0]
    ifo = oracle.get_elm_ifo(w["elm"], n)
    alp, info = gpu.run(w["q0"], None, w["elm"], ifo, w["tgt_cen"], w["tgt_vol"], w["tgt_dir"], w["cen_type"],
                        w["cen_width"], w["dir_type"], w["dir_width"], w["dis_wgt"], w["knl_order"], w["knl_width"],
                        w["time_stp"], w["time_num"], 30, 0.0, 0.0, 0.5, 5, precision="fp32", refine_rel_tol=1e-2,
                        verbose=True)
    assert info["status"] == 0 and 1 <= info["fp32_iterations"] < info["iterations"] == 30


def test_circle_to_square_200_iterations_match_oracle_run(gpu):
    """Config 1 in full: 400 points, T = 21, memory 5, 200 iterations from
    zero momentum, against the oracle's run (tests/golden/oracle_runs.npz,
    written by tests/golden/make_oracle_runs.py; E = 0.42053...).

    This run is sensitive to rounding: the oracle itself, started from
    momenta of size 1e-16 instead of 0, takes different line-search steps
    from about iteration 30 and ends up to 0.8% away (c1_spread_f).  So the
    GPU run must follow the oracle's energies to 1e-6 through iteration 25,
    and end, with the same status and iteration count, within twice the
    largest relative distance of those rounding-perturbed oracle runs."""
    import os
    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "oracle_runs.npz"))
    print("lbfgs circle_square 400 points, 200 iterations vs oracle trace", flush=True)
    w = circle_square_args()
    n = w["q0"].shape[0]
    ifo = oracle.get_elm_ifo(w["elm"], n)
    alp, info = gpu.run(w["q0"], None, w["elm"], ifo, w["tgt_cen"], w["tgt_vol"], w["tgt_dir"], w["cen_type"],
                        w["cen_width"], w["dir_type"], w["dir_width"], w["dis_wgt"], w["knl_order"], w["knl_width"],
                        w["time_stp"], w["time_num"], 200, 0.0, 0.0, 0.5, 5)
    ef = float(ref["c1_f"][-1])
    spread = float(np.max(np.abs(ref["c1_spread_f"] - ef)) / ef)
    print("GPU E = %.15e, oracle E = %.15e, rel %.2e (oracle's own spread %.2e); evaluations %d vs %d" %
          (info["fcn"], ef, (info["fcn"] - ef) / ef, spread, info["evaluations"], int(np.sum(ref["c1_evals"]))),
          flush=True)
    assert info["status"] == int(ref["c1_status"]) == 0
    assert info["iterations"] == 200
    assert spread > 1e-4   # the fixture shows the run's sensitivity
    assert abs(info["fcn"] - ef) <= 2 * spread * abs(ef)
    obj = gpu.Objective(w["q0"], w["elm"], ifo, w["tgt_cen"], w["tgt_vol"], w["tgt_dir"], w["cen_type"],
                        w["cen_width"], w["dir_type"], w["dir_width"], w["dis_wgt"], w["knl_order"], w["knl_width"],
                        w["time_stp"], w["time_num"])
    f, _ = obj(alp)
    assert abs(f - info["fcn"]) <= 1e-12 * abs(f)
    obj.close()
    for k in (10, 25):
        _, ik = gpu.run(w["q0"], None, w["elm"], ifo, w["tgt_cen"], w["tgt_vol"], w["tgt_dir"], w["cen_type"],
                        w["cen_width"], w["dir_type"], w["dir_width"], w["dis_wgt"], w["knl_order"], w["knl_width"],
                        w["time_stp"], w["time_num"], k, 0.0, 0.0, 0.5, 5)
        assert abs(ik["fcn"] - ref["c1_f"][k]) <= 1e-6 * ref["c1_f"][k], (k, ik["fcn"], ref["c1_f"][k])
        assert ik["evaluations"] == int(np.sum(ref["c1_evals"][:k + 1]))


def test_forced_line_search_failure_matches_oracle(gpu):
    """Wolfe constants (0.9, 0.1): on a locally quadratic objective no step
    satisfies both conditions (Armijo needs t < 0.2 t*, the curvature test
    t > 0.9 t* along the line), so the first line search brackets and
    bisects until bisMax and returns code 3 (Cpp/src/lineSearch.cpp:104-109).
    The GPU run gives the oracle's status and evaluation count."""
    print("lbfgs forced line-search failure", flush=True)
    w = circle_square_args(100, T=6)
    n = w["q0"].shape[0]
    ifo = oracle.get_elm_ifo(w["elm"], n)
    alp, info = gpu.run(w["q0"], None, w["elm"], ifo, w["tgt_cen"], w["tgt_vol"], w["tgt_dir"], w["cen_type"],
                        w["cen_width"], w["dir_type"], w["dir_width"], w["dis_wgt"], w["knl_order"], w["knl_width"],
                        w["time_stp"], w["time_num"], 5, 0.0, 0.9, 0.1, 5)
    P = oracle.OracleProblem(w["q0"], w["elm"], w["tgt_cen"], w["tgt_dir"], w["tgt_vol"], w["cen_type"],
                             w["cen_width"], w["dir_type"], w["dir_width"], w["dis_wgt"], w["knl_order"],
                             w["knl_width"], w["time_stp"], w["time_num"], ifo=ifo)
    ealp, ef, tr = P.run(itr_max=5, tol=0.0, wolfe1=0.9, wolfe2=0.1, mem=5)
    print("GPU status %d, %d evaluations; oracle status %d, %d evaluations" %
          (info["status"], info["evaluations"], tr["status"], tr["total_evals"]), flush=True)
    assert tr["status"] == 3 and info["status"] == 3
    assert info["evaluations"] == tr["total_evals"] == 52   # the start and bisMax + 1 = 51 trials
    assert info["iterations"] == tr["iters"] == 0
    # the output is the last accepted iterate: the (zero) start
    assert not np.any(alp) and not np.any(ealp)


def test_reference_example_optimiser_settings_match_oracle(gpu):
    """The optimiser settings of the reference's examples
    (Cpp/examples/sphere2cube/run_sphere2cube.m:11-31: 'C' 0.1 x 'B'
    matching, Matern order 3, tolerance 1e-2, Wolfe (0, 0.5), a memory far
    above the iteration count) on a 120-point curve with memory 50, against
    the oracle's two-loop trace (tests/golden/oracle_runs.npz, ex_*).  This
    run is very sensitive to rounding: the oracle itself, started from momenta
    of size 1e-16, takes other line-search steps from iteration 13 on.  So
    the GPU run must take the oracle's evaluations per iteration with its
    energies to 1e-8 through iteration 10, and run its 60 iterations with
    status 0."""
    import os
    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "oracle_runs.npz"))
    w = circle_square_args(120, T=21)
    w["cen_type"], w["cen_width"], w["knl_order"], w["knl_width"] = "C", 0.1, 3, 0.25
    n = w["q0"].shape[0]
    ifo = oracle.get_elm_ifo(w["elm"], n)
    print("lbfgs memory 50, tolerance 1e-2, 60 iterations", flush=True)
    alp, info = gpu.run(w["q0"], None, w["elm"], ifo, w["tgt_cen"], w["tgt_vol"], w["tgt_dir"], w["cen_type"],
                        w["cen_width"], w["dir_type"], w["dir_width"], w["dis_wgt"], w["knl_order"], w["knl_width"],
                        w["time_stp"], w["time_num"], 60, 1e-2, 0.0, 0.5, 50)
    ef, evals = float(ref["ex_f"][-1]), ref["ex_evals"]
    print("GPU: %d iterations, %d evaluations, E = %.15e; oracle: %d, %d, E = %.15e" %
          (info["iterations"], info["evaluations"], info["fcn"], len(evals) - 1, int(np.sum(evals)), ef),
          flush=True)
    assert info["status"] == int(ref["ex_status"]) == 0
    assert info["iterations"] == len(evals) - 1 == 60
    gtr = gpu_trace_of(gpu, w, ifo, 10, 1e-2, 50)
    assert np.array_equal(gtr["evals"], evals[:11])
    assert np.max(np.abs(gtr["f"] - ref["ex_f"][:11]) / np.abs(ref["ex_f"][:11])) <= 1e-8


def test_memory_50_landmark_run_matches_oracle(gpu):
    """Memory 50 over 80 iterations (every direction from up to 50 pairs; the
    ring wraps at iteration 51) on a landmark problem (the l2Dist matching
    term, MATLAB/src/objgrd.m:29-34) whose L-BFGS run is insensitive to
    rounding: oracle runs from momenta of size 1e-16 stay within 2e-10 of it
    with the same evaluations.  The GPU's compact recursion gives the oracle's
    two-loop evaluation count and energy to 1e-8."""
    n, T = 60, 11
    q, _ = G.circle_curve(n)
    t = q + 0.05 * G.normal(3, q.size).reshape(q.shape)
    none = np.zeros((0, 2))
    P = oracle.OracleProblem(q, np.zeros((0, 2), dtype=np.int32), none, none, np.zeros(0), "G", 0.3, "B", 0.0,
                             50.0, -1, 0.3, 1.0 / (T - 1), T, ifo=np.zeros((n, 1), dtype=np.int32), tgt_nde=t)
    for itr in (30, 80):
        _, ef, tr = P.run(itr_max=itr, tol=0.0, wolfe1=0.0, wolfe2=0.5, mem=50)
        alp, info = gpu.run_l2(q, None, t, 50.0, -1, 0.3, 1.0 / (T - 1), T, itr, 0.0, 0.0, 0.5, 50)
        print("l2 memory 50, %d iterations: GPU E = %.15e (%d evaluations), oracle E = %.15e (%d)" %
              (itr, info["fcn"], info["evaluations"], ef, tr["total_evals"]), flush=True)
        assert info["status"] == tr["status"] == 0 and info["iterations"] == tr["iters"] == itr
        assert info["evaluations"] == tr["total_evals"]
        assert abs(info["fcn"] - ef) <= 1e-8 * abs(ef)


def gpu_trace_of(gpu, w, ifo, itr, tol, mem):
    """Energies and evaluation counts of the GPU run after 0..itr iterations
    (runs of 0, 1, ..., itr iterations; each one's count minus the previous)."""
    f, ev, total = [], [], 0
    for k in range(itr + 1):
        _, info = gpu.run(w["q0"], None, w["elm"], ifo, w["tgt_cen"], w["tgt_vol"], w["tgt_dir"], w["cen_type"],
                          w["cen_width"], w["dir_type"], w["dir_width"], w["dis_wgt"], w["knl_order"],
                          w["knl_width"], w["time_stp"], w["time_num"], k, tol, 0.0, 0.5, mem)
        f.append(info["fcn"])
        ev.append(info["evaluations"] - total)
        total = info["evaluations"]
    return dict(f=np.array(f), evals=np.array(ev))
