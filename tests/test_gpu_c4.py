"""The batched configuration at its stated size: problems of the
256-problem, seed-4 batch (vertex counts 492 to 20,252), T = 21, Gaussian
deformation kernels of widths in [0.15, 0.35], 'G' 0.1 x 'B' varifold.

The oracle takes seconds per evaluation only on the small problems, so the
smallest ones are compared with it directly at T = 21; the large ones are
checked through properties that do not depend on size: the FP32 policy
against FP64 on the same momenta, the batched path against the
single-problem path, and the registration protocol (FP32 runs stop only at
an energy confirmed in FP64 at or below the FP64 run's)."""
import numpy as np
import pytest

import oracle
from conftest import rel_err
from lddmm_amd import geometry as G

pytestmark = pytest.mark.gpu

COUNT, SEED, NMIN, NMAX = 256, 4, 500, 20000


def pick():
    n = G.batch_sizes(COUNT, SEED, NMIN, NMAX)
    order = np.argsort(n, kind="stable")
    return [int(i) for i in (order[-1], order[-3], order[len(order) // 2], order[1], order[0])], n


def c4_batch(gpu, idx, precision="fp64"):
    probs = G.batch_problems(COUNT, SEED, NMIN, NMAX, indices=idx)
    k = G.BATCH_KERNELS
    B = gpu.Batch(probs, k["cen_type"], k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"], k["knl_order"],
                  k["time_stp"], k["time_num"], precision=precision)
    return probs, k, B


def test_c4_smallest_problems_match_oracle_at_full_time_grid(gpu):
    idx, n = pick()
    small = idx[-2:]
    probs, k, B = c4_batch(gpu, small)
    print("c4 vs oracle: problems %s, N = %s" % (small, [int(n[i]) for i in small]), flush=True)
    for scale in (0.0, 0.02):
        alps = [scale * G.normal(500 + i, B.sizes[i]).reshape(20, B.n[i], 3) for i in range(B.P)]
        f, g = B.objgrd(alps)
        for i, p in enumerate(probs):
            ifo = oracle.get_elm_ifo(p["elm"], len(p["q0"]))
            P = oracle.OracleProblem(p["q0"], p["elm"], p["tgt_cen"], p["tgt_dir"], p["tgt_vol"], k["cen_type"],
                                     k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"], k["knl_order"],
                                     p["knl_width"], k["time_stp"], k["time_num"], ifo=ifo)
            ef, eg = P.objgrd(alps[i].reshape(-1))
            print("  scale %.2f N=%d: energy %.2e gradient %.2e" % (scale, B.n[i], abs(f[i] - ef) / abs(ef),
                                                                    rel_err(g[i].reshape(-1), eg)), flush=True)
            assert abs(f[i] - ef) <= 1e-10 * abs(ef)
            assert rel_err(g[i].reshape(-1), eg) <= 1e-10
    B.close()


def test_c4_fp32_policy_against_fp64_up_to_20k_vertices(gpu):
    """Five problems from 492 to 20,252 vertices in one batch: the FP32
    policy's energy within 1e-6 and gradient within 1e-4 (max-norm) of
    FP64 on the same momenta; the largest problem's batched FP64 result
    equals its single-problem evaluation."""
    idx, n = pick()
    probs, k, B64 = c4_batch(gpu, idx)
    _, _, B32 = c4_batch(gpu, idx, "fp32")
    print("c4 fp32 vs fp64: N = %s" % [int(n[i]) for i in idx], flush=True)
    for scale in (0.0, 0.02):
        alps = [scale * G.normal(700 + i, B64.sizes[i]).reshape(20, B64.n[i], 3) for i in range(B64.P)]
        f64, g64 = B64.objgrd(alps)
        f32, g32 = B32.objgrd(alps)
        for i in range(B64.P):
            e, gr = abs(f32[i] - f64[i]) / abs(f64[i]), rel_err(g32[i].reshape(-1), g64[i].reshape(-1))
            print("  scale %.2f N=%d: energy %.2e gradient %.2e" % (scale, B64.n[i], e, gr), flush=True)
            assert e <= 1e-6 and gr <= 1e-4
        if scale > 0:
            p = probs[0]
            ifo = oracle.get_elm_ifo(p["elm"], len(p["q0"]))
            obj = gpu.Objective(p["q0"], p["elm"], ifo, p["tgt_cen"], p["tgt_vol"], p["tgt_dir"], k["cen_type"],
                                k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"], k["knl_order"],
                                p["knl_width"], k["time_stp"], k["time_num"])
            fs, gs = obj(alps[0])
            assert abs(fs - f64[0]) <= 1e-12 * abs(fs) and rel_err(g64[0].reshape(-1), gs) <= 1e-11
            obj.close()
    B64.close()
    B32.close()


def test_c4_fp32_runs_reach_fp64_energies_confirmed_in_fp64(gpu):
    """The registrations/s protocol on problems of every size class: a
    20-iteration batched FP64 run sets each problem's energy; the batched
    FP32 policy runs until each problem's energy, confirmed by an FP64
    evaluation, is at or below it x (1 + 1e-6)."""
    idx, n = pick()
    probs, k, B = c4_batch(gpu, idx)
    _, i64, r64 = B.run(20, 0.0, 0.0, 0.5, 5)
    targets = np.array([r["fcn"] for r in i64]) * (1 + 1e-6)
    a32, i32, r32 = B.run(200, 0.0, 0.0, 0.5, 5, precision="fp32", targets=targets)
    f, _ = B.objgrd(a32)
    print("c4 protocol: fp64 rounds %d, fp32 rounds %d, iterations %s" %
          (r64, r32, [r["iterations"] for r in i32]), flush=True)
    for i in range(B.P):
        assert i32[i]["reached_target"], i
        assert f[i] <= targets[i] * (1 + 1e-13)
    B.close()
