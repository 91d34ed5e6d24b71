"""Trains the learned warm start of the batched configuration and writes
its parameters (lddmm_amd/data/warmstart_c4.npz by default):

    python -m lddmm_amd.train_warmstart --problems 512 --seed 1004

Training problems come from the batched-configuration generator
(geometry.batch_problems: jittered icospheres to seeded superquadrics,
vertex counts log-uniform in [500, 20000], Gaussian kernel widths in
[0.15, 0.35]) with seeds other than the benchmark's (4).  Each is solved by
the plain method -- FP64 arithmetic, zero start, L-BFGS memory 5, weak
Wolfe (0, 0.5), at most 300 iterations, stop at a relative decrease below
1e-6 -- with the batched evaluation (one launch per kernel per Euler step
for all problems of a chunk; per-problem results equal the single-problem
path's to rounding).  The network is then fitted on the GPU (BF16 MFMA
forward and backward, Adam) to the solutions' time-mean momentum density.
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

import lddmm_amd  # noqa: E402
from lddmm_amd import geometry as G  # noqa: E402
from lddmm_amd import warmstart as W  # noqa: E402

DEFAULT_OUT = os.path.join(HERE, "data", "warmstart_c4.npz")


def solve(problems, itr=300, precision="fp64"):
    k = G.BATCH_KERNELS
    B = lddmm_amd.Batch(problems, k["cen_type"], k["cen_width"], k["dir_type"], k["dir_width"], k["dis_wgt"],
                        k["knl_order"], k["time_stp"], k["time_num"], precision=precision)
    sol, info, rounds = B.run(itr, 0.0, 0.0, 0.5, 5, precision=precision, rel_energy_tol=1e-6)
    B.close()
    return sol, info


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--problems", type=int, default=512)
    ap.add_argument("--chunk", type=int, default=128)
    ap.add_argument("--seed", type=int, default=1004)
    ap.add_argument("--nmin", type=int, default=500)
    ap.add_argument("--nmax", type=int, default=20000)
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--batch-rows", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=DEFAULT_OUT)
    a = ap.parse_args(argv)
    k = G.BATCH_KERNELS
    ws = W.WarmStart(hidden=128, layers=4, seed=1, cen_type=k["cen_type"], cen_width=k["cen_width"],
                     dir_type=k["dir_type"], dir_width=k["dir_width"])
    t0 = time.time()
    probs, sols, its = [], [], []
    for c0 in range(0, a.problems, a.chunk):
        idx = list(range(c0, min(a.problems, c0 + a.chunk)))
        ps = G.batch_problems(a.problems, a.seed, a.nmin, a.nmax, indices=idx)
        sol, info = solve(ps)
        probs += ps
        sols += sol
        its += [r["iterations"] for r in info]
        print("solved %d of %d problems (%.0f s)" % (len(probs), a.problems, time.time() - t0), flush=True)
    t_solve = time.time() - t0
    t0 = time.time()
    losses = ws.fit(probs, sols, steps=a.steps, batch_rows=a.batch_rows, lr=a.lr)
    t_fit = time.time() - t0
    print("fit: %d rows, %d Adam steps, loss %.4f -> %.4f (%.1f s)" %
          (ws.meta["rows"], a.steps, losses[0], float(np.mean(losses[-50:])), t_fit), flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    ws.save(a.out, seed=a.seed, nmin=a.nmin, nmax=a.nmax, solve_seconds=t_solve, fit_seconds=t_fit,
            loss_first=float(losses[0]), loss_last=float(np.mean(losses[-50:])),
            median_iterations=float(np.median(its)))
    print("wrote", a.out, flush=True)


if __name__ == "__main__":
    main()
