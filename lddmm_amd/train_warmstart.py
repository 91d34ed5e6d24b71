"""Trains the learned warm start of the batched configuration and writes
its parameters (lddmm_amd/data/warmstart_c4.npz by default):

    python -m lddmm_amd.train_warmstart --first 896 --problems 896 --save-solved part.npz
    python -m lddmm_amd.train_warmstart --problems 896 --add-solved part.npz

(1,792 problems in two stages of about 16 GPU-minutes each; seed 1004,
source chunk 64.)

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

A set larger than one session's solves is built in stages: a run with
--first F --problems M --save-solved part.npz solves problems F .. F+M-1 and
writes their time-mean momenta without fitting; a later run given
--add-solved part.npz fits on its own problems and those.
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
    ap.add_argument("--problems", type=int, default=896)
    ap.add_argument("--chunk", type=int, default=64)
    ap.add_argument("--seed", type=int, default=1004)
    ap.add_argument("--nmin", type=int, default=500)
    ap.add_argument("--nmax", type=int, default=20000)
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--batch-rows", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--first", type=int, default=0,
                    help="generator index of the first problem solved here (problems first .. first + problems - 1)")
    ap.add_argument("--save-solved", default=None,
                    help="write the solved problems' time-mean momenta to this .npz and stop (no fit)")
    ap.add_argument("--add-solved", nargs="*", default=[],
                    help=".npz files written by --save-solved (same seed and sizes) whose problems join the fit")
    a = ap.parse_args(argv)
    k = G.BATCH_KERNELS
    ws = W.WarmStart(hidden=128, layers=4, seed=1, cen_type=k["cen_type"], cen_width=k["cen_width"],
                     dir_type=k["dir_type"], dir_width=k["dir_width"])
    t0 = time.time()
    # the fit uses the time mean of each solution only, so a solution is kept
    # as its mean, one slice (T-1 = 1)
    probs, sols, its = [], [], []
    last = a.first + a.problems
    for c0 in range(a.first, last, a.chunk):
        idx = list(range(c0, min(last, c0 + a.chunk)))
        ps = G.batch_problems(last, a.seed, a.nmin, a.nmax, indices=idx)
        sol, info = solve(ps)
        probs += ps
        sols += [np.asarray(x).mean(axis=0)[None] for x in sol]
        its += [r["iterations"] for r in info]
        print("solved %d of %d problems (%.0f s)" % (len(probs), a.problems, time.time() - t0), flush=True)
    t_solve = time.time() - t0
    if a.save_solved:
        np.savez(a.save_solved, indices=np.arange(a.first, last), sizes=np.array([len(x[0]) for x in sols]),
                 means=np.concatenate([x[0] for x in sols]).astype(np.float32), iterations=np.array(its),
                 seed=a.seed, nmin=a.nmin, nmax=a.nmax, solve_seconds=t_solve)
        print("wrote", a.save_solved, flush=True)
        return
    for path in a.add_solved:
        z = np.load(path, allow_pickle=False)
        if (int(z["seed"]), int(z["nmin"]), int(z["nmax"])) != (a.seed, a.nmin, a.nmax):
            raise SystemExit("%s: solved with another generator setting" % path)
        idx = [int(i) for i in z["indices"]]
        ps = G.batch_problems(max(idx) + 1, a.seed, a.nmin, a.nmax, indices=idx)
        offs = np.concatenate([[0], np.cumsum(z["sizes"])])
        for j, p in enumerate(ps):
            if len(p["q0"]) != int(z["sizes"][j]):
                raise SystemExit("%s: problem %d has another size" % (path, idx[j]))
            sols.append(z["means"][offs[j]:offs[j + 1]].astype(np.float64)[None])
        probs += ps
        its += [int(i) for i in z["iterations"]]
        t_solve += float(z["solve_seconds"])
        print("added %d solved problems from %s" % (len(ps), path), flush=True)
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
