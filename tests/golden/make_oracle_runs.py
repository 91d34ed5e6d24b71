"""Writes oracle_runs.npz: L-BFGS traces of the CPU oracle (oracle/, the FP64
restatement of Cpp/src pinned by the tests in tests/test_oracle_*.py) on
runs too long to repeat inside a GPU test.

    python tests/golden/make_oracle_runs.py

c1_*: circle -> square, 400 points (lddmm_amd.geometry.circle_square), the
Gaussian kernel of width 0.25, 'G' 0.1 x 'B' varifold, weight 100, T = 21,
h = 0.05, zero initial momentum, L-BFGS memory 5, 200 iterations, tolerance
0, Wolfe (0, 0.5): energy, gradient norm, step length and evaluation count
per iteration (index 0 is the initial point), the final momentum's norm;
c1_spread_f: the final energies of the same run started from momenta of
size 1e-16 (four seeds) -- the spread that rounding alone gives this
200-iteration run (its line searches start to differ near iteration 30).

ex_*: the optimiser settings of the reference's examples
(Cpp/examples/sphere2cube/run_sphere2cube.m:11-31: 'C' 0.1 x 'B' varifold,
Matern order 3, tolerance 1e-2, Wolfe (0, 0.5)) on the 120-point circle ->
square curve (kernel width 0.25, T = 21), memory 50, 60 iterations: the same
trace fields.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from lddmm_amd import geometry as G  # noqa: E402


def c1(itr=200):
    w = G.circle_square(400)
    T = 21
    ifo = oracle.get_elm_ifo(w["elm"], 400)
    P = oracle.OracleProblem(w["q0"], w["elm"], w["tgt_cen"], w["tgt_dir"], w["tgt_vol"], w["cen_type"],
                             w["cen_width"], w["dir_type"], w["dir_width"], w["dis_wgt"], w["knl_order"],
                             w["knl_width"], 1.0 / (T - 1), T, ifo=ifo)
    t0 = time.time()
    alp, f, tr = P.run(itr_max=itr, tol=0.0, wolfe1=0.0, wolfe2=0.5, mem=5)
    print("c1: %d iterations, %d evaluations, E = %.15e (%.1f s)" % (tr["iters"], tr["total_evals"], f,
                                                                      time.time() - t0))
    spread = []
    for seed in range(4):
        _, fs, _ = P.run(alp0=1e-16 * G.normal(900 + seed, P.varnum), itr_max=itr, tol=0.0, wolfe1=0.0,
                         wolfe2=0.5, mem=5)
        spread.append(fs)
        print("c1 from a 1e-16 start (seed %d): E = %.15e, rel %.2e" % (seed, fs, (fs - f) / f))
    return dict(c1_f=tr["f"], c1_gnorm=tr["gnorm"], c1_step=tr["step"], c1_evals=tr["evals"],
                c1_status=np.array(tr["status"]), c1_alp_norm=np.array(np.linalg.norm(alp)),
                c1_spread_f=np.array(spread))


def example_settings(itr=60):
    w = G.circle_square(120)
    T = 21
    ifo = oracle.get_elm_ifo(w["elm"], 120)
    P = oracle.OracleProblem(w["q0"], w["elm"], w["tgt_cen"], w["tgt_dir"], w["tgt_vol"], "C", 0.1, "B", 0.0,
                             w["dis_wgt"], 3, 0.25, 1.0 / (T - 1), T, ifo=ifo)
    t0 = time.time()
    alp, f, tr = P.run(itr_max=itr, tol=1e-2, wolfe1=0.0, wolfe2=0.5, mem=50)
    print("ex: %d iterations, %d evaluations, E = %.15e (%.1f s)" % (tr["iters"], tr["total_evals"], f,
                                                                      time.time() - t0))
    return dict(ex_f=tr["f"], ex_gnorm=tr["gnorm"], ex_step=tr["step"], ex_evals=tr["evals"],
                ex_status=np.array(tr["status"]), ex_alp_norm=np.array(np.linalg.norm(alp)))


if __name__ == "__main__":
    oracle.build()
    out = c1()
    out.update(example_settings())
    np.savez_compressed(os.path.join(HERE, "oracle_runs.npz"), **out)
