"""The l2Dist landmark matching term on the GPU (MATLAB/src/l2Dist.m:1-5,
as the disHdl of MATLAB/src/objgrd.m) against the oracle, which is pinned by
finite differences and closed forms in test_oracle_math.py."""
import numpy as np
import pytest

import oracle
from conftest import rel_err
from lddmm_amd import geometry as G

pytestmark = pytest.mark.gpu


def landmarks(dim, n, seed):
    q = (0.4 * G.normal(seed, n * dim)).reshape(n, dim)
    t = q + (0.2 * G.normal(seed + 1, n * dim)).reshape(n, dim)
    return q, t


@pytest.mark.parametrize("dim,n", [(2, 1), (3, 7), (3, 1000), (2, 3001)])
def test_l2_dist_operator(gpu, dim, n):
    q, t = landmarks(dim, n, 3 * n + dim)
    f, g = gpu.l2Dist(q, t, 2.5)
    ef, eg = oracle.l2_dist(q, t, 2.5)
    assert abs(f - ef) <= 1e-13 * abs(ef)
    assert np.max(np.abs(g - eg)) == 0.0


@pytest.mark.parametrize("dim,order", [(2, -1), (3, -1), (3, 1), (2, 2)])
def test_l2_objective_matches_oracle(gpu, dim, order):
    n, T = 300, 6
    q, t = landmarks(dim, n, 40 + dim)
    obj = gpu.Objective.l2(q, t, 7.0, order, 0.3, 1.0 / (T - 1), T)
    none = np.zeros((0, dim))
    P = oracle.OracleProblem(q, np.zeros((0, dim), dtype=np.int32), none, none, np.zeros(0), "G", 0.3, "B", 0.0,
                             7.0, order, 0.3, 1.0 / (T - 1), T, ifo=np.zeros((n, 1), dtype=np.int32), tgt_nde=t)
    print("l2 objective dim=%d order=%d" % (dim, order), flush=True)
    for seed, scale in ((0, 0.0), (5, 0.1), (6, 0.5)):
        a = scale * G.normal(seed + 200, obj.varnum)
        f, g = obj(a)
        ef, eg = P.objgrd(a)
        assert abs(f - ef) <= 1e-10 * abs(ef), (seed, f, ef)
        assert rel_err(g, eg) <= 1e-10
    obj.set_precision("fp32")
    a = 0.1 * G.normal(300, obj.varnum)
    f32, g32 = obj(a)
    ef, eg = P.objgrd(a)
    assert abs(f32 - ef) <= 1e-6 * abs(ef) and rel_err(g32, eg) <= 1e-4
    obj.close()


def test_l2_registration_moves_landmarks_onto_targets(gpu):
    """Targets from a smooth deformation of the source (rotation, scaling and
    a sine warp): the registration brings the landmarks onto them."""
    q = (0.4 * G.normal(77, 1500)).reshape(500, 3)
    c, s_ = np.cos(0.3), np.sin(0.3)
    R = np.array([[c, -s_, 0.0], [s_, c, 0.0], [0.0, 0.0, 1.0]])
    t = 1.1 * q @ R.T + 0.05 * np.sin(2.0 * q[:, [1, 2, 0]])
    T = 11
    alp, info = gpu.run_l2(q, None, t, 100.0, -1, 0.3, 1.0 / (T - 1), T, 200, 0.0, 0.0, 0.5, 5, rel_energy_tol=1e-9)
    nde, _ = gpu.computeNodePosition(q, alp, -1, 0.3, 1.0 / (T - 1), T)
    err0 = np.sqrt(np.mean(np.sum((q - t) ** 2, axis=1)))
    err = np.sqrt(np.mean(np.sum((nde[-1] - t) ** 2, axis=1)))
    print("l2 registration: %d iterations, RMS landmark error %.3e -> %.3e" % (info["iterations"], err0, err),
          flush=True)
    assert info["status"] == 0
    assert err < 0.05 * err0
    # the run's energy is the objective at its result
    obj = gpu.Objective.l2(q, t, 100.0, -1, 0.3, 1.0 / (T - 1), T)
    f, _ = obj(alp)
    assert abs(f - info["fcn"]) <= 1e-12 * abs(f)
    obj.close()
