"""The reference's own sphere->cube example data (tests/golden/sphere2cube.npz,
converted from Cpp/examples/sphere2cube/*.mat by tests/golden/make_golden.py):
the target preprocessing, the MAT reader and the oracle against it."""
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, rel_err
from lddmm_amd import geometry as G
from lddmm_amd import io as lio


@pytest.fixture(scope="module")
def example():
    z = np.load(os.path.join(GOLDEN, "sphere2cube.npz"))
    return {k: z[k] for k in z.files}


def test_target_varifold_of_the_example_matches_its_mesh(example):
    """tgtCenPosMat / tgtUniDirMat / tgtElmVolVec of cubeData.mat are the
    element varifold of its tgtNdeMat / tgtElmVtxMat (the formulas of
    Cpp/src/varifold3D.cpp:8-39)."""
    p = lio.problem_from_arrays(example)
    tc, tu, ta = G.triangle_varifold(p["tgt_nde"], p["tgt_elm"])
    assert np.max(np.abs(tc - p["tgt_cen"])) <= 1e-14
    assert np.max(np.abs(tu - p["tgt_dir"])) <= 1e-12
    assert np.max(np.abs(ta - p["tgt_vol"])) <= 1e-15


def test_mat_reader_round_trip(example, tmp_path):
    import scipy.io as sio
    src = {k: example[k] for k in ("iniNdeMat", "elmVtxMat")}
    tgt = {k: example[k] for k in ("tgtCenPosMat", "tgtUniDirMat", "tgtElmVolVec", "tgtNdeMat", "tgtElmVtxMat")}
    sio.savemat(tmp_path / "s.mat", src)
    sio.savemat(tmp_path / "t.mat", tgt)
    p = lio.load_example(str(tmp_path / "s.mat"), str(tmp_path / "t.mat"))
    assert p["q0"].shape == (600, 3) and p["elm"].shape == (1196, 3) and p["elm"].dtype == np.int32
    assert p["elm"].min() == 0 and p["elm"].max() == 599
    assert np.array_equal(p["q0"], example["iniNdeMat"].T)
    assert p["tgt_cen"].shape == (30000, 3) and p["tgt_vol"].shape == (30000,)


def test_oracle_varifold_on_the_example_matches_the_reference(example):
    p = lio.problem_from_arrays(example)
    ifo = oracle.get_elm_ifo(p["elm"], len(p["q0"]))
    val, dq = oracle.varifold(p["q0"], p["elm"], ifo, p["tgt_cen"], p["tgt_dir"], p["tgt_vol"], "C", 0.1, "B", 0.0,
                              grad=True)
    # the target-target term sums 9e8 Cauchy-kernel pairs (heavy tails: all
    # contribute); the reference and the oracle add them in different orders,
    # so their rounding differs at ~3e-11 of the terms' size
    dd, dt, tt = oracle.varifold_terms(p["q0"], p["elm"], p["tgt_cen"], p["tgt_dir"], p["tgt_vol"], "C", 0.1, "B",
                                       0.0)
    scale = abs(dd) + 2 * abs(dt) + abs(tt)
    print("example varifold: V = %.15e, terms %.3e %.3e %.3e" % (val, dd, dt, tt))
    assert abs(val - float(example["ref_vf_val"])) <= 1e-10 * scale
    assert rel_err(dq, example["ref_vf_dq"]) <= 1e-12
