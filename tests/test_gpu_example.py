"""The GPU path on the reference's own sphere->cube example
(tests/golden/sphere2cube.npz, from Cpp/examples/sphere2cube/*.mat; matching
and deformation parameters of Cpp/examples/sphere2cube/run_sphere2cube.m:
'C' 0.1 x 'B', weight 100, Matern order 3 of width 0.25, h = 0.05, T = 21)."""
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, rel_err
from lddmm_amd import io as lio

pytestmark = pytest.mark.gpu

K = dict(cen_type="C", cen_width=0.1, dir_type="B", dir_width=0.0, dis_wgt=100.0, knl_order=3, knl_width=0.25,
         time_stp=0.05, time_num=21)


@pytest.fixture(scope="module")
def example():
    z = np.load(os.path.join(GOLDEN, "sphere2cube.npz"))
    arrs = {k: z[k] for k in z.files}
    return arrs, lio.problem_from_arrays(arrs)


def test_element_varifold_of_the_example_target(gpu, example):
    arrs, p = example
    cen, udir, vol = gpu.elementVarifold(p["tgt_nde"], p["tgt_elm"])
    assert np.max(np.abs(cen - p["tgt_cen"])) <= 1e-14
    assert np.max(np.abs(udir - p["tgt_dir"])) <= 1e-12
    assert np.max(np.abs(vol - p["tgt_vol"])) <= 1e-15


def test_varifold_on_the_example_matches_the_reference(gpu, example):
    arrs, p = example
    ifo = gpu.getElmIfo(p["elm"], len(p["q0"]), one_based=False)
    val, dq = gpu.varifold(p["q0"], p["elm"], ifo, p["tgt_cen"], p["tgt_dir"], p["tgt_vol"], "C", 0.1, "B", 0.0)
    dd, dt, tt = oracle.varifold_terms(p["q0"], p["elm"], p["tgt_cen"], p["tgt_dir"], p["tgt_vol"], "C", 0.1, "B",
                                       0.0)
    print("example varifold: GPU %.15e reference %.15e" % (val, float(arrs["ref_vf_val"])), flush=True)
    # 9e8 target-target pairs summed in different orders (see test_example_data.py)
    assert abs(val - float(arrs["ref_vf_val"])) <= 1e-10 * (abs(dd) + 2 * abs(dt) + abs(tt))
    assert rel_err(dq, arrs["ref_vf_dq"]) <= 1e-12


def test_objective_on_the_example_matches_the_oracle(gpu, example):
    arrs, p = example
    n = len(p["q0"])
    ifo = oracle.get_elm_ifo(p["elm"], n)
    T = K["time_num"]
    rng = np.random.default_rng(7)
    alp = 0.02 * rng.standard_normal((T - 1, n, 3))
    obj = gpu.Objective(p["q0"], p["elm"], ifo, p["tgt_cen"], p["tgt_vol"], p["tgt_dir"], K["cen_type"],
                        K["cen_width"], K["dir_type"], K["dir_width"], K["dis_wgt"], K["knl_order"], K["knl_width"],
                        K["time_stp"], T)
    f, g = obj(alp)
    ref = oracle.OracleProblem(p["q0"], p["elm"], p["tgt_cen"], p["tgt_dir"], p["tgt_vol"], K["cen_type"],
                               K["cen_width"], K["dir_type"], K["dir_width"], K["dis_wgt"], K["knl_order"],
                               K["knl_width"], K["time_stp"], T, ifo=ifo)
    fo, go = ref.objgrd(alp)
    print("example objective: GPU %.15e oracle %.15e, gradient %.2e" % (f, fo, rel_err(g, go)), flush=True)
    assert abs(f - fo) <= 1e-10 * abs(fo)
    assert rel_err(g, go) <= 1e-10
    obj.set_precision("fp32")
    f32, g32 = obj(alp)
    assert abs(f32 - f) <= 1e-6 * abs(f) and rel_err(g32, g) <= 1e-4
    obj.close()


def test_registration_on_the_example(gpu, example):
    """The example registered for 40 iterations in FP64 and in the FP32
    policy run to the FP64 energy: monotone decrease, FP32 reaches it."""
    arrs, p = example
    n = len(p["q0"])
    ifo = gpu.getElmIfo(p["elm"], n, one_based=False)
    common = (p["q0"], None, p["elm"], ifo, p["tgt_cen"], p["tgt_vol"], p["tgt_dir"], K["cen_type"],
              K["cen_width"], K["dir_type"], K["dir_width"], K["dis_wgt"], K["knl_order"], K["knl_width"],
              K["time_stp"], K["time_num"])
    a64, i64 = gpu.run(*common, 40, 0.0, 0.0, 0.5, 5)
    a32, i32 = gpu.run(*common, 200, 0.0, 0.0, 0.5, 5, precision="fp32", target_energy=i64["fcn"] * (1 + 1e-6))
    print("example registration: FP64 40 it E=%.10e (%.2f s); FP32 %d it E=%.10e (%.2f s)" %
          (i64["fcn"], i64["seconds"], i32["iterations"], i32["fcn"], i32["seconds"]), flush=True)
    assert i64["status"] == 0 and i64["iterations"] == 40
    assert i32["reached_target"] and i32["fcn"] <= i64["fcn"] * (1 + 1e-6) * (1 + 1e-13)
