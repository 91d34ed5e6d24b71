"""The staged training set of lddmm_amd.train_warmstart (host logic only: the
solves and the network are replaced by stand-ins, so this runs without a
GPU): problems solved by a --save-solved run join a later run's fit through
--add-solved, in generator order, with their time-mean momenta."""
import numpy as np

import lddmm_amd.train_warmstart as TW
from lddmm_amd import geometry as G
from lddmm_amd import warmstart as W

T = 21


def stand_in_solve(ps, itr=300, precision="fp64"):
    sols = [np.random.default_rng(len(p["q0"]) * 7 + int(1e3 * p["knl_width"])).normal(size=(T - 1, len(p["q0"]), 3))
            for p in ps]
    return sols, [dict(iterations=5 + i) for i in range(len(ps))]


def test_saved_problems_join_a_later_fit(monkeypatch, tmp_path):
    got = {}

    def fit(self, problems, solutions, **kw):
        got["sizes"] = [len(p["q0"]) for p in problems]
        got["widths"] = [p["knl_width"] for p in problems]
        got["means"] = [np.asarray(a).mean(axis=0) for a in solutions]
        self.meta = dict(problems=len(problems), rows=sum(got["sizes"]))
        return np.array([1.0, 0.5])

    class StandInMlp:
        def __init__(self, *a, **k):
            pass

    monkeypatch.setattr(TW, "solve", stand_in_solve)
    monkeypatch.setattr(W, "Mlp", StandInMlp)
    monkeypatch.setattr(W.WarmStart, "fit", fit)
    monkeypatch.setattr(W.WarmStart, "save", lambda self, path, **extra: got.update(saved=extra))
    part = str(tmp_path / "part.npz")
    TW.main(["--first", "5", "--problems", "4", "--chunk", "3", "--nmax", "900", "--save-solved", part])
    z = np.load(part)
    assert list(z["indices"]) == [5, 6, 7, 8] and list(z["iterations"]) == [5, 6, 7, 5]
    TW.main(["--problems", "5", "--chunk", "2", "--nmax", "900", "--add-solved", part, "--out", str(tmp_path / "w.npz")])
    ref = G.batch_problems(9, 1004, 500, 900)
    assert got["sizes"] == [len(p["q0"]) for p in ref]
    assert got["widths"] == [p["knl_width"] for p in ref]
    for p, m in zip(ref, got["means"]):
        exact = stand_in_solve([p])[0][0].mean(axis=0)
        assert np.abs(m - exact).max() <= 1e-6 * np.abs(exact).max()   # FP32 storage of the saved means
    # the iteration counts of both stages: chunks of 2 here, of 3 in the saved run
    assert got["saved"]["median_iterations"] == float(np.median([5, 6, 5, 6, 5] + [5, 6, 7, 5]))


def test_vertex_area_is_a_third_of_the_adjacent_triangles():
    """WarmStart's vertex areas A_i (the density scale of its targets): a
    third of the areas of the triangles at each vertex, against a plain loop,
    and they add up to the surface area."""
    q, tri = G.geodesic_sphere(3)
    q = q + 0.01 * G.normal(9, q.size).reshape(q.shape)
    A = W.vertex_area(q, tri)
    ref = np.zeros(len(q))
    total = 0.0
    for t in np.asarray(tri):
        a = 0.5 * np.linalg.norm(np.cross(q[t[1]] - q[t[0]], q[t[2]] - q[t[0]]))
        total += a
        for v in t:
            ref[v] += a / 3.0
    assert np.allclose(A, ref, rtol=1e-14, atol=0.0)
    assert abs(A.sum() - total) <= 1e-12 * total
