"""The spatial order in which the library holds a problem (host code of the
C-ABI library, no GPU): a permutation whose aligned 64- and 128-slot blocks
are compact patches, independent of the input order up to the order inside
a block."""
import numpy as np
import pytest

import lddmm_amd as L
from lddmm_amd import geometry as G


def block_radii(x, ti):
    return np.array([np.max(np.linalg.norm(x[s:s + ti] - x[s:s + ti].mean(0), axis=1))
                     for s in range(0, len(x), ti)])


@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 128, 130, 256, 257, 1000, 4099])
def test_is_a_permutation(n):
    pts = G.normal(3, 3 * n).reshape(n, 3)
    p = L.spatial_order(pts)
    assert p.dtype == np.int32 and np.array_equal(np.sort(p), np.arange(n))


@pytest.mark.parametrize("ti", [64, 128])
def test_sphere_tiles_are_compact_patches(ti):
    """Icosphere (10,242 vertices; 20,480 element centres): every aligned
    tile lies in a ball of at most 2.5x the radius of the disk that holds
    `ti` points' share of the sphere's area (the bisection's cells have
    bounded aspect ratio; tiles of the mesh's own order reach 6-10x)."""
    q, tri = G._geodesic_cached(32)
    for pts in (q, q[tri].mean(1)):
        share = np.sqrt(ti * 4.0 / len(pts))
        r = block_radii(pts[L.spatial_order(pts)], ti)
        assert r.max() <= 2.5 * share, (r.max(), share)


def test_shuffled_circle_tiles_are_arcs():
    """4,000 points of a circle in random order: each aligned 64-block is an
    arc of 64 consecutive points (radius = half its chord)."""
    t = np.linspace(0.0, 2 * np.pi, 4000, endpoint=False)
    c = np.c_[np.cos(t), np.sin(t)]
    c = c[np.random.default_rng(0).permutation(len(c))]
    r = block_radii(c[L.spatial_order(c)], 64)
    assert r.max() <= np.sin(np.pi * 64 / 4000) * 1.001


def test_tile_sets_do_not_depend_on_the_input_order():
    """With distinct coordinates every median split is unique, so the set of
    points in each aligned block does not depend on the input order (ties
    are broken by input index)."""
    q, _ = G._geodesic_cached(12)
    q = q + 1e-6 * G.normal(9, q.size).reshape(q.shape)
    perm = np.random.default_rng(4).permutation(len(q))
    a = q[L.spatial_order(q)]
    b = q[perm][L.spatial_order(q[perm])]
    for s in range(0, len(q), 64):
        ka = sorted(map(tuple, a[s:s + 64]))
        kb = sorted(map(tuple, b[s:s + 64]))
        assert ka == kb


def test_argument_errors():
    with pytest.raises(L.LddmmError):
        L.spatial_order(np.zeros((5, 4)))
