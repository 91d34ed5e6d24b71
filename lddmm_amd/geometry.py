"""Seeded synthetic inputs for the registration workloads.

Every generator is deterministic: the only randomness comes from a
splitmix64 counter stream (`uniform`, `normal`), so the same seed gives the
same bits on every machine.  Shapes follow the package conventions: points
(N, D), 0-based element tables (E, V), target varifolds as centres (Et, D),
unit directions (Et, D) and volumes (Et,).

Workloads (names used by bench.py and the tests):
  circle_square   400-point unit circle -> 400-segment square [-0.9, 0.9]^2
  sphere_cube     icosphere (5 subdivisions: 10,242 vertices, 20,480
                  triangles) -> cube [-0.8, 0.8]^3, 58 x 58 cells per face,
                  2 triangles per cell (40,368 triangles)
  superquadric    geodesic sphere of frequency f mapped onto a seeded
                  superquadric
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def _splitmix(x):
    x = (x + _GOLDEN).astype(np.uint64)
    z = x.copy()
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def uniform(seed, n, offset=0):
    """n doubles in [0, 1) from the splitmix64 stream of `seed`."""
    with np.errstate(over="ignore"):
        ctr = np.uint64(seed) * _GOLDEN + np.arange(offset, offset + n, dtype=np.uint64) * _GOLDEN
        z = _splitmix(ctr)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, n, offset=0):
    """n standard normals (Box-Muller on the uniform stream)."""
    m = (n + 1) // 2
    u = uniform(seed, 2 * m, 2 * offset)
    u1, u2 = 1.0 - u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    return z[:n]


# ---------------------------------------------------------------- 2D curves
def circle_curve(n=400, radius=1.0):
    """Closed polyline through n points evenly spaced on a circle."""
    ang = 2.0 * np.pi * np.arange(n) / n
    q = radius * np.stack([np.cos(ang), np.sin(ang)], axis=1)
    elm = np.stack([np.arange(n), (np.arange(n) + 1) % n], axis=1).astype(np.int32)
    return q, elm


def square_polyline(n=400, half=0.9):
    """Vertices of a closed n-segment polyline on the square [-half, half]^2,
    evenly spaced by arclength, counter-clockwise from (-half, -half)
    (the target construction of Cpp/examples/circle2square/run_circle2square.m:12-19)."""
    if n % 4:
        raise ValueError("n must be a multiple of 4")
    s = n // 4
    side = np.linspace(-half, half, s + 1)[:s]
    x = np.concatenate([side, np.full(s, half), -side, np.full(s, -half)])
    y = np.concatenate([np.full(s, -half), side, np.full(s, half), -side])
    return np.stack([x, y], axis=1)


def polyline_varifold(v):
    """Centres, unit directions and lengths of the closed polyline v
    (Cpp/examples/circle2square/run_circle2square.m:21-24)."""
    w = np.roll(v, -1, axis=0)
    c = 0.5 * (v + w)
    d = w - v
    a = np.sqrt((d ** 2).sum(1))
    return c, d / a[:, None], a


# ---------------------------------------------------------------- 3D surfaces
def _icosahedron():
    t = (1.0 + np.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t],
                  [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                  [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def icosphere(subdivisions=5, radius=1.0):
    """Recursive midpoint subdivision of the icosahedron with projection to
    the sphere after every level: 10*4^s + 2 vertices, 20*4^s triangles,
    outward orientation."""
    v, f = _icosahedron()
    verts = [tuple(x) for x in v]
    for _ in range(subdivisions):
        cache = {}
        newf = []

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = 0.5 * (np.asarray(verts[a]) + np.asarray(verts[b]))
                verts.append(tuple(m / np.linalg.norm(m)))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            newf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = np.array(newf, dtype=np.int64)
    return radius * np.array(verts), f.astype(np.int32)


def geodesic_sphere(freq, radius=1.0):
    """Class-I geodesic sphere of frequency `freq`: every icosahedron face
    cut into freq^2 triangles, vertices projected to the sphere;
    10 freq^2 + 2 vertices, 20 freq^2 triangles, outward orientation."""
    v, f = _icosahedron()
    n = int(freq)
    index = {}
    pts = []

    def vid(key):
        # key: sorted tuple of (corner vertex, integer weight) with weights summing to n
        if key not in index:
            p = sum(w * v[c] for c, w in key) / n
            pts.append(p / np.linalg.norm(p))
            index[key] = len(pts) - 1
        return index[key]

    def key_of(a, b, c, i, j):
        w = {a: n - i - j}
        w[b] = w.get(b, 0) + i
        w[c] = w.get(c, 0) + j
        return tuple(sorted((k, x) for k, x in w.items() if x))

    tris = []
    for a, b, c in f:
        grid = {}
        for i in range(n + 1):
            for j in range(n + 1 - i):
                grid[i, j] = vid(key_of(a, b, c, i, j))
        for i in range(n):
            for j in range(n - i):
                tris.append((grid[i, j], grid[i + 1, j], grid[i, j + 1]))
                if i + j < n - 1:
                    tris.append((grid[i + 1, j], grid[i + 1, j + 1], grid[i, j + 1]))
    return radius * np.array(pts), np.array(tris, dtype=np.int32)


def cube_surface(cells=58, half=0.8):
    """Triangulated surface of the cube [-half, half]^3: each face is a
    cells x cells grid, two triangles per cell, outward normals.  Returns
    vertices (with duplicates along face seams) and triangles."""
    g = np.linspace(-half, half, cells + 1)
    U, V = np.meshgrid(g, g, indexing="ij")
    u, w = U.ravel(), V.ravel()
    verts, tris = [], []
    base = 0
    # (axis, sign): the face x_axis = sign*half; (e1, e2) oriented so e1 x e2 points outward
    for axis in range(3):
        for sign in (-1.0, 1.0):
            a1, a2 = (axis + 1) % 3, (axis + 2) % 3
            p = np.zeros(((cells + 1) ** 2, 3))
            p[:, axis] = sign * half
            if sign > 0:
                p[:, a1], p[:, a2] = u, w
            else:
                p[:, a1], p[:, a2] = w, u
            verts.append(p)
            idx = np.arange((cells + 1) ** 2).reshape(cells + 1, cells + 1) + base
            i00, i10, i01, i11 = idx[:-1, :-1], idx[1:, :-1], idx[:-1, 1:], idx[1:, 1:]
            t1 = np.stack([i00.ravel(), i10.ravel(), i11.ravel()], 1)
            t2 = np.stack([i00.ravel(), i11.ravel(), i01.ravel()], 1)
            tris += [t1, t2]
            base += (cells + 1) ** 2
    return np.concatenate(verts), np.concatenate(tris).astype(np.int32)


def triangle_varifold(q, tri):
    """Centres, unit normals and areas of triangles
    (the element geometry of Cpp/src/varifold3D.cpp:8-39)."""
    p0, p1, p2 = q[tri[:, 0]], q[tri[:, 1]], q[tri[:, 2]]
    c = (p0 + p1 + p2) / 3.0
    nrm = np.cross(p1 - p0, p2 - p0)
    ln = np.linalg.norm(nrm, axis=1)
    return c, nrm / ln[:, None], 0.5 * ln


def _spow(x, e):
    return np.sign(x) * np.abs(x) ** e


def superquadric(freq, seed, exps=None, radii=None):
    """Geodesic sphere of frequency `freq` mapped onto a superquadric
    x = a1 s(cos eta)^e1 s(cos om)^e2, y = a2 s(cos eta)^e1 s(sin om)^e2,
    z = a3 s(sin eta)^e1 with signed powers; exponents uniform in
    [0.3, 2.0] and radii uniform in [0.7, 1.3] from `seed` unless given."""
    u = uniform(seed, 5)
    e1, e2 = (0.3 + 1.7 * u[:2]) if exps is None else exps
    a = (0.7 + 0.6 * u[2:5]) if radii is None else np.asarray(radii)
    s, tri = geodesic_sphere(freq)
    eta = np.arcsin(np.clip(s[:, 2], -1.0, 1.0))
    om = np.arctan2(s[:, 1], s[:, 0])
    ce = _spow(np.cos(eta), e1)
    q = np.stack([a[0] * ce * _spow(np.cos(om), e2), a[1] * ce * _spow(np.sin(om), e2),
                  a[2] * _spow(np.sin(eta), e1)], axis=1)
    return q, tri


# ---------------------------------------------------------------- workloads
def circle_square(n=400):
    """2D curve matching workload: source, 0-based segments, target varifold."""
    q, elm = circle_curve(n)
    tc, tu, ta = polyline_varifold(square_polyline(n, 0.9))
    return dict(q0=q, elm=elm, tgt_cen=tc, tgt_dir=tu, tgt_vol=ta, cen_type="G", cen_width=0.1, dir_type="B",
                dir_width=0.0, dis_wgt=100.0, knl_order=-1, knl_width=0.25, time_stp=0.05, time_num=21)


def sphere_cube(subdivisions=5, cells=58):
    """3D surface matching workload: icosphere -> cube, Matern nu = 3/2."""
    q, tri = icosphere(subdivisions)
    cv, ct = cube_surface(cells, 0.8)
    tc, tu, ta = triangle_varifold(cv, ct)
    return dict(q0=q, elm=tri, tgt_cen=tc, tgt_dir=tu, tgt_vol=ta, cen_type="G", cen_width=0.15, dir_type="B",
                dir_width=0.0, dis_wgt=100.0, knl_order=1, knl_width=0.3, time_stp=0.05, time_num=21)


_GEO_CACHE = {}


def _geodesic_cached(freq):
    if freq not in _GEO_CACHE:
        _GEO_CACHE[freq] = geodesic_sphere(freq)
    return _GEO_CACHE[freq]


def batch_sizes(count, seed, nmin=500, nmax=20000):
    """Vertex counts of batch_problems(count, seed, nmin, nmax), without
    building the problems."""
    u = uniform(seed, 2 * count)
    out = []
    for p in range(count):
        nn = np.exp(np.log(nmin) + (np.log(nmax) - np.log(nmin)) * u[2 * p])
        f = max(1, int(round(np.sqrt((nn - 2) / 10.0))))
        out.append(10 * f * f + 2)
    return np.array(out)


def batch_problems(count, seed, nmin=500, nmax=20000, jitter=0.01, indices=None):
    """Uneven batch of 3D registrations: vertex counts log-uniform in
    [nmin, nmax] (rounded to a geodesic sphere of frequency
    f = round(sqrt((N - 2) / 10)), N = 10 f^2 + 2); sources are jittered
    unit spheres (Gaussian vertex noise of std `jitter`), targets seeded
    superquadrics of the same frequency; deformation kernel widths uniform
    in [0.15, 0.35].  `indices` builds only those problems of the batch."""
    u = uniform(seed, 2 * count)
    out = []
    for p in (range(count) if indices is None else indices):
        nn = np.exp(np.log(nmin) + (np.log(nmax) - np.log(nmin)) * u[2 * p])
        f = max(1, int(round(np.sqrt((nn - 2) / 10.0))))
        s, tri = _geodesic_cached(f)
        q = s + jitter * normal(seed * 1000 + 7 * p + 1, s.size).reshape(s.shape)
        tq, ttri = superquadric(f, seed * 1000 + 7 * p + 3)
        tc, tu, ta = triangle_varifold(tq, ttri)
        out.append(dict(q0=q, elm=tri, tgt_cen=tc, tgt_dir=tu, tgt_vol=ta, knl_width=0.15 + 0.2 * u[2 * p + 1]))
    return out


BATCH_KERNELS = dict(cen_type="G", cen_width=0.1, dir_type="B", dir_width=0.0, dis_wgt=100.0, knl_order=-1,
                     time_stp=0.05, time_num=21)


def superquadric_pair(freq, seed, knl_width=0.2):
    """Single large 3D problem: two independent seeded superquadrics."""
    q, tri = superquadric(freq, seed)
    tq, ttri = superquadric(freq, seed + 1)
    tc, tu, ta = triangle_varifold(tq, ttri)
    return dict(q0=q, elm=tri, tgt_cen=tc, tgt_dir=tu, tgt_vol=ta, cen_type="G", cen_width=0.1, dir_type="B",
                dir_width=0.0, dis_wgt=100.0, knl_order=-1, knl_width=knl_width, time_stp=0.05, time_num=21)
