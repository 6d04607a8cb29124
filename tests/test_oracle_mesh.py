"""Pins the oracle's restatement of mesh preparation (mesh.go: BoundingBox,
Transform, FitInside, SmoothNormals, SmoothNormalsThreshold) on the CPU:
against a pure-Python transcription of mesh.go's map code on small meshes
(Python dicts treat -0.0 == 0.0 as one key and a fresh NaN as a key never
found again, as Go maps do), against the independent numpy implementation
in fauxgl_amd/meshes.py, and against closed forms and special cases."""
import math

import numpy as np

import oracle as O
from fauxgl_amd import meshes as M
from fauxgl_amd.api import Prims, V

F = np.float64


def _norm(v):
    with np.errstate(divide="ignore", invalid="ignore"):
        r = F(1) / np.sqrt(F(v[0]) * F(v[0]) + F(v[1]) * F(v[1]) + F(v[2]) * F(v[2]))
        return (F(v[0]) * r, F(v[1]) * r, F(v[2]) * r)


def _key(p):
    return tuple(float(c) for c in p)  # fresh float objects: a NaN key is never found again


def go_smooth_threshold(pos, nrm, threshold):
    """mesh.go:80-103, transcribed."""
    lookup = {}
    for p, n in zip(pos, nrm):
        k = _key(p)
        lookup[k] = lookup.get(k, []) + [tuple(F(c) for c in n)]
    out = []
    for p, n in zip(pos, nrm):
        lst = lookup.get(_key(p), [])
        res = (F(0), F(0), F(0))
        for x in lst:
            if x[0] * F(n[0]) + x[1] * F(n[1]) + x[2] * F(n[2]) >= threshold:
                res = (res[0] + x[0], res[1] + x[1], res[2] + x[2])
        out.append(_norm(res))
    return np.array(out, dtype=np.float64)


def go_smooth(pos, nrm):
    """mesh.go:105-125, transcribed."""
    lookup = {}
    for p, n in zip(pos, nrm):
        k = _key(p)
        s = lookup.get(k, (F(0), F(0), F(0)))
        lookup[k] = (s[0] + F(n[0]), s[1] + F(n[1]), s[2] + F(n[2]))
    lookup = {k: _norm(v) for k, v in lookup.items()}
    return np.array([lookup.get(_key(p), (F(0), F(0), F(0))) for p in pos], dtype=np.float64)


def shared_mesh(seed, ntri=40):
    """Triangles over a small pool of positions (so positions repeat), with
    -0 coordinates, a NaN position and random unit-ish normals."""
    rng = np.random.Generator(np.random.PCG64(seed))
    pool = rng.integers(-2, 3, size=(12, 3)).astype(np.float64) * 0.5
    pool[0] = (0.0, 0.5, 1.0)
    pool[1] = (-0.0, 0.5, 1.0)  # same key as pool[0]
    pool[2] = (np.nan, 0.0, 0.0)
    idx = rng.integers(0, len(pool), size=(ntri, 3))
    pos = pool[idx]
    nrm = rng.normal(size=(ntri, 3, 3))
    nrm /= np.linalg.norm(nrm, axis=2, keepdims=True)
    return pos, nrm


def bits(a):
    return np.asarray(a, dtype=np.float64).view(np.uint64)


def same(a, b):
    """Bit-equal, NaNs of any payload counted equal."""
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    na, nb = np.isnan(a), np.isnan(b)
    return bool(np.array_equal(na, nb) and np.array_equal(bits(np.where(na, 0, a)), bits(np.where(nb, 0, b))))


def test_smooth_threshold_matches_transcription():
    for seed in range(6):
        pos, nrm = shared_mesh(seed)
        for deg in (0.0, 30.0, 75.0, 180.0):
            thr = math.cos(math.radians(deg))
            got = O.mesh_smooth_normals(pos, nrm, thr).reshape(-1, 3)
            want = go_smooth_threshold(pos.reshape(-1, 3), nrm.reshape(-1, 3), thr)
            assert same(got, want), (seed, deg)


def test_smooth_matches_transcription():
    for seed in range(6):
        pos, nrm = shared_mesh(seed)
        got = O.mesh_smooth_normals(pos, nrm).reshape(-1, 3)
        want = go_smooth(pos.reshape(-1, 3), nrm.reshape(-1, 3))
        assert same(got, want), seed
    # a NaN position reads the zero Vector (never normalized)
    pos, nrm = shared_mesh(0)
    flat = pos.reshape(-1, 3)
    nan_rows = np.isnan(flat).any(axis=1)
    assert nan_rows.any()
    assert not O.mesh_smooth_normals(pos, nrm).reshape(-1, 3)[nan_rows].any()


def test_smooth_threshold_matches_numpy_generator():
    """fauxgl_amd.meshes.smooth_normals_threshold is a separate (vectorised)
    implementation of the same code; bit-equal on a displaced icosphere."""
    pos = M.sphere_positions(4)
    pos = pos * (1.0 + 0.1 * np.sin(7 * pos[..., :1]))
    p = M.Prims(pos, M.with_face_normals(pos))
    want = M.smooth_normals_threshold(p, math.radians(30)).normal
    got = O.mesh_smooth_normals(p.position, p.normal, math.cos(math.radians(30)))
    assert same(got, want)


def test_smooth_special_thresholds():
    pos, nrm = shared_mesh(3)
    flat = pos.reshape(-1, 3)
    ok = ~np.isnan(flat).any(axis=1)
    # threshold above any dot product: nothing passes, Normalize(0) = NaN
    got = O.mesh_smooth_normals(pos, nrm, 2.0).reshape(-1, 3)
    assert np.isnan(got).all()
    # threshold below any dot product: every vertex gets SmoothNormals' value
    a = O.mesh_smooth_normals(pos, nrm, -2.0).reshape(-1, 3)
    b = O.mesh_smooth_normals(pos, nrm).reshape(-1, 3)
    assert same(a[ok], b[ok])
    # a position used once keeps its own normal (normalized once more)
    lone = np.array([[[5.0, 5, 5], [6, 6, 6], [7, 7, 7]]])
    n = np.array([[[0.0, 0, 2], [0, 3, 0], [4, 0, 0]]])
    assert same(O.mesh_smooth_normals(lone, n, 0.5)[0], [[0, 0, 1], [0, 1, 0], [1, 0, 0]])


def test_bounding_box():
    rng = np.random.Generator(np.random.PCG64(1))
    t = rng.normal(size=(50, 3, 3))
    lo, hi = O.mesh_bounding_box(t)
    assert same(lo, t.reshape(-1, 3).min(axis=0)) and same(hi, t.reshape(-1, 3).max(axis=0))
    # lines count too
    l = np.full((1, 2, 3), 9.0)
    lo, hi = O.mesh_bounding_box(t, l)
    assert same(hi, [9, 9, 9])
    # Box.Extend replaces a running box equal to EmptyBox: leading triangles
    # at the origin do not count
    z = np.zeros((2, 3, 3))
    far = np.full((1, 3, 3), 2.0)
    far[0, 0] = 1.0
    lo, hi = O.mesh_bounding_box(np.concatenate([z, far]))
    assert same(lo, [1, 1, 1]) and same(hi, [2, 2, 2])
    # ... but once the box is not empty, an origin triangle does count
    lo, hi = O.mesh_bounding_box(np.concatenate([far, z]))
    assert same(lo, [0, 0, 0]) and same(hi, [2, 2, 2])
    # all at the origin: EmptyBox
    lo, hi = O.mesh_bounding_box(z)
    assert same(lo, [0, 0, 0]) and same(hi, [0, 0, 0])
    # math.Min: NaN poisons, but -Inf wins over NaN
    t2 = t.copy()
    t2[3, 1, 0] = np.nan
    lo, hi = O.mesh_bounding_box(t2)
    assert np.isnan(lo[0]) and np.isnan(hi[0])
    t2[4, 2, 0] = -np.inf
    lo, hi = O.mesh_bounding_box(t2)
    assert lo[0] == -np.inf and np.isnan(hi[0])
    # -0 against +0: Min gives -0, Max gives +0
    s = np.array([[[1.0, 1, 1], [0.0, 2, 2], [-0.0, 3, 3]]])
    lo, hi = O.mesh_bounding_box(s)
    assert bits(lo)[0] == bits(-0.0) and hi[0] == 1.0


def test_transform_closed_forms():
    rng = np.random.Generator(np.random.PCG64(2))
    p = rng.normal(size=(30, 3))
    n = rng.normal(size=(30, 3))
    ident = np.eye(4).reshape(-1)
    P, Nn = O.mesh_transform(p, n, ident)
    assert same(P, p + 0.0)  # identity: x*1 + y*0 + z*0 + 0
    r = 1.0 / np.sqrt(n[:, 0] * n[:, 0] + n[:, 1] * n[:, 1] + n[:, 2] * n[:, 2])
    assert same(Nn, n * r[:, None])
    T = np.eye(4)
    T[:3, 3] = (0.5, -1.0, 2.0)
    P, Nn = O.mesh_transform(p, n, T.reshape(-1))
    assert same(P, p + np.array([0.5, -1.0, 2.0]))
    # matches the generator's numpy transform bit for bit
    m = rng.normal(size=16)
    pr = Prims(p.reshape(10, 3, 3), n.reshape(10, 3, 3))
    want = M.transform(pr, list(m))
    P, Nn = O.mesh_transform(p, n, m)
    assert same(P.reshape(-1, 3), want.position.reshape(-1, 3))
    assert same(Nn.reshape(-1, 3), want.normal.reshape(-1, 3))


def test_fit_inside():
    rng = np.random.Generator(np.random.PCG64(3))
    t = rng.normal(size=(60, 3, 3)) * np.array([3.0, 1.0, 0.5]) + 4.0
    tn = rng.normal(size=(60, 3, 3))
    tp, tnn, _, _, m = O.mesh_fit_inside(t, tn, (-1, -1, -1), (1, 1, 1), (0.5, 0.5, 0.5))
    want = M.bi_unit_cube(Prims(t, tn))
    assert same(tp, want.position) and same(tnn, want.normal)
    lo, hi = tp.reshape(-1, 3).min(axis=0), tp.reshape(-1, 3).max(axis=0)
    # the longest axis spans [-1, 1]; the box is centred
    assert abs(lo[0] + 1) < 1e-12 and abs(hi[0] - 1) < 1e-12
    assert np.all(np.abs(lo + hi) < 1e-12)
    # the matrix is Translate(-min), Scale(s), Translate(...) composed
    bmin = t.reshape(-1, 3).min(axis=0)
    s = 2.0 / (t.reshape(-1, 3).max(axis=0) - bmin).max()
    assert abs(m[0] - s) < 1e-15 and m[1] == 0 and m[4] == 0
    # anchor 0 puts the min corner on the box's min corner
    tp0, _, _, _, _ = O.mesh_fit_inside(t, tn, (-1, -1, -1), (1, 1, 1), (0, 0, 0))
    assert np.all(np.abs(tp0.reshape(-1, 3).min(axis=0) + 1) < 1e-12)
