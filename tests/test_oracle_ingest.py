"""Pins the oracle's restatement of LoadSTL / LoadOBJ (stl.go, obj.go) on
the CPU: exact round trips of seeded files, strconv.ParseFloat cases stated
from Go's documentation, the reference's example assets against the arrays
the golden fixtures hold (parsed there by tests/golden/make_golden.py), face
normals against the numpy formula, and the inputs the reference rejects."""
import numpy as np
import pytest

import ingest_cases as C
import oracle as O
from fauxgl_amd import meshes as M
from test_oracle_mesh import same


def test_binary_stl_round_trip():
    t, data = C.random_binary_stl(1)
    pos, nrm = O.load_stl(data)
    assert same(pos, t.astype(np.float64))
    assert same(nrm, M.with_face_normals(pos))
    assert np.isnan(nrm[0]).all()  # degenerate triangle: Normalize(0)


def test_ascii_stl_tokens():
    data, xs, zs = C.odd_ascii_stl()
    pos, nrm = O.load_stl(data)
    flat = pos.reshape(-1, 3)
    assert same(flat[:, 0], xs)
    assert (flat[:, 1] == 1.25).all()
    assert same(flat[:, 2], zs)
    assert np.signbit(flat[3, 0])  # "-0" keeps its sign


def test_example_assets_match_golden_arrays():
    g = np.load(C.os.path.join(C.GOLDEN, "golden.npz"))
    pos, nrm = O.load_stl(C.fixture("cube.stl"))
    assert same(pos, g["cube_pos"])
    pos, nrm = O.load_stl(C.fixture("hello.stl"))  # ASCII
    assert same(nrm, M.with_face_normals(pos))
    # the golden arrays hold hello.stl after BiUnitCube and SmoothNormalsThreshold(30 degrees)
    hp = M.smooth_normals_threshold(M.bi_unit_cube(M.Prims(pos, nrm)), np.radians(30.0))
    assert same(hp.position, g["hello_pos"]) and same(hp.normal, g["hello_nrm"])
    pos, nrm, uv = O.load_obj(C.fixture("square.obj"))
    assert same(pos, g["square_pos"]) and same(uv, g["square_uv"])
    assert (nrm == np.array([0.0, 0.0, 1.0])).all()


def test_obj_indices_and_fix_normals():
    data = C.random_obj(3)
    pos, nrm, uv = O.load_obj(data)
    # rebuild by hand: fan (0, i, i+1), index forms decoded with Python
    v, vt, vn, tris = [np.zeros(3)], [np.zeros(2)], [np.zeros(3)], []
    for line in data.decode().splitlines():
        f = line.split()
        if not f:
            continue
        if f[0] == "v":
            v.append(np.array([float(x) for x in f[1:4]]))
        elif f[0] == "vt":
            vt.append(np.array([float(x) for x in f[1:3]]))
        elif f[0] == "vn":
            vn.append(np.array([float(x) for x in f[1:4]]))
        elif f[0] == "f":
            idx = []
            for a in f[1:]:
                p = (a + "//").split("/")[:3]
                ii = []
                for q, lst in zip(p, (v, vt, vn)):
                    k = int(q, 0) if q else 0
                    ii.append(k + len(lst) if k < 0 else k)
                idx.append(ii)
            for i in range(1, len(idx) - 1):
                tris.append([idx[0], idx[i], idx[i + 1]])
    tris = np.array(tris)
    assert pos.shape[0] == len(tris)
    assert same(pos, np.stack([np.stack([v[c[0]] for c in t]) for t in tris]))
    assert same(uv, np.stack([np.stack([vt[c[1]] for c in t]) for t in tris]))
    want_n = np.stack([np.stack([vn[c[2]] for c in t]) for t in tris])
    fn = M.with_face_normals(pos)
    zero = (want_n == 0).all(axis=2)
    want_n[zero] = fn[zero]
    assert same(nrm, want_n)


def test_parse_int_base_zero():
    # "010" is octal 8 in Go's base-0 ParseInt; "08" is an error (index 0)
    obj = b"v 1 0 0\nv 2 0 0\nv 3 0 0\nv 4 0 0\nv 5 0 0\nv 6 0 0\nv 7 0 0\nv 8 0 0\nv 9 0 0\n" \
          b"f 010 0b1 0x2\nf 0o3 1_0 1\n"
    with pytest.raises(ValueError):
        O.load_obj(obj)  # "1_0" is valid (10) but out of range: a panic in Go
    pos, _, _ = O.load_obj(b"v 1 0 0\nv 2 0 0\nv 3 0 0\nv 4 0 0\nv 5 0 0\nv 6 0 0\nv 7 0 0\nv 8 0 0\n"
                           b"f 010 0b1 0x2\nf 08 1 2\n")
    assert same(pos[0, :, 0], [8, 1, 2])
    assert same(pos[1, :, 0], [0, 1, 2])  # "08": error -> index 0 -> the zero vertex


def test_rejected_inputs():
    with pytest.raises(ValueError):
        O.load_stl(b"short")
    with pytest.raises(ValueError):
        O.load_stl(b"solid x\nvertex 1 2 3\nvertex 4 5 6\n" + b" " * 80)  # 2 vertices
    with pytest.raises(ValueError):
        O.load_obj(b"v 1 2 3\nf 1 2 3\n")  # index out of range
    with pytest.raises(ValueError):
        O.load_obj(b"v 1 2\n")  # f[2] out of range
    pos, _, _ = O.load_obj(b"v 1 2 3\nf 1 1\n")  # two corners: no triangle
    assert pos.shape[0] == 0
