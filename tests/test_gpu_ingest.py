"""GPU parity of mesh ingest (fgl_mesh_load_stl / fgl_mesh_load_obj) against
the oracle's LoadSTL / LoadOBJ: bit-exact positions, normals and texture
coordinates (NaNs of any payload counted equal), the same inputs rejected.
Each case prints its name before touching the GPU."""
import sys
import time

import numpy as np
import pytest

import ingest_cases as C
import oracle as O
from fauxgl_amd import _native as N
from fauxgl_amd import api as A
from fauxgl_amd import meshes as M
from test_oracle_mesh import same

pytestmark = pytest.mark.gpu


def _say(name):
    print(f"[case] {name}", flush=True)
    sys.stdout.flush()


def check_stl(ctx, data):
    pos, nrm = O.load_stl(data)
    dm = A.LoadSTL(ctx, data)
    got = dm.Download().Triangles
    dm.close()
    if pos.shape[0] == 0:
        assert got is None
        return
    assert same(got.position, pos) and same(got.normal, nrm)


@pytest.mark.parametrize("name", ["cube.stl", "hello.stl"])
def test_example_stl(name):
    _say(f"ingest {name}")
    check_stl(A.NewContext(8, 8), C.fixture(name))


@pytest.mark.parametrize("seed", [1, 2])
def test_binary_stl(seed):
    _say(f"ingest binary stl seed {seed}")
    _, data = C.random_binary_stl(seed, n=3000)
    check_stl(A.NewContext(8, 8), data)


def test_ascii_stl_tokens():
    _say("ingest ascii stl tokens")
    data, _, _ = C.odd_ascii_stl()
    check_stl(A.NewContext(8, 8), data)


@pytest.mark.parametrize("seed", [3, 4, 5])
def test_obj(seed):
    _say(f"ingest obj seed {seed}")
    data = C.random_obj(seed, nv=200, nf=400)
    pos, nrm, uv = O.load_obj(data)
    dm = A.LoadOBJ(A.NewContext(8, 8), data)
    got = dm.Download().Triangles
    assert same(got.position, pos) and same(got.normal, nrm) and same(got.texture, uv)
    dm.close()


def test_example_obj():
    _say("ingest square.obj")
    data = C.fixture("square.obj")
    pos, nrm, uv = O.load_obj(data)
    got = A.LoadOBJ(A.NewContext(8, 8), data).Download().Triangles
    assert same(got.position, pos) and same(got.normal, nrm) and same(got.texture, uv)


def test_rejected_inputs():
    _say("ingest rejected inputs")
    ctx = A.NewContext(8, 8)
    for bad in (b"short", b"solid x\nvertex 1 2 3\nvertex 4 5 6\n" + b" " * 80):
        with pytest.raises(N.FglError):
            A.LoadSTL(ctx, bad)
    for bad in (b"v 1 2 3\nf 1 2 3\n", b"v 1 2\n"):
        with pytest.raises(N.FglError):
            A.LoadOBJ(ctx, bad)


def test_large_binary_stl():
    """The 871,306-triangle noise sphere written as binary STL (43.6 MB):
    device decode against the oracle, with the time of each."""
    _say("ingest large binary stl")
    p = M.noise_sphere_mesh().Triangles.position
    data = C.binary_stl(p.astype(np.float32))
    t0 = time.perf_counter()
    pos, nrm = O.load_stl(data)
    t_cpu = time.perf_counter() - t0
    ctx = A.NewContext(8, 8)
    A.LoadSTL(ctx, data).close()  # warm-up
    t0 = time.perf_counter()
    dm = A.LoadSTL(ctx, data)
    t_gpu = time.perf_counter() - t0
    got = dm.Download().Triangles
    print(f"binary STL, {len(p)} triangles: device load {t_gpu * 1e3:.1f} ms (upload + decode), "
          f"oracle {t_cpu * 1e3:.0f} ms", flush=True)
    assert same(got.position, pos) and same(got.normal, nrm)
