"""GPU parity of device-side mesh preparation (fgl_mesh_*: BoundingBox,
Transform, FitInside, SmoothNormals, SmoothNormalsThreshold) against the
oracle: bit-exact (NaNs of any payload counted equal).  Each case prints
its name before touching the GPU."""
import math
import sys
import time

import numpy as np
import pytest

import oracle as O
from fauxgl_amd import api as A
from fauxgl_amd import meshes as M
from test_oracle_mesh import same, shared_mesh

pytestmark = pytest.mark.gpu


def _say(name):
    print(f"[case] {name}", flush=True)
    sys.stdout.flush()


def upload(tpos, tnrm, lpos=None, lnrm=None):
    ctx = A.NewContext(8, 8)
    tris = A.Prims(tpos, tnrm) if tpos is not None else None
    lines = A.Prims(lpos, lnrm, nv=2) if lpos is not None else None
    return ctx, ctx.Upload(A.Mesh(tris, lines))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_bounding_box_transform_fit(seed):
    _say(f"mesh box/transform/fit seed {seed}")
    rng = np.random.Generator(np.random.PCG64(seed))
    t = rng.normal(size=(300, 3, 3)) * 2 + 1
    t[:2] = 0.0  # leading triangles at the origin do not count (Box.Extend)
    tn = rng.normal(size=(300, 3, 3))
    l = rng.normal(size=(7, 2, 3)) * 3
    ln = rng.normal(size=(7, 2, 3))
    ctx, dm = upload(t, tn, l, ln)
    lo, hi = dm.BoundingBox()
    wlo, whi = O.mesh_bounding_box(t, l)
    assert same(list(lo) + list(hi), list(wlo) + list(whi))
    m = rng.normal(size=16)
    dm.Transform(list(m))
    tp, tnn = O.mesh_transform(t, tn, m)
    lp, lnn = O.mesh_transform(l, ln, m)
    got = dm.Download()
    assert same(got.Triangles.position, tp) and same(got.Triangles.normal, tnn)
    assert same(got.Lines.position, lp) and same(got.Lines.normal, lnn)
    mat = dm.BiUnitCube()
    wtp, wtn, wlp, wln, wm = O.mesh_fit_inside(tp, tnn, (-1, -1, -1), (1, 1, 1), (0.5, 0.5, 0.5), lp, lnn)
    got = dm.Download()
    assert same(list(mat), wm)
    assert same(got.Triangles.position, wtp) and same(got.Triangles.normal, wtn)
    assert same(got.Lines.position, wlp) and same(got.Lines.normal, wln)
    dm.close()


def test_box_of_origin_mesh_and_missing_normals():
    _say("mesh empty box, missing normals")
    z = np.zeros((3, 3, 3))
    ctx, dm = upload(z, None)
    lo, hi = dm.BoundingBox()
    assert list(lo) == [0, 0, 0] and list(hi) == [0, 0, 0]
    # a zero normal through MulDirection normalizes to NaN, as in Go
    dm.Transform(list(np.eye(4).reshape(-1)))
    assert np.isnan(dm.Download().Triangles.normal).all()
    dm.close()


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_smooth_normals(seed):
    _say(f"mesh smooth seed {seed}")
    pos, nrm = shared_mesh(seed, ntri=2000)
    for deg in (None, 0.0, 30.0, 120.0):
        ctx, dm = upload(pos, nrm)
        if deg is None:
            dm.SmoothNormals()
            want = O.mesh_smooth_normals(pos, nrm)
        else:
            thr = math.cos(math.radians(deg))
            dm.SmoothNormalsThresholdCos(thr)
            want = O.mesh_smooth_normals(pos, nrm, thr)
        assert same(dm.Download().Triangles.normal, want), (seed, deg)
        dm.close()


def test_noise_sphere_preparation():
    """The 871,306-triangle workload's own preparation (BiUnitCube, then
    SmoothNormalsThreshold(30 degrees)) on the GPU equals the generator's
    numpy path bit for bit."""
    _say("mesh noise sphere preparation")
    raw = M.noise_sphere_prims()
    t0 = time.perf_counter()
    want = M.smooth_normals_threshold(M.bi_unit_cube(raw), math.radians(30.0))
    t_cpu = time.perf_counter() - t0
    ctx, dm = upload(raw.position, raw.normal)
    dm.BiUnitCube()  # warm-up of the kernels and the host path
    ctx2, dm2 = upload(raw.position, raw.normal)
    t0 = time.perf_counter()
    dm2.BiUnitCube()
    dm2.SmoothNormalsThreshold(math.radians(30.0))
    t_gpu = time.perf_counter() - t0
    got = dm2.Download().Triangles
    print(f"noise sphere preparation: GPU {t_gpu * 1e3:.1f} ms, numpy {t_cpu * 1e3:.0f} ms", flush=True)
    assert same(got.position, want.position)
    assert same(got.normal, want.normal)
    dm.close()
    dm2.close()
