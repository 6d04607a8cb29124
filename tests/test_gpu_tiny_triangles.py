"""Sub-pixel and pixel-sized triangles against the oracle.

The vertex stage drops a raster triangle whose small bounding box has no
pixel center that the reference's inside test (context.go:184-228) can
accept, by a margin above the rounding of any order the walk evaluates the
edge functions in.  These scenes put many such triangles next to ones that
do cover centers, including triangles whose vertices and edges lie exactly
on pixel centers (the b == 0 case the reference counts as inside), and
check the frame and RasterizeInfo bit for bit against the oracle.

Each case prints its name before touching the GPU."""
import sys

import numpy as np
import pytest

from fauxgl_amd import api as A
from fauxgl_amd import meshes as M
from helpers import compare, run_gpu, run_oracle

pytestmark = pytest.mark.gpu

ORTHO = A.Identity()


def _say(name):
    print(f"[case] {name}", flush=True)
    sys.stdout.flush()


def tiny(n, seed, W, px_lo, px_hi, snap=False):
    """n triangles of extent px_lo..px_hi pixels anywhere on a W x W screen
    (NDC with the identity matrix: screen x = (ndc + 1) * W / 2).  snap puts
    every vertex on a pixel center or corner (exact in binary)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    c = rng.uniform(0, W, size=(n, 1, 2))
    d = rng.uniform(px_lo, px_hi, size=(n, 1, 1)) * rng.uniform(-1, 1, size=(n, 3, 2))
    xy = c + d
    if snap:
        xy = np.round(xy * 2) / 2
    pos = np.empty((n, 3, 3))
    pos[..., :2] = xy / (W / 2) - 1
    pos[..., 2] = rng.uniform(-0.9, 0.9, size=(n, 1))
    col = rng.uniform(0, 1, size=(n, 3, 4))
    col[..., 3] = 1
    return A.Prims(pos, M.with_face_normals(pos), None, col)


SHADER = A.NewPhongShader(ORTHO, A.V(0.2, 0.4, 1).Normalize(), A.V(0, 0, 3))


@pytest.mark.parametrize("case", [
    ("sub-pixel", 20000, 0.02, 0.6, False, 1),
    ("pixel-sized", 20000, 0.3, 2.0, False, 2),
    ("on pixel centers", 20000, 0.5, 1.5, True, 3),
    ("mixed", 30000, 0.01, 4.0, False, 4),
])
def test_tiny_triangles_match_the_oracle(case):
    name, n, lo, hi, snap, seed = case
    W = 256
    _say(f"tiny triangles: {name}")
    for cull in (A.CullNone, A.CullBack):
        ops = [("shader", SHADER), ("set", {"Cull": cull}), ("clear_color", A.HexColor("#FFF8E3")),
               ("tris", tiny(n, seed, W, lo, hi, snap))]
        ref = run_oracle(W, W, ops)
        got = run_gpu(W, W, ops)
        res = compare(ref, got, color_tol=0)
        assert res["ok"], (name, cull, res)
        assert ref[2][0][0] > 0  # some triangles do cover centers


def test_tiny_triangles_at_16x():
    """Small triangles at large screen coordinates (the margin grows with
    them): a 4096 x 2304 frame."""
    W, H = 4096, 2304
    _say("tiny triangles: 4096x2304")
    rng = np.random.Generator(np.random.PCG64(11))
    n = 40000
    c = np.stack([rng.uniform(0, W, n), rng.uniform(0, H, n)], axis=1)[:, None, :]
    xy = c + rng.uniform(0.05, 2.5, size=(n, 1, 1)) * rng.uniform(-1, 1, size=(n, 3, 2))
    pos = np.empty((n, 3, 3))
    pos[..., 0] = xy[..., 0] / (W / 2) - 1
    pos[..., 1] = xy[..., 1] / (H / 2) - 1
    pos[..., 2] = rng.uniform(-0.9, 0.9, size=(n, 1))
    ops = [("shader", SHADER), ("set", {"Cull": A.CullNone}),
           ("tris", A.Prims(pos, M.with_face_normals(pos)))]
    ref = run_oracle(W, H, ops)
    got = run_gpu(W, H, ops)
    res = compare(ref, got, color_tol=0)
    assert res["ok"], res
