"""Tile shapes and the autotuner on the GPU.  The draw kernels are compiled
for several tile shapes; the shape only changes how the framebuffer is cut
into tiles, so every shape must give bit-identical color, depth and
RasterizeInfo (the 32 x 32 default is the one the parity tests check against
the oracle).  The autotuner must leave the framebuffer as it found it and
its table must survive a save / load.  Each case prints its name before
touching the GPU."""
import math
import sys

import numpy as np
import pytest

from fauxgl_amd import api as A
from fauxgl_amd import meshes as M

pytestmark = pytest.mark.gpu

BG = A.HexColor("FFF8E3")


def _say(name):
    print(f"[case] {name}", flush=True)
    sys.stdout.flush()


def _tris(n, seed, size=None, colors=False, uvs=False):
    rng = np.random.Generator(np.random.PCG64(seed))
    c = rng.uniform(-1.2, 1.2, size=(n, 1, 3))
    c[..., 2] = rng.uniform(-0.9, 0.9, size=(n, 1))
    pos = c + rng.uniform(-size, size, size=(n, 3, 3)) if size else rng.uniform(-1.2, 1.2, size=(n, 3, 3))
    nrm = M.with_face_normals(pos)
    col = rng.uniform(0, 1, size=(n, 3, 4)) if colors else None
    uv = rng.uniform(-0.5, 1.5, size=(n, 3, 2)) if uvs else None
    return A.Prims(pos, nrm, uv, col)


def _scenes():
    W, H = 300, 200
    cam = M.camera(W, H)
    out = []
    # Phong sphere and cube under the standard camera (clipped triangles)
    sh = M.phong_shader(W, H)
    out.append(("phong_sphere_cube", W, H, sh, {}, M.sphere_and_cube_mesh(4)))
    # textured, blended soup with a depth bias
    rng = np.random.Generator(np.random.PCG64(5))
    tex = A.ImageTexture(rng.integers(0, 256, size=(32, 32, 4), dtype=np.uint8), "NRGBA")
    out.append(("texture_blend_soup", W, H, A.NewTextureShader(cam, tex), {"DepthBias": 1e-4},
                A.Mesh(_tris(3000, 6, size=0.3, uvs=True))))
    # vertex colors (ObjectColor = Discard), no culling
    sh2 = A.NewPhongShader(A.Orthographic(-1, 1, -1, 1, -1, 1), A.V(0.2, 0.3, 1).Normalize(), A.V(0, 0, 2))
    out.append(("vertex_colors_cullnone", 170, 130, sh2, {"Cull": A.CullNone},
                A.Mesh(_tris(2500, 7, size=0.25, colors=True))))
    # wireframe (side records)
    sh3 = A.NewPhongShader(cam, A.V(0.75, 0.5, 1).Normalize(), M.EYE)
    out.append(("wireframe_sphere", 257, 131, sh3, {"Wireframe": True, "LineWidth": 1.5},
                M.sphere_and_cube_mesh(2)))
    return out


def _draw(W, H, shader, state, mesh, shape):
    ctx = A.NewContext(W, H)
    ctx.SetTileShape(*shape)
    for k, v in state.items():
        setattr(ctx, k, v)
    ctx.Shader = shader
    ctx.ClearColorBufferWith(BG)
    ctx.ClearDepthBuffer()
    dm = ctx.Upload(mesh)
    info = tuple(ctx.DrawMesh(dm))
    info2 = tuple(ctx.DrawMesh(dm))  # a second draw over the first
    return ctx.Image(), ctx.DepthBuffer, (info, info2)


def _same(a, b):
    return (np.array_equal(a[0], b[0]) and np.array_equal(a[1].view(np.uint64), b[1].view(np.uint64))
            and a[2] == b[2])


@pytest.mark.parametrize("scene", range(4))
def test_every_shape_gives_the_same_frame(scene):
    base = None
    for shape in A.TileShapes():
        # a fresh scene per context: a texture belongs to the context it was
        # first drawn with
        name, W, H, sh, st, mesh = _scenes()[scene]
        _say(f"{name} tile {shape}")
        got = _draw(W, H, sh, st, mesh, shape)
        if base is None:
            base = got
            assert got[2][0][0] > 0
        else:
            assert _same(base, got), (name, shape)


def test_autotune_leaves_the_frame_and_fills_the_table(tmp_path):
    _say("autotune 320x240 noise sphere")
    A.ClearTuning()
    W, H = 320, 240
    ctx = A.NewContext(W, H)
    ctx.Shader = M.phong_shader(W, H)
    prims = M.noise_sphere_prims(ntri=20000, detail=5)
    dm = ctx.Upload(A.Mesh(prims))
    dm.BiUnitCube()
    ctx.ClearColorBufferWith(BG)
    ctx.ClearDepthBuffer()
    want_info = tuple(ctx.DrawMesh(dm))
    img, dep = ctx.Image(), ctx.DepthBuffer
    assert ctx.TileShape(20000) == (32, 32)
    shape = ctx.Autotune(dm, reps=2)
    assert shape in A.TileShapes()
    # the frame is untouched, and the table now answers for this workload
    assert np.array_equal(ctx.Image(), img) and np.array_equal(ctx.DepthBuffer.view(np.uint64), dep.view(np.uint64))
    assert ctx.TileShape(20000) == shape
    assert ctx.TileShape(100) == (32, 32)  # another size class is not tuned
    # a draw with the tuned shape gives the same frame
    ctx.ClearColorBufferWith(BG)
    ctx.ClearDepthBuffer()
    assert tuple(ctx.DrawMesh(dm)) == want_info
    assert np.array_equal(ctx.Image(), img) and np.array_equal(ctx.DepthBuffer.view(np.uint64), dep.view(np.uint64))
    # save / clear / load
    path = tmp_path / "tune.txt"
    A.SaveTuning(path)
    A.ClearTuning()
    assert ctx.TileShape(20000) == (32, 32)
    A.LoadTuning(path)
    assert ctx.TileShape(20000) == shape
    # a context's own shape wins over the table
    other = (64, 64) if shape != (64, 64) else (32, 16)
    ctx.SetTileShape(*other)
    assert ctx.TileShape(20000) == other
    ctx.SetTileShape()
    assert ctx.TileShape(20000) == shape
    with pytest.raises(Exception):
        ctx.SetTileShape(48, 48)
    A.ClearTuning()


def test_batch_with_mixed_shapes():
    _say("batch of 4 views with different tile shapes")
    shapes = A.TileShapes()
    prims = M.noise_sphere_prims(ntri=8000, detail=5)
    home = A.NewContext(8, 8)
    dm = home.Upload(A.Mesh(prims))
    dm.BiUnitCube()
    dm.SmoothNormalsThreshold(math.radians(30.0))
    views = [(200, 150), (96, 160), (257, 99), (128, 128)]
    want, ctxs = [], []
    b = A.Batch()
    for i, (w, h) in enumerate(views):
        eye = A.V(3.5 * math.cos(i), 1.0, 3.5 * math.sin(i))
        c = A.NewContext(w, h)
        c.Shader = M.batch_shader(w, h, eye)
        c.ClearColorBufferWith(M.BACKGROUND)
        c.ClearDepthBuffer()
        info = c.DrawMesh(dm)
        want.append((tuple(info), c.Image(), c.DepthBuffer))
        c2 = A.NewContext(w, h)
        c2.SetTileShape(*shapes[i % len(shapes)])
        c2.Shader = M.batch_shader(w, h, eye)
        b.Add(c2, dm, clear_color=M.BACKGROUND)
        ctxs.append(c2)
    for rep in range(2):
        infos = b.Run()
        for i, c in enumerate(ctxs):
            assert tuple(infos[i]) == want[i][0], (rep, i)
            assert np.array_equal(c.Image(), want[i][1]), (rep, i)
            assert np.array_equal(c.DepthBuffer.view(np.uint64), want[i][2].view(np.uint64)), (rep, i)
