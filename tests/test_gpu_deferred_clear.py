"""Deferred clears per block: a draw that consumes a pending clear leaves the
tiles no primitive reaches unwritten (marked per 32 x 16 block) instead of
filling them, and every later user of the framebuffer honours the marks.

- Draw sequences whose later draws land in tiles an earlier draw left at the
  clear values, with color-only and depth-only clears between them: every
  prefix of the sequence is bit-exact against the oracle (context.go:119-145,
  context.go:184-273).
- Readers of the whole buffer (Image, DepthBuffer, DepthImage, the
  allocation-free DepthImage) see the clear values in marked blocks.
- Writers of a whole channel (SetColorBuffer / SetDepthBuffer, a resolve or network
  output into a context with marked blocks) are not overwritten later by the
  clear values.
- The autotuner's timing draws leave the frame and its marks as they were.

Each case prints its name before touching the GPU."""
import ctypes
import sys

import numpy as np
import pytest
import torch

from fauxgl_amd import _native as N
from fauxgl_amd import api as A
from fauxgl_amd import meshes as M
from helpers import compare, run_oracle

pytestmark = pytest.mark.gpu

ORTHO = A.Identity()


def _say(name):
    print(f"[case] {name}", flush=True)
    sys.stdout.flush()


def cluster(n, seed, x0, x1, y0=-1.0, y1=1.0, size=0.08):
    """n small triangles with centres uniform in [x0, x1] x [y0, y1] (NDC)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    c = np.stack([rng.uniform(x0, x1, n), rng.uniform(y0, y1, n), rng.uniform(-0.9, 0.9, n)], axis=1)[:, None, :]
    pos = c + rng.uniform(-size, size, size=(n, 3, 3))
    col = rng.uniform(0, 1, size=(n, 3, 4))
    col[..., 3] = rng.uniform(0.3, 1.0, size=(n, 3))
    return A.Prims(pos, M.with_face_normals(pos), None, col)


SHADER = A.NewPhongShader(ORTHO, A.V(0.2, 0.4, 1).Normalize(), A.V(0, 0, 3))


def sequence():
    left = cluster(150, 1, -1.0, -0.5)
    right = cluster(150, 2, 0.4, 1.0)
    middle = cluster(150, 3, -0.2, 0.2, -0.5, 0.5)
    top = cluster(100, 4, -1.0, 1.0, 0.7, 1.0)
    return [("shader", SHADER), ("set", {"Cull": A.CullNone}),
            ("clear_color", A.HexColor("#FFF8E3")), ("clear_depth", A.MaxFloat64), ("tris", left),
            ("tris", right),                                    # lands in blocks the first draw left marked
            ("clear_color", A.Color(0.1, 0.2, 0.3, 1.0)), ("tris", middle),   # color-only clear
            ("clear_depth", 0.5), ("tris", top),                 # depth-only clear
            ("tris", left)]


@pytest.mark.parametrize("tile", [(32, 16), (64, 32), (128, 64)])
def test_draws_over_blocks_left_at_the_clear_values(tile):
    W, H = 300, 170
    ops = sequence()
    for k in range(len(ops)):
        if ops[k][0] != "tris":
            continue
        _say(f"deferred clear: tile {tile}, prefix through op {k}")
        ref = run_oracle(W, H, ops[:k + 1])
        ctx = A.NewContext(W, H)
        ctx.SetTileShape(*tile)
        infos = []
        for op, val in ops[:k + 1]:
            if op == "clear_color":
                ctx.ClearColorBufferWith(val)
            elif op == "clear_depth":
                ctx.ClearDepthBufferWith(val)
            elif op == "set":
                for kk, vv in val.items():
                    setattr(ctx, kk, vv)
            elif op == "shader":
                ctx.Shader = val
            else:
                infos.append(tuple(ctx.DrawTriangles(val)))
        got = (ctx.Image(), ctx.DepthBuffer, infos)
        res = compare(ref, got, color_tol=0)
        assert res["ok"], (tile, k, res)
        assert np.array_equal(ref[3].depth_image(), ctx.DepthImage()), (tile, k)


def frame_with_marks(W=256, H=128):
    """A context whose last draw consumed a clear and left most blocks marked."""
    ctx = A.NewContext(W, H)
    ctx.SetTileShape(32, 16)
    ctx.Shader = SHADER
    ctx.Cull = A.CullNone
    ctx.ClearColorBufferWith(A.HexColor("#FFF8E3"))
    ctx.ClearDepthBufferWith(0.75)
    ctx.DrawTriangles(cluster(60, 7, -0.9, -0.6, -0.9, -0.6))
    return ctx


def test_noalloc_depth_image_fills_marked_blocks():
    _say("deferred clear: allocation-free DepthImage")
    ctx = frame_with_marks()
    out = torch.zeros(ctx.Width * ctx.Height, dtype=torch.int16, device="cuda:0")
    torch.cuda.synchronize()
    lib = N.load()
    assert lib.fgl_depth_image_noalloc_async(ctx._ptr(), ctypes.c_void_p(out.data_ptr())) == 0
    ctx.Sync()
    got = out.cpu().numpy().view(np.uint16).reshape(ctx.Height, ctx.Width)
    assert np.array_equal(got, ctx.DepthImage())
    d = ctx.DepthBuffer
    assert np.count_nonzero(d == 0.75) > d.size // 2


@pytest.mark.parametrize("channel", ["color", "depth"])
def test_whole_channel_writes_are_kept(channel):
    _say(f"deferred clear: Set{channel}Buffer over marked blocks")
    ctx = frame_with_marks()
    rng = np.random.Generator(np.random.PCG64(9))
    if channel == "color":
        x = rng.integers(0, 256, size=(ctx.Height, ctx.Width, 4), dtype=np.uint8)
        ctx.SetColorBuffer(x)
        assert np.array_equal(ctx.Image(), x)
    else:
        x = rng.uniform(0, 1, size=(ctx.Height, ctx.Width))
        ctx.SetDepthBuffer(x)
        assert np.array_equal(ctx.DepthBuffer, x)
    # a draw after the write starts its tiles from the written values
    t = cluster(80, 8, -0.2, 0.9)
    ctx.DrawTriangles(t)
    # the reference frame: the same draws on a context whose buffers were
    # materialised before the write (no marks involved)
    alt = A.NewContext(ctx.Width, ctx.Height)
    alt.SetTileShape(32, 16)
    alt.Shader = SHADER
    alt.Cull = A.CullNone
    alt.ClearColorBufferWith(A.HexColor("#FFF8E3"))
    alt.ClearDepthBufferWith(0.75)
    alt.DrawTriangles(cluster(60, 7, -0.9, -0.6, -0.9, -0.6))
    alt.Image()  # materialise every block
    if channel == "color":
        alt.SetColorBuffer(x)
    else:
        alt.SetDepthBuffer(x)
    alt.DrawTriangles(t)
    assert np.array_equal(ctx.Image(), alt.Image())
    assert np.array_equal(ctx.DepthBuffer.view(np.uint64), alt.DepthBuffer.view(np.uint64))


def test_resolve_into_a_context_with_marked_blocks():
    _say("deferred clear: resolve into marked blocks")
    src = A.NewContext(512, 256)
    src.Shader = SHADER
    src.Cull = A.CullNone
    src.ClearColorBufferWith(A.Color(0.3, 0.6, 0.9, 1))
    src.DrawTriangles(cluster(300, 10, -1.0, 1.0))
    dst = frame_with_marks()  # 256 x 128, marked
    fresh = A.NewContext(256, 128)
    src.Resolve(dst, "tent")
    src.Resolve(fresh, "tent")
    assert np.array_equal(dst.Image(), fresh.Image())
    # dst's depth keeps its own marks: the clear depth where nothing was drawn
    assert np.count_nonzero(dst.DepthBuffer == 0.75) > 0


def test_autotune_keeps_frame_and_marks():
    _say("deferred clear: autotune keeps the marks")
    ctx = frame_with_marks()
    before_img, before_depth = ctx.Image(), ctx.DepthBuffer  # materialises
    ctx2 = frame_with_marks()
    mesh = ctx2.Upload(A.Mesh(cluster(200, 11, -1.0, 1.0), None))
    ctx2.Autotune(mesh, reps=1)
    ctx2.SetTileShape(32, 16)
    assert np.array_equal(ctx2.Image(), before_img)
    assert np.array_equal(ctx2.DepthBuffer.view(np.uint64), before_depth.view(np.uint64))
    mesh.close()


@pytest.mark.parametrize("size", [(256, 128), (75, 43)])
def test_4x_resolve_reads_marked_blocks(size):
    """The 4x tent resolve takes the clear color for source blocks a draw left
    unwritten: the same result as resolving the materialised frame."""
    w, h = size
    _say(f"deferred clear: 4x resolve of marked blocks to {w}x{h}")

    def source():
        src = A.NewContext(4 * w, 4 * h)
        src.SetTileShape(32, 16)
        src.Shader = SHADER
        src.Cull = A.CullNone
        src.ClearColorBufferWith(A.Color(0.3, 0.6, 0.9, 0.5))
        src.DrawTriangles(cluster(200, 12, -0.95, 0.1, -0.9, 0.2, size=0.05))
        return src

    marked, plain = source(), source()
    plain.Image()  # every block written
    a, b = A.NewContext(w, h), A.NewContext(w, h)
    marked.Resolve(a, "tent")
    plain.Resolve(b, "tent")
    assert np.array_equal(a.Image(), b.Image())
    assert np.array_equal(marked.Image(), plain.Image())
