"""fgl_depth_image_noalloc_async, the allocation-free variant of
fgl_depth_image (DepthImage, context.go:87-117): through the C ABI, its
Gray16 values equal fgl_depth_image's bit for bit on two calls in a row, the
second on a smaller frame (the per-GPU scratch keeps the first call's
partials, which the second must not see), and again with a pending clear.

Each case prints its name before touching the GPU."""
import ctypes
import sys

import numpy as np
import pytest
import torch

from fauxgl_amd import _native as N
from fauxgl_amd import api as A
from fauxgl_amd import meshes as M

pytestmark = pytest.mark.gpu


def _say(name):
    print(f"[case] {name}", flush=True)
    sys.stdout.flush()


def render(W, H, ntri, seed):
    ctx = A.NewContext(W, H)
    ctx.ClearColorBufferWith(M.BACKGROUND)
    ctx.ClearDepthBuffer()
    ctx.Shader = M.phong_shader(W, H)
    ctx.DrawMesh(M.noise_sphere_mesh(ntri=ntri, detail=5, amplitude=0.15, seed=seed))
    return ctx


def noalloc_depth_image(ctx):
    W, H = ctx.Width, ctx.Height
    out = torch.full((H * W,), 0x5a5a, dtype=torch.int32, device="cuda:0").to(torch.int16)
    torch.cuda.synchronize()
    lib = N.load()
    rc = lib.fgl_depth_image_noalloc_async(ctx._ptr(), ctypes.c_void_p(out.data_ptr()))
    assert rc == 0, lib.fgl_last_error(ctx._ptr())
    ctx.Sync()
    return out.cpu().numpy().view(np.uint16).reshape(H, W)


def test_noalloc_depth_image_matches_twice_then_smaller():
    _say("noalloc depth image: 640x480 then 200x120")
    big = render(640, 480, 20000, 3)
    small = render(200, 120, 3000, 4)
    for ctx in (big, small):
        want = ctx.DepthImage()
        got = noalloc_depth_image(ctx)
        assert got.shape == want.shape
        assert np.array_equal(got, want)
        assert np.count_nonzero(got != 0xffff) > 0


def test_noalloc_depth_image_applies_a_pending_clear():
    _say("noalloc depth image: pending clear")
    ctx = render(160, 96, 3000, 5)
    ctx.ClearDepthBufferWith(0.25)
    got = noalloc_depth_image(ctx)
    want = ctx.DepthImage()
    assert np.array_equal(got, want)


def test_noalloc_depth_image_rejects_a_null_buffer():
    _say("noalloc depth image: null buffer")
    ctx = A.NewContext(8, 8)
    lib = N.load()
    assert lib.fgl_depth_image_noalloc_async(ctx._ptr(), None) == N.FGL_E_INVALID
