"""Batched views whose draws need different kernels: a replay launches each
pipeline stage once per group of draws with one tile shape and kernel
variant (primitive kind, wireframe, shader kind, texture), in rounds when
views share a context.  These views mix Phong, solid and textured shaders,
wireframe, a mesh with lines, three tile shapes and a shared context; every
replay must equal drawing each view on its own (the path the parity tests
check against the oracle).  Each case prints its name before touching the
GPU."""
import math
import sys

import numpy as np

from fauxgl_amd import api as A
from fauxgl_amd import meshes as M

import pytest

pytestmark = pytest.mark.gpu


def _say(name):
    print(f"[case] {name}", flush=True)
    sys.stdout.flush()


def sphere(ctx, ntri, seed):
    dm = ctx.Upload(A.Mesh(M.noise_sphere_prims(ntri=ntri, detail=5, amplitude=0.15, seed=seed)))
    dm.BiUnitCube()
    dm.SmoothNormalsThreshold(math.radians(30.0))
    return dm


def soup_with_lines(ctx):
    tris = M.triangle_soup(n=3000, seed=5, lmin=0.01, lmax=0.3).Triangles
    rng = np.random.Generator(np.random.PCG64(6))
    lp = rng.uniform(-1, 1, size=(400, 2, 3))
    lines = A.Prims(lp, None, None, rng.uniform(0, 1, size=(400, 2, 4)), nv=2)
    return ctx.Upload(A.Mesh(tris, lines))


def views(home):
    """(context, mesh, configure(ctx)) per view, in batch order."""
    s1 = sphere(home, 8000, 31)
    s2 = sphere(home, 2500, 32)
    soup = soup_with_lines(home)
    tex = A.ImageTexture(M.noise_texture(64, 3), "NRGBA")
    eye1, eye2 = A.V(3.2, 1.0, 0.5), A.V(-1.0, 0.6, 3.4)

    def phong(w, h, eye, tile=None, wire=False):
        def f(c):
            c.Shader = M.batch_shader(w, h, eye)
            c.Wireframe = wire
            if tile:
                c.SetTileShape(*tile)
        return f

    def solid(w, h, eye, tile=None):
        def f(c):
            c.Shader = A.NewSolidColorShader(M.camera(w, h, eye=eye, center=A.V(0, 0, 0)), A.HexColor("#468966"))
            c.Cull = A.CullNone
            if tile:
                c.SetTileShape(*tile)
        return f

    def textured(w, h, eye, tile=None):
        def f(c):
            c.Shader = A.NewTextureShader(M.camera(w, h, eye=eye, center=A.V(0, 0, 0)), tex)
            c.Cull = A.CullNone
            if tile:
                c.SetTileShape(*tile)
        return f

    shared = (96, 80)
    out = [
        ((160, 120), s1, phong(160, 120, eye1)),
        ((160, 120), s2, phong(160, 120, eye2, tile=(64, 32))),
        ((128, 128), s1, phong(128, 128, eye2, wire=True)),
        ((200, 100), soup, solid(200, 100, eye1)),
        ((200, 100), soup, textured(200, 100, eye2, tile=(32, 16))),
        (shared, s2, phong(96, 80, eye1)),
        (shared, soup, solid(96, 80, eye2)),  # same context: drawn after the view above
        ((64, 64), s1, textured(64, 64, eye1)),
    ]
    return out


def test_mixed_kernel_variants_equal_single_draws():
    _say("batch of mixed kernel variants, tile shapes and a shared context, 3 replays")
    home = A.NewContext(8, 8)
    vs = views(home)
    # each view drawn on its own (a context of its own per view)
    want = []
    for (w, h), dm, conf in vs:
        c = A.NewContext(w, h)
        conf(c)
        c.ClearColorBufferWith(M.BACKGROUND)
        c.ClearDepthBuffer()
        info = c.DrawMesh(dm)
        want.append((info, c.Image(), c.DepthBuffer))
    # the batch: the two views of size `shared` share one context
    b = A.Batch()
    ctxs, shared_ctx = [], None
    for k, ((w, h), dm, conf) in enumerate(vs):
        if (w, h) == (96, 80):
            if shared_ctx is None:
                shared_ctx = A.NewContext(w, h)
            c = shared_ctx
        else:
            c = A.NewContext(w, h)
        conf(c)
        b.Add(c, dm, clear_color=M.BACKGROUND)
        ctxs.append(c)
    last_on = {id(c): k for k, c in enumerate(ctxs)}
    for rep in range(3):  # direct draws and recording, then two replays
        infos = b.Run()
        for k, c in enumerate(ctxs):
            winfo, wimg, wdepth = want[k]
            assert infos[k] == winfo, (rep, k)
            if last_on[id(c)] == k:  # a shared context ends with its last view
                assert np.array_equal(c.Image(), wimg), (rep, k)
                assert np.array_equal(c.DepthBuffer.view(np.uint64), wdepth.view(np.uint64)), (rep, k)
    b.close()
