"""Batched views (fgl_batch): a batch of frames recorded into one HIP graph
and replayed gives, on every replay, the same pixels, depth and
RasterizeInfo as drawing each view on its own (the path the parity tests
check against the oracle).  Each case prints its name before touching the
GPU."""
import math
import sys

import numpy as np
import pytest

from fauxgl_amd import api as A
from fauxgl_amd import meshes as M

pytestmark = pytest.mark.gpu


def _say(name):
    print(f"[case] {name}", flush=True)
    sys.stdout.flush()


def small_workload():
    return M.batch_workload(nmesh=5, views_per_mesh=3, seed=9, nmin=200, nmax=30000,
                            resolutions=[(64, 64), (128, 96), (96, 160)])


def prepare(ctx, prims):
    dm = ctx.Upload(A.Mesh(prims))
    dm.BiUnitCube()
    dm.SmoothNormalsThreshold(math.radians(30.0))
    return dm


def test_batch_equals_single_draws():
    _say("batch of 15 views, 3 replays")
    meshes, views = small_workload()
    home = A.NewContext(8, 8)
    dms = [prepare(home, p) for p in meshes]
    # each view drawn on its own
    want = []
    for m, eye, w, h in views:
        c = A.NewContext(w, h)
        c.ClearColorBufferWith(M.BACKGROUND)
        c.ClearDepthBuffer()
        c.Shader = M.batch_shader(w, h, eye)
        info = c.DrawMesh(dms[m])
        want.append((info, c.Image(), c.DepthBuffer))
    # the same views as one batch
    ctxs = []
    b = A.Batch()
    for m, eye, w, h in views:
        c = A.NewContext(w, h)
        c.Shader = M.batch_shader(w, h, eye)
        b.Add(c, dms[m], clear_color=M.BACKGROUND)
        ctxs.append(c)
    for rep in range(3):  # eager + record, then two replays
        infos = b.Run()
        for i, c in enumerate(ctxs):
            winfo, wimg, wdepth = want[i]
            assert infos[i] == winfo, (rep, i)
            assert np.array_equal(c.Image(), wimg), (rep, i)
            assert np.array_equal(c.DepthBuffer, wdepth), (rep, i)
        # scribble on the framebuffers: the next replay must clear them again
        for c in ctxs:
            c.ClearColorBufferWith(A.Color(1, 0, 0, 1))
            c.Image()
    b.close()


def test_views_sharing_a_context_and_added_later():
    """Two views on one context run in order (the second one's frame is what
    remains); adding a view after a run records the graph again."""
    _say("batch shared context, late add")
    meshes, views = small_workload()
    home = A.NewContext(8, 8)
    dms = [prepare(home, p) for p in meshes[:2]]
    shared = A.NewContext(96, 64)
    shared.Shader = M.batch_shader(96, 64, views[0][1])
    b = A.Batch()
    b.Add(shared, dms[0], clear_color=M.BACKGROUND)
    shared.Shader = M.batch_shader(96, 64, views[1][1])
    b.Add(shared, dms[1], clear_color=M.BACKGROUND)
    other = A.NewContext(64, 64)
    other.Shader = M.batch_shader(64, 64, views[2][1])
    b.Run()
    b.Run()
    b.Add(other, dms[0], clear_color=M.BACKGROUND)
    b.Run()
    infos = b.Run()
    ref = A.NewContext(96, 64)
    ref.ClearColorBufferWith(M.BACKGROUND)
    ref.ClearDepthBuffer()
    ref.Shader = M.batch_shader(96, 64, views[1][1])
    info = ref.DrawMesh(dms[1])
    assert infos[1] == info and np.array_equal(shared.Image(), ref.Image())
    ref2 = A.NewContext(64, 64)
    ref2.ClearColorBufferWith(M.BACKGROUND)
    ref2.ClearDepthBuffer()
    ref2.Shader = M.batch_shader(64, 64, views[2][1])
    assert infos[2] == ref2.DrawMesh(dms[0]) and np.array_equal(other.Image(), ref2.Image())
    b.close()
