"""Batched views (fgl_batch): a batch of frames prepared once and replayed
(one launch per pipeline stage for each group of views; one recorded HIP
graph when a view's mesh is empty) gives, on every replay, the same pixels,
depth and RasterizeInfo as drawing each view on its own (the path the parity
tests check against the oracle).  Each case prints its name before touching
the GPU."""
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


def single(dm, w, h, eye, clear=M.BACKGROUND):
    c = A.NewContext(w, h)
    c.ClearColorBufferWith(clear)
    c.ClearDepthBuffer()
    c.Shader = M.batch_shader(w, h, eye)
    info = c.DrawMesh(dm)
    return info, c.Image(), c.DepthBuffer


def test_clear_between_runs_is_replaced_by_the_replay():
    """A clear issued on a view's context between runs, with no readback in
    between, must not survive the replay (the view clears its own frame)."""
    _say("batch: clear between runs, no readback")
    meshes, views = small_workload()
    home = A.NewContext(8, 8)
    dms = [prepare(home, p) for p in meshes[:3]]
    b = A.Batch()
    ctxs = []
    for m, eye, w, h in views[:6]:
        if m >= 3:
            continue
        c = A.NewContext(w, h)
        c.Shader = M.batch_shader(w, h, eye)
        b.Add(c, dms[m], clear_color=M.BACKGROUND)
        ctxs.append((c, m, eye, w, h))
    b.Run()
    b.Run()
    for c, *_ in ctxs:
        c.ClearColorBufferWith(A.Color(1, 0, 0, 1))
        c.ClearDepthBufferWith(0.25)
    infos = b.Run()
    for i, (c, m, eye, w, h) in enumerate(ctxs):
        info, img, depth = single(dms[m], w, h, eye)
        assert infos[i] == info
        assert np.array_equal(c.Image(), img) and np.array_equal(c.DepthBuffer, depth), i
    b.close()


def test_view_with_an_empty_mesh_still_clears():
    _say("batch: empty-mesh view")
    meshes, views = small_workload()
    home = A.NewContext(8, 8)
    dm = prepare(home, meshes[0])
    empty = home.Upload(A.Mesh(None, None))
    m, eye, w, h = views[0]
    c = A.NewContext(w, h)
    c.Shader = M.batch_shader(w, h, eye)
    c.DrawMesh(dm)  # something for the clear to replace
    other = A.NewContext(64, 64)
    other.Shader = M.batch_shader(64, 64, eye)
    clear = A.Color(0.25, 0.5, 0.75, 1)
    b = A.Batch()
    b.Add(other, dm, clear_color=M.BACKGROUND)
    b.Add(c, empty, clear_color=clear, clear_depth=0.5)
    want = np.empty((h, w, 4), np.uint8)
    want[:] = clear.NRGBA()
    for rep in range(3):
        infos = b.Run()
        assert infos[1] == (0, 0)
        assert np.array_equal(c.Image(), want), rep
        assert np.all(c.DepthBuffer == 0.5), rep
        c.DrawMesh(dm)  # scribble: the next replay must clear it again
    b.close()


def test_replay_after_a_context_grew_its_buffers():
    """A direct draw of a larger mesh on a view's context reallocates its
    buffers; the next run must not replay the old addresses."""
    _say("batch: direct draw grows a view's buffers between runs")
    meshes, views = small_workload()
    home = A.NewContext(8, 8)
    dms = [prepare(home, p) for p in meshes[:2]]
    big = prepare(home, M.noise_sphere_prims(ntri=200000, detail=7, amplitude=0.15, seed=77))
    b = A.Batch()
    ctxs = []
    for m, eye, w, h in views:
        if m >= 2:
            continue
        c = A.NewContext(w, h)
        c.Shader = M.batch_shader(w, h, eye)
        b.Add(c, dms[m], clear_color=M.BACKGROUND)
        ctxs.append((c, m, eye, w, h))
    b.Run()
    b.Run()
    c0 = ctxs[0][0]
    c0.ClearColorBufferWith(M.BACKGROUND)
    c0.ClearDepthBuffer()
    c0.DrawMesh(big)
    for rep in range(2):
        infos = b.Run()
        for i, (c, m, eye, w, h) in enumerate(ctxs):
            info, img, depth = single(dms[m], w, h, eye)
            assert infos[i] == info, (rep, i)
            assert np.array_equal(c.Image(), img) and np.array_equal(c.DepthBuffer, depth), (rep, i)
    b.close()


def test_overflow_in_a_replay_on_a_shared_context():
    """A view whose frame outgrew the recorded capacities (its mesh was
    scaled up in place) is redrawn with every later view on its context:
    the context must end with its last view's frame."""
    _say("batch: replay overflow on a shared context")
    home = A.NewContext(8, 8)
    # whole detail-4 spheres (5,120 triangles), shrunk to radius 0.4 so that
    # their first frames fit the capacities the first draws reserve
    a = prepare(home, M.noise_sphere_prims(ntri=5120, detail=4, amplitude=0.15, seed=5))
    bm = prepare(home, M.noise_sphere_prims(ntri=5120, detail=4, amplitude=0.15, seed=6))
    a.Transform(A.Scale(A.V(0.4, 0.4, 0.4)))
    bm.Transform(A.Scale(A.V(0.4, 0.4, 0.4)))
    eye0, eye1, eye2 = A.V(3.5, 0.3, 0.2), A.V(-0.4, 0.5, 3.4), A.V(0.2, 3.4, 0.5)
    S = A.NewContext(1024, 1024)
    T = A.NewContext(256, 256)
    b = A.Batch()
    S.Shader = M.batch_shader(1024, 1024, eye0)
    b.Add(S, a, clear_color=M.BACKGROUND)
    S.Shader = M.batch_shader(1024, 1024, eye1)
    b.Add(S, bm, clear_color=M.BACKGROUND)
    T.Shader = M.batch_shader(256, 256, eye2)
    b.Add(T, a, clear_color=M.BACKGROUND)
    b.Run()
    b.Run()
    a.Transform(A.Scale(A.V(2.5, 2.5, 2.5)))  # in place (no reallocation): now fills the 1024x1024 view
    for rep in range(2):
        infos = b.Run()
        i0, _, _ = single(a, 1024, 1024, eye0)
        i1, img1, d1 = single(bm, 1024, 1024, eye1)
        i2, img2, d2 = single(a, 256, 256, eye2)
        # more fragments than the 1024x1024 context's fragment capacity after its
        # first frames: 2 * 3 * (2 * 5120 triangles + tiles + 65536), at most
        # 466,944 with the smallest (32x16) tiles
        assert i0[0] > 466944, i0
        assert infos == [i0, i1, i2], (rep, infos)
        assert np.array_equal(S.Image(), img1) and np.array_equal(S.DepthBuffer, d1), rep
        assert np.array_equal(T.Image(), img2) and np.array_equal(T.DepthBuffer, d2), rep
    b.close()
