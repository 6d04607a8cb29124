"""The C ABI as a Go shim drives it, and the draw path's failure handling.

- Go's enum values (context.go:11-26: FaceCW = 1, FaceCCW = 2; CullNone = 1,
  CullFront = 2, CullBack = 3) passed unchanged through fgl_draw_triangles,
  plus values Go would accept but never names (0, 7, -1): bit-exact against
  the oracle given the same ints (context.go:328-336 compares, it does not
  validate).
- Capacity overflow forced from the host (fgl_ctx_force_capacities): each of
  the five capacity bits is hit, the kernels exit normally, the host grows
  the buffers and reruns, and the result is bit-exact against the oracle
  with no error recorded.
- The host mirror's residency policy: a mesh drawn, changed in place by
  Mesh.Transform, and drawn again shows the change (examples/animate.go:59-66).

Each case prints its name before touching the GPU."""
import ctypes
import sys

import numpy as np
import pytest

from fauxgl_amd import _native as N
from fauxgl_amd import api as A
from fauxgl_amd import meshes as M
from helpers import compare, run_oracle

pytestmark = pytest.mark.gpu

ORTHO = A.Identity()
BG = A.HexColor("#FFF8E3")


def _say(name):
    print(f"[case] {name}", flush=True)
    sys.stdout.flush()


def rand_tris(n, seed, size=0.35, colors=True):
    rng = np.random.Generator(np.random.PCG64(seed))
    c = rng.uniform(-1.1, 1.1, size=(n, 1, 3))
    c[..., 2] = rng.uniform(-0.9, 0.9, size=(n, 1))
    pos = c + rng.uniform(-size, size, size=(n, 3, 3))
    col = rng.uniform(0, 1, size=(n, 3, 4)) if colors else None
    return A.Prims(pos, M.with_face_normals(pos), None, col)


def gpu_frame(ctx):
    return ctx.Image(), ctx.DepthBuffer


# ---------------------------------------------------------------------------
# enums as the Go shim passes them


def draw_raw(ctx, shader, prims, front_face, cull):
    """fgl_draw_triangles with a hand-built fgl_state, as the cgo shim fills
    it from dc.FrontFace and dc.Cull (no translation)."""
    st = N.FglState(1, 1, 1, 1, 0, (ctypes.c_uint8 * 3)(), front_face, cull, 2.0, 0.0)
    sh = A.shader_struct(shader, ctx)
    p = prims.struct()
    info = (ctypes.c_uint64 * 2)()
    rc = N.load().fgl_draw_triangles(ctx._ptr(), ctypes.byref(st), ctypes.byref(sh), ctypes.byref(p), info)
    assert rc == 0, N.load().fgl_last_error(ctx._ptr())
    return int(info[0]), int(info[1])


GO_FACE_CW, GO_FACE_CCW = 1, 2                        # context.go:11-16
GO_CULL_NONE, GO_CULL_FRONT, GO_CULL_BACK = 1, 2, 3   # context.go:18-26


@pytest.mark.parametrize("front_face,cull", [
    (GO_FACE_CCW, GO_CULL_BACK), (GO_FACE_CCW, GO_CULL_NONE), (GO_FACE_CCW, GO_CULL_FRONT),
    (GO_FACE_CW, GO_CULL_BACK), (GO_FACE_CW, GO_CULL_NONE), (GO_FACE_CW, GO_CULL_FRONT),
    (0, 0), (7, GO_CULL_FRONT), (GO_FACE_CW, -1), (-3, 9)])
def test_go_enum_values_pass_through_unchanged(front_face, cull):
    _say(f"go enums front_face={front_face} cull={cull}")
    W, H = 96, 72
    tris = rand_tris(600, 17)
    sh = A.NewPhongShader(ORTHO, A.V(0.2, 0.4, 1).Normalize(), A.V(0, 0, 3))
    ctx = A.NewContext(W, H)
    ctx.ClearColorBufferWith(BG)
    info = draw_raw(ctx, sh, tris, front_face, cull)
    got = (*gpu_frame(ctx), [info])
    ref = run_oracle(W, H, [("clear_color", BG), ("set", {"FrontFace": front_face, "Cull": cull}), ("shader", sh),
                            ("tris", tris)])
    res = compare(ref, got, color_tol=0)
    assert res["ok"], res
    assert info[0] > 0


def test_enum_semantics_follow_the_reference_comparisons():
    # CullNone keeps both faces (more pixels than either cull); FaceCW with
    # CullBack culls what FaceCCW with CullFront culls
    _say("go enums: face/cull relations")
    W, H = 96, 72
    tris = rand_tris(600, 18)
    sh = A.NewSolidColorShader(ORTHO, A.Color(0.2, 0.5, 0.9, 1.0))
    totals = {}
    for ff, cu in ((2, 1), (2, 2), (2, 3), (1, 3), (1, 2), (2, 0)):
        ctx = A.NewContext(W, H)
        totals[(ff, cu)] = draw_raw(ctx, sh, tris, ff, cu)[0]
    assert totals[(2, 1)] == totals[(2, 2)] + totals[(2, 3)]
    assert totals[(1, 3)] == totals[(2, 2)] and totals[(1, 2)] == totals[(2, 3)]
    assert totals[(2, 0)] == totals[(2, 3)]  # an unnamed value acts as CullBack


# ---------------------------------------------------------------------------
# capacity overflow forced from the host


def clip_scene():
    rng = np.random.Generator(np.random.PCG64(91))
    W, H = 160, 120
    pos = rng.uniform(-4, 4, size=(2500, 3, 3))
    mat = A.LookAt(A.V(0, 0, 3), A.V(0, 0, 0), A.V(0, 1, 0)).Perspective(60, W / H, 1, 6)
    p = A.Prims(pos, M.with_face_normals(pos), None, rng.uniform(0, 1, size=(2500, 3, 4)))
    sh = A.NewPhongShader(mat, A.V(0, 0, 1), A.V(0, 0, 3))
    return W, H, p, sh


CAPS = {1: "side_triangles", 2: "clipped_vertices", 4: "pairs", 16: "spans", 32: "fragments"}


@pytest.mark.parametrize("bit", sorted(CAPS))
@pytest.mark.parametrize("prim", ["tris", "wireframe", "lines"])
def test_forced_capacity_overflow_reruns_exactly(bit, prim):
    if bit == 2 and prim == "lines":
        pytest.skip("lines are clipped without new vertices (clipping.go:76-98)")
    _say(f"forced overflow bit {bit} ({CAPS[bit]}) on {prim}")
    W, H, p, sh = clip_scene()
    ops = [("clear_color", BG), ("set", {"Cull": A.CullNone}), ("shader", sh)]
    if prim == "wireframe":
        ops.append(("set", {"Wireframe": True, "LineWidth": 3}))
    ops.append(("lines", A.Prims(p.position[:, :2].copy(), None, None, None, nv=2)) if prim == "lines"
               else ("tris", p))
    ref = run_oracle(W, H, ops)
    ctx = A.NewContext(W, H)
    ctx.ClearColorBufferWith(BG)
    ctx.Cull = A.CullNone
    ctx.Shader = sh
    if prim == "wireframe":
        ctx.Wireframe, ctx.LineWidth = True, 3
    # a first draw sizes the buffers; the forced one must overflow anyway
    (ctx.DrawLines if prim == "lines" else ctx.DrawTriangles)(ops[-1][1])
    ctx.ClearColorBufferWith(BG)
    ctx.ClearDepthBuffer()
    ctx.ForceCapacities(**{CAPS[bit]: 1})
    info = (ctx.DrawLines if prim == "lines" else ctx.DrawTriangles)(ops[-1][1])
    st = ctx.DrawStats()
    assert st["overflow_mask"] & bit, st
    assert st["reruns"] >= 1, st
    assert N.load().fgl_last_error(ctx._ptr()) == b""
    res = compare(ref, (*gpu_frame(ctx), [tuple(info)]), color_tol=0)
    assert res["ok"], res
    # the next draw runs with its grown capacities and no rerun
    ctx.ClearColorBufferWith(BG)
    ctx.ClearDepthBuffer()
    info2 = (ctx.DrawLines if prim == "lines" else ctx.DrawTriangles)(ops[-1][1])
    assert tuple(info2) == tuple(info) and ctx.DrawStats()["reruns"] == 0


def test_forced_overflow_of_every_capacity_at_once():
    _say("forced overflow of all five capacities")
    W, H, p, sh = clip_scene()
    ops = [("clear_color", BG), ("set", {"Cull": A.CullNone}), ("shader", sh), ("tris", p)]
    ref = run_oracle(W, H, ops)
    ctx = A.NewContext(W, H)
    ctx.ClearColorBufferWith(BG)
    ctx.Cull = A.CullNone
    ctx.Shader = sh
    ctx.ForceCapacities(1, 1, 1, 1, 1)
    info = ctx.DrawTriangles(p)
    st = ctx.DrawStats()
    assert st["overflow_mask"] & 3 and st["reruns"] >= 1, st
    res = compare(ref, (*gpu_frame(ctx), [tuple(info)]), color_tol=0)
    assert res["ok"], res


def test_forced_overflow_on_the_headline_scene_at_16x_size():
    # the fragment and span capacities of the 871,306-triangle frame at
    # 7680x4320, forced well under the need: same frame as without
    _say("forced overflow: noise sphere 871k at 7680x4320")
    mesh = M.noise_sphere_mesh()
    ctx = A.NewContext(7680, 4320)
    ctx.Shader = M.phong_shader(1920, 1080)
    dm = ctx.Upload(mesh)
    ctx.ClearColorBufferWith(M.BACKGROUND)
    ctx.ClearDepthBuffer()
    want_info = ctx.DrawMesh(dm)
    want = gpu_frame(ctx)
    for caps in ({"fragments": 1 << 20}, {"spans": 1 << 18}, {"pairs": 4096}):
        ctx.ClearColorBufferWith(M.BACKGROUND)
        ctx.ClearDepthBuffer()
        ctx.ForceCapacities(**caps)
        info = ctx.DrawMesh(dm)
        st = ctx.DrawStats()
        assert st["reruns"] >= 1, (caps, st)
        got = gpu_frame(ctx)
        assert info == want_info and np.array_equal(got[0], want[0]), caps
        assert np.array_equal(got[1].view(np.uint64), want[1].view(np.uint64)), caps


# ---------------------------------------------------------------------------
# residency of host meshes (the Go shim's cache policy)


@pytest.mark.parametrize("resident", [False, True])
def test_draw_transform_draw(resident):
    _say(f"draw, Transform, draw (resident={resident})")
    W, H = 128, 96
    mesh = A.Mesh(rand_tris(400, 23, colors=False), None)
    sh = A.NewPhongShader(ORTHO, A.V(0.2, 0.4, 1).Normalize(), A.V(0, 0, 3))
    sh.ObjectColor = A.HexColor("#468966")  # opaque: the vertices carry no color
    ctx = A.NewContext(W, H)
    ctx.Shader = sh
    if resident:
        ctx.KeepResident(mesh)
    frames = []
    m = A.Translate(A.V(0.3, -0.2, 0)).Mul(A.Scale(A.V(0.6, 0.8, 0.5)))
    snapshots = []
    for k in range(3):
        snapshots.append(A.Mesh(A.Prims(mesh.Triangles.position.copy(), mesh.Triangles.normal.copy()), None))
        ctx.ClearColorBufferWith(BG)
        ctx.ClearDepthBuffer()
        info = ctx.DrawMesh(mesh)
        frames.append((*gpu_frame(ctx), [tuple(info)]))
        mesh.Transform(m)  # in place, as examples/animate.go:66 does
    for k in range(3):
        ref = run_oracle(W, H, [("clear_color", BG), ("shader", sh), ("mesh", snapshots[k])])
        res = compare(ref, frames[k], color_tol=0)
        assert res["ok"], (k, res)
    assert not np.array_equal(frames[0][0], frames[1][0])
