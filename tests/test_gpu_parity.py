"""GPU parity: the gfx950 draw path against the oracle on the same inputs.

Bar: depth buffer bit-exact, coverage exact, RasterizeInfo equal, 8-bit
color within 1 LSB (the only color differences allowed come from math.Pow
with a non-integer exponent, whose Go bits are unpinned).  Each case prints
its name before touching the GPU.
"""
import math
import sys

import numpy as np
import pytest

from fauxgl_amd import api as A
from fauxgl_amd import meshes as M
from helpers import compare, run_gpu, run_oracle

pytestmark = pytest.mark.gpu


def _say(name):
    print(f"[case] {name}", flush=True)
    sys.stdout.flush()


def check(name, W, H, ops, color_tol=0, resident=False, oracle_threads=1):
    _say(name)
    ref = run_oracle(W, H, ops, oracle_threads)
    got = run_gpu(W, H, ops, resident=resident)
    res = compare(ref, got, color_tol)
    assert res["ok"], f"{name}: {res}"
    return ref, got


def rand_tris(n, seed, lo=-1.2, hi=1.2, z=(-0.9, 0.9), size=None, colors=False, uvs=False):
    rng = np.random.Generator(np.random.PCG64(seed))
    if size is None:
        pos = rng.uniform(lo, hi, size=(n, 3, 3))
        pos[..., 2] = rng.uniform(z[0], z[1], size=(n, 3))
    else:
        c = rng.uniform(lo, hi, size=(n, 1, 3))
        c[..., 2] = rng.uniform(z[0], z[1], size=(n, 1))
        pos = c + rng.uniform(-size, size, size=(n, 3, 3))
    nrm = M.with_face_normals(pos)
    col = rng.uniform(0, 1, size=(n, 3, 4)) if colors else None
    uv = rng.uniform(-0.5, 1.5, size=(n, 3, 2)) if uvs else None
    return A.Prims(pos, nrm, uv, col)


ORTHO = A.Identity()
BG = A.HexColor("#FFF8E3")


def test_sphere_and_cube_512():
    W = H = 512
    mesh = M.sphere_and_cube_mesh()
    ops = [("clear_color", M.BACKGROUND), ("shader", M.phong_shader(W, H)), ("mesh", mesh)]
    check("sphere_and_cube_512", W, H, ops)


def test_sphere_and_cube_resident():
    W = H = 256
    mesh = M.sphere_and_cube_mesh()
    ops = [("clear_color", M.BACKGROUND), ("shader", M.phong_shader(W, H)), ("mesh", mesh)]
    check("sphere_and_cube_resident", W, H, ops, resident=True)


@pytest.mark.parametrize("W,H", [(1, 1), (7, 5), (33, 31), (64, 64), (100, 37)])
def test_random_solid_sizes(W, H):
    tris = rand_tris(300, 11 + W)
    ops = [("clear_color", BG), ("shader", A.NewSolidColorShader(ORTHO, A.Color(0.2, 0.5, 0.9, 1.0))),
           ("tris", tris)]
    check(f"random_solid_{W}x{H}", W, H, ops)


def test_random_phong_vertex_colors_cull_none():
    W, H = 160, 120
    tris = rand_tris(2000, 5, size=0.3, colors=True)
    sh = A.NewPhongShader(ORTHO, A.V(0.3, 0.5, 1).Normalize(), A.V(0, 0, 3))
    ops = [("clear_color", A.Black), ("set", {"Cull": A.CullNone}), ("shader", sh), ("tris", tris)]
    check("random_phong_vertex_colors_cull_none", W, H, ops)


@pytest.mark.parametrize("cull,face", [(A.CullFront, A.FaceCCW), (A.CullBack, A.FaceCW), (A.CullNone, A.FaceCW)])
def test_cull_modes(cull, face):
    W, H = 96, 80
    tris = rand_tris(800, 7, size=0.4)
    ops = [("clear_color", BG), ("set", {"Cull": cull, "FrontFace": face}),
           ("shader", A.NewSolidColorShader(ORTHO, A.Color(1, 0, 0, 1))), ("tris", tris)]
    check(f"cull_{cull}_{face}", W, H, ops)


def test_depth_state_toggles():
    W, H = 96, 96
    t1 = rand_tris(600, 21, size=0.5)
    t2 = rand_tris(600, 22, size=0.5)
    red = A.NewSolidColorShader(ORTHO, A.Color(1, 0, 0, 1))
    blue = A.NewSolidColorShader(ORTHO, A.Color(0, 0, 1, 1))
    ops = [("clear_color", BG), ("set", {"Cull": A.CullNone}), ("shader", red), ("tris", t1),
           ("set", {"ReadDepth": False}), ("shader", blue), ("tris", t2),
           ("set", {"ReadDepth": True, "WriteDepth": False}), ("shader", red), ("tris", t2),
           ("set", {"WriteDepth": True, "WriteColor": False}), ("tris", t1)]
    check("depth_state_toggles", W, H, ops)


@pytest.mark.parametrize("bias", [1e-4, -1e-5, 0.25, -0.3])
def test_depth_bias(bias):
    W, H = 120, 90
    t = rand_tris(1500, 31, size=0.35, colors=True)
    sh = A.NewPhongShader(ORTHO, A.V(0, 0, 1), A.V(0, 0, 3))
    ops = [("clear_color", BG), ("set", {"Cull": A.CullNone, "DepthBias": bias}), ("shader", sh), ("tris", t),
           ("tris", t)]
    check(f"depth_bias_{bias}", W, H, ops)


def test_alpha_blend_order():
    W, H = 128, 96
    t = rand_tris(3000, 41, size=0.3, colors=True)
    sh = A.NewSolidColorShader(ORTHO, A.Color(0.9, 0.3, 0.1, 0.35))
    ops = [("clear_color", A.Color(0.1, 0.2, 0.3, 0.5)), ("set", {"Cull": A.CullNone, "ReadDepth": False}),
           ("shader", sh), ("tris", t)]
    check("alpha_blend_order", W, H, ops)


def test_alpha_blend_vertex_alpha_phong():
    W, H = 128, 96
    t = rand_tris(3000, 43, size=0.3, colors=True)
    sh = A.NewPhongShader(ORTHO, A.V(0.2, 0.4, 1).Normalize(), A.V(0, 0, 3))
    ops = [("clear_color", BG), ("set", {"Cull": A.CullNone, "DepthBias": 1e-4}), ("shader", sh), ("tris", t)]
    check("alpha_blend_vertex_alpha_phong", W, H, ops)


def test_discard_color():
    W, H = 64, 64
    t = rand_tris(400, 51, size=0.5)
    ops = [("clear_color", BG), ("shader", A.NewSolidColorShader(ORTHO, A.Discard)), ("tris", t)]
    ref, got = check("discard_color", W, H, ops)
    assert ref[2][0][1] == 0 and ref[2][0][0] > 0


@pytest.mark.parametrize("kind", ["NRGBA", "RGBA", "RGBA16"])
def test_texture_kinds(kind):
    W, H = 128, 128
    rng = np.random.Generator(np.random.PCG64(61))
    if kind == "RGBA16":
        texels = rng.integers(0, 65536, size=(37, 53, 4), dtype=np.uint16)
    else:
        texels = rng.integers(0, 256, size=(37, 53, 4), dtype=np.uint8)
        if kind == "RGBA":  # premultiplied: channels <= alpha
            texels[..., :3] = np.minimum(texels[..., :3], texels[..., 3:4])
    tex = A.ImageTexture(texels, kind)
    t = rand_tris(1200, 62, size=0.45, uvs=True)
    ops = [("clear_color", BG), ("set", {"Cull": A.CullNone}), ("shader", A.NewTextureShader(ORTHO, tex)),
           ("tris", t)]
    check(f"texture_{kind}", W, H, ops)


def test_phong_textured_and_powers():
    W, H = 120, 100
    rng = np.random.Generator(np.random.PCG64(71))
    tex = A.ImageTexture(rng.integers(0, 256, size=(16, 16, 4), dtype=np.uint8), "NRGBA")
    t = rand_tris(1500, 72, size=0.4, uvs=True)
    sh = A.NewPhongShader(ORTHO, A.V(-0.3, 0.6, 1).Normalize(), A.V(0.2, 0.1, 2.5))
    sh.Texture = tex
    ops = [("clear_color", BG), ("set", {"Cull": A.CullNone}), ("shader", sh), ("tris", t)]
    check("phong_textured", W, H, ops)
    for power in (0.0, 1.0, 7.0, 128.0):
        sh2 = A.NewPhongShader(ORTHO, A.V(0.1, 0.2, 1).Normalize(), A.V(0, 0, 2))
        sh2.ObjectColor = A.Color(0.3, 0.6, 0.2, 1.0)
        sh2.SpecularPower = power
        check(f"phong_power_{power}", W, H, [("clear_color", BG), ("set", {"Cull": A.CullNone}), ("shader", sh2),
                                             ("tris", t)])
    # non-integer exponents (the values the oracle's pow is pinned on against
    # exact powers): the device's exp / log may differ from the C library's
    # in the last bit, 1 LSB allowed; depth and coverage stay bit-exact
    for power in (12.5, 0.75, 2.5, 33.3, 7.6):
        sh3 = A.NewPhongShader(ORTHO, A.V(0.1, 0.2, 1).Normalize(), A.V(0, 0, 2))
        sh3.ObjectColor = A.Color(0.3, 0.6, 0.2, 1.0)
        sh3.SpecularPower = power
        check(f"phong_power_{power}", W, H, [("clear_color", BG), ("set", {"Cull": A.CullNone}), ("shader", sh3),
                                             ("tris", t)], color_tol=1)


def _rand_lines(n, seed, spread=1.4):
    rng = np.random.Generator(np.random.PCG64(seed))
    pos = rng.uniform(-spread, spread, size=(n, 2, 3))
    pos[..., 2] = rng.uniform(-0.9, 0.9, size=(n, 2))
    col = rng.uniform(0, 1, size=(n, 2, 4))
    return A.Prims(pos, None, None, col, nv=2)


@pytest.mark.parametrize("lw", [1.0, 2.0, 5.5])
def test_lines_with_offscreen_wrap(lw):
    W, H = 90, 70
    lines = _rand_lines(500, 81 + int(lw))
    sh = A.NewPhongShader(ORTHO, A.V(0, 0, 1), A.V(0, 0, 3))
    ops = [("clear_color", BG), ("set", {"LineWidth": lw}), ("shader", sh), ("lines", lines)]
    check(f"lines_lw{lw}", W, H, ops)
    # lines exactly along the screen borders: their quads reach x < 0 and
    # x >= W, where the reference's index check wraps to adjacent rows
    border = np.array([[[-1, -1, 0.1], [1, -1, 0.1]], [[-1, 1, 0.2], [1, 1, 0.2]], [[-1, -1, 0.3], [-1, 1, 0.3]],
                       [[1, -1, 0.4], [1, 1, 0.4]], [[-1, 0.013, 0.5], [1, 0.013, 0.5]]], dtype=np.float64)
    bl = A.Prims(border, None, None, None, nv=2)
    for blend in (False, True):
        sh2 = A.NewSolidColorShader(ORTHO, A.Color(0.2, 0.9, 0.4, 0.6 if blend else 1.0))
        check(f"border_lines_lw{lw}_blend{blend}", W, H,
              [("clear_color", BG), ("set", {"LineWidth": lw, "ReadDepth": False}), ("shader", sh2), ("lines", bl)])


def test_wireframe_sphere_over_solid():
    W, H = 200, 160
    mesh = A.Mesh(M.new_sphere(3))
    mat = M.camera(W, H, eye=A.V(2.2, 1.5, 2.8), center=A.V(0, 0, 0), up=A.V(0, 1, 0))
    sh = A.NewPhongShader(mat, A.V(0.75, 0.5, 1).Normalize(), A.V(2.2, 1.5, 2.8))
    sh.ObjectColor = A.HexColor("FFFF9D").Alpha(0.65)
    sh.SpecularPower = 0
    ops = [("clear_color", A.Black), ("shader", sh), ("mesh", mesh),
           ("set", {"Wireframe": True, "DepthBias": -0.00001}), ("mesh", mesh)]
    check("wireframe_sphere_over_solid", W, H, ops)


def test_clipping_near_far_and_behind_camera():
    W, H = 160, 120
    rng = np.random.Generator(np.random.PCG64(91))
    pos = rng.uniform(-4, 4, size=(2500, 3, 3))
    mat = A.LookAt(A.V(0, 0, 3), A.V(0, 0, 0), A.V(0, 1, 0)).Perspective(60, W / H, 1, 6)
    p = A.Prims(pos, M.with_face_normals(pos), None, rng.uniform(0, 1, size=(2500, 3, 4)))
    sh = A.NewPhongShader(mat, A.V(0, 0, 1), A.V(0, 0, 3))
    ops = [("clear_color", BG), ("set", {"Cull": A.CullNone}), ("shader", sh), ("tris", p)]
    check("clipping_near_far", W, H, ops)
    ops = [("clear_color", BG), ("set", {"Cull": A.CullNone, "Wireframe": True, "LineWidth": 3}), ("shader", sh),
           ("tris", p)]
    check("clipping_wireframe", W, H, ops)
    lines = A.Prims(pos[:, :2].copy(), None, None, None, nv=2)
    check("clipping_lines", W, H, [("clear_color", BG), ("shader", sh), ("lines", lines)])


def test_clipping_vertices_on_planes_and_corners():
    """Clip-space vertices (identity matrix, w = 1) on a lattice that puts
    many of them exactly on the planes x, y, z = +-1 and makes triangles
    cross two or three planes at once (polygons of up to 9 points, several
    fans each), plus triangles with a vertex behind the eye (w <= 0 through
    a perspective matrix): the clipping kernel's register form and its
    fallback must give the oracle's polygons, fans and records."""
    W, H = 96, 80
    rng = np.random.Generator(np.random.PCG64(97))
    lattice = np.array([-3.0, -2.0, -1.0, -0.5, 0.0, 0.5, 1.0, 2.0, 3.0])
    pos = lattice[rng.integers(0, len(lattice), size=(3000, 3, 3))]
    p = A.Prims(pos, M.with_face_normals(pos), None, rng.uniform(0, 1, size=(3000, 3, 4)))
    sh = A.NewSolidColorShader(ORTHO, A.Color(0.2, 0.6, 0.3, 1))
    check("clipping_lattice_on_planes", W, H, [("clear_color", BG), ("set", {"Cull": A.CullNone}), ("shader", sh),
                                              ("tris", p)])
    # big triangles around a perspective eye: vertices behind it and on the
    # near plane's side of it
    mat = A.LookAt(A.V(0, 0, 0.5), A.V(0, 0, -1), A.V(0, 1, 0)).Perspective(75, W / H, 0.25, 4)
    big = rng.uniform(-6, 6, size=(1500, 3, 3))
    q = A.Prims(big, M.with_face_normals(big), None, rng.uniform(0, 1, size=(1500, 3, 4)))
    ph = A.NewPhongShader(mat, A.V(0.3, 0.5, 1), A.V(0, 0, 0.5))
    check("clipping_behind_eye", W, H, [("clear_color", BG), ("set", {"Cull": A.CullNone}), ("shader", ph),
                                       ("tris", q)])


def test_mesh_triangles_then_lines():
    W, H = 128, 128
    t = rand_tris(500, 101, size=0.4, colors=True)
    lines = _rand_lines(200, 102, spread=1.0)
    mesh = A.Mesh(t, lines)
    sh = A.NewPhongShader(ORTHO, A.V(0, 0, 1), A.V(0, 0, 3))
    ops = [("clear_color", BG), ("set", {"Cull": A.CullNone}), ("shader", sh), ("mesh", mesh)]
    check("mesh_tris_then_lines", W, H, ops)
    check("mesh_tris_then_lines_resident", W, H, ops, resident=True)


def test_clears_and_accumulation():
    W, H = 70, 50
    t1 = rand_tris(300, 111, size=0.5)
    t2 = rand_tris(300, 112, size=0.5)
    red = A.NewSolidColorShader(ORTHO, A.Color(1, 0, 0, 1))
    ops = [("shader", red), ("tris", t1),                      # no clear: initial zero color, MaxFloat64 depth
           ("clear_depth", 0.25), ("tris", t2),                  # custom depth clear
           ("clear_color", A.Color(0.5, 0.5, 0.5, 1)), ("tris", t1), ("tris", t2),
           ("clear_depth", A.MaxFloat64), ("clear_color", A.Color(2.0, -1.0, math.nan, 0.5)), ("tris", t2)]
    check("clears_and_accumulation", W, H, ops)


def test_depth_image():
    W, H = 96, 64
    t = rand_tris(700, 121, size=0.4)
    ops = [("clear_color", BG), ("set", {"Cull": A.CullNone}),
           ("shader", A.NewSolidColorShader(ORTHO, A.Color(1, 1, 1, 1))), ("tris", t)]
    _say("depth_image")
    ref = run_oracle(W, H, ops)
    got = run_gpu(W, H, ops)
    assert np.array_equal(ref[3].depth_image(), got[3].DepthImage())


def test_degenerate_triangles():
    W, H = 64, 48
    rng = np.random.Generator(np.random.PCG64(131))
    pos = rng.uniform(-1, 1, size=(200, 3, 3))
    pos[:50, 1] = pos[:50, 0]                     # repeated vertex
    pos[50:100, 2] = (pos[50:100, 0] + pos[50:100, 1]) / 2  # collinear
    pos[100:120, :, 0] = 0.5                      # vertical sliver
    pos[120:130] = 0                              # all at the origin
    p = A.Prims(pos, M.with_face_normals(pos))
    sh = A.NewSolidColorShader(ORTHO, A.Color(0.3, 0.3, 0.9, 1))
    check("degenerate_triangles", W, H, [("clear_color", BG), ("set", {"Cull": A.CullNone}), ("shader", sh),
                                        ("tris", p)])


def test_pixel_center_ties():
    # vertices on pixel centers and edges exactly: equal-depth ties must go
    # to the later triangle (<=, context.go:248)
    W, H = 32, 32
    g = np.linspace(-1, 1, 9)
    tris = []
    for i in range(8):
        for j in range(8):
            a, b, c, d = (g[i], g[j]), (g[i + 1], g[j]), (g[i + 1], g[j + 1]), (g[i], g[j + 1])
            tris.append([[*a, 0.0], [*b, 0.0], [*c, 0.0]])
            tris.append([[*a, 0.0], [*c, 0.0], [*d, 0.0]])
    pos = np.array(tris)
    rng = np.random.Generator(np.random.PCG64(141))
    col = np.repeat(rng.uniform(0, 1, size=(len(tris), 1, 4)), 3, axis=1)
    col[..., 3] = 1
    p = A.Prims(pos, np.zeros_like(pos), None, col)
    sh = A.NewPhongShader(ORTHO, A.V(0, 0, 1), A.V(0, 0, 3))
    check("pixel_center_ties", W, H, [("clear_color", BG), ("set", {"Cull": A.CullNone}), ("shader", sh),
                                     ("tris", p), ("tris", p)])


def test_noise_sphere_reduced():
    W, H = 480, 270
    mesh = M.noise_sphere_mesh(ntri=60000, detail=6)
    ops = [("clear_color", M.BACKGROUND), ("shader", M.phong_shader(W, H)), ("mesh", mesh)]
    check("noise_sphere_60k_480x270", W, H, ops)


def test_soup_blend_texture_reduced():
    W, H = 480, 270
    mesh, sh = M.soup_scene(W, H, n=20000, tex_size=128)
    ops = [("clear_color", M.BACKGROUND), ("set", {"DepthBias": 1e-4}), ("shader", sh), ("mesh", mesh)]
    check("soup_blend_texture_20k_480x270", W, H, ops)


def test_noise_sphere_full_1080p():
    W, H = 1920, 1080
    mesh = M.noise_sphere_mesh()
    ops = [("clear_color", M.BACKGROUND), ("shader", M.phong_shader(W, H)), ("mesh", mesh)]
    check("noise_sphere_871k_1920x1080", W, H, ops, resident=True)


def test_soup_blend_texture_full_1080p():
    W, H = 1920, 1080
    mesh, sh = M.soup_scene(W, H)
    ops = [("clear_color", M.BACKGROUND), ("set", {"DepthBias": 1e-4}), ("shader", sh), ("mesh", mesh)]
    check("soup_blend_texture_200k_1920x1080", W, H, ops, resident=True)


def test_noise_sphere_full_16x():
    """The headline draw itself: 871,306 triangles at 7680x4320 (4x4 SSAA
    internal size), bit-exact against the single-threaded oracle."""
    W, H = 7680, 4320
    mesh = M.noise_sphere_mesh()
    ops = [("clear_color", M.BACKGROUND), ("shader", M.phong_shader(1920, 1080)), ("mesh", mesh)]
    check("noise_sphere_871k_7680x4320", W, H, ops, resident=True)
