"""The host-side setup that feeds both the GPU path and the oracle: the
Python mirror's matrices (matrix.go:69-128, 188-222), HexColor
(color.go:31-54), the NewPhongShader defaults (shader.go:61-68) and the
host Mesh.Transform (mesh.go:167-175).  A wrong matrix here would reach the
GPU and the oracle alike and pass every parity test, so each is checked
against the oracle's pinned restatement bit for bit, or against a closed
form where the oracle has none.  CPU only."""
import math
import struct

import numpy as np

import oracle as O
from fauxgl_amd import api as A


def bits(v):
    return [struct.unpack("<Q", struct.pack("<d", float(x)))[0] for x in np.ravel(v)]


def same_bits(a, b):
    return bits(a) == bits(b)


def cameras():
    rng = np.random.Generator(np.random.PCG64(21))
    out = [((-3.0, 1.0, -0.75), (0.0, -0.07, 0.0), (0.0, 1.0, 0.0))]
    for _ in range(40):
        eye = tuple(float(v) for v in rng.uniform(-6, 6, 3))
        center = tuple(float(v) for v in rng.uniform(-1, 1, 3))
        up = tuple(float(v) for v in rng.normal(size=3))
        out.append((eye, center, up))
    return out


def test_look_at_bits_equal_the_oracle():
    for eye, center, up in cameras():
        got = A.LookAt(A.V(*eye), A.V(*center), A.V(*up))
        ref = O.look_at(np.array(eye), np.array(center), np.array(up))
        assert same_bits(got, ref), (eye, center, up)


def test_frustum_perspective_screen_bits_equal_the_oracle():
    rng = np.random.Generator(np.random.PCG64(22))
    for _ in range(40):
        l, b, n = (float(v) for v in rng.uniform(-2, -0.1, 3))
        r, t = (float(v) for v in rng.uniform(0.1, 2, 2))
        n = -n
        f = n + float(rng.uniform(0.5, 50))
        assert same_bits(A.Frustum(l, r, b, t, n, f), O.frustum(l, r, b, t, n, f))
        fovy, aspect = float(rng.uniform(10, 120)), float(rng.uniform(0.3, 3))
        assert same_bits(A.Perspective(fovy, aspect, n, f), O.perspective(fovy, aspect, n, f))
    for w, h in ((512, 512), (1920, 1080), (7680, 4320), (1, 1), (37, 5)):
        assert same_bits(A.Screen(w, h), O.screen(w, h))


def test_configured_camera_matrix_bits_equal_the_oracle():
    # LookAt(...).Perspective(...) (matrix.go:167-169) = Perspective.Mul(LookAt)
    for (eye, center, up), aspect in zip(cameras(), (1.0, 16 / 9, 0.5, 2.0) * 11):
        got = A.LookAt(A.V(*eye), A.V(*center), A.V(*up)).Perspective(30, aspect, 1, 10)
        ref = O.matrix_mul(O.perspective(30, aspect, 1, 10), O.look_at(np.array(eye), np.array(center), np.array(up)))
        assert same_bits(got, ref)


def test_matrix_mul_bits_equal_the_oracle():
    rng = np.random.Generator(np.random.PCG64(23))
    for _ in range(50):
        a, b = rng.normal(size=16), rng.normal(size=16)
        assert same_bits(A.Matrix(a).Mul(A.Matrix(b)), O.matrix_mul(a, b))
        p = rng.normal(size=3)
        got = A.Matrix(a).MulPosition(A.V(*p))
        ref = O.mul_position_w(a, p)[:3]
        assert same_bits(got, ref)


def test_orthographic_is_gl_ortho():
    # glOrtho; the operands are dyadic so every entry is exact
    l, r, b, t, n, f = -2.0, 2.0, -1.0, 1.0, 1.0, 9.0
    m = np.array(A.Orthographic(l, r, b, t, n, f)).reshape(4, 4)
    gl = np.array([[2 / (r - l), 0, 0, -(r + l) / (r - l)], [0, 2 / (t - b), 0, -(t + b) / (t - b)],
                   [0, 0, -2 / (f - n), -(f + n) / (f - n)], [0, 0, 0, 1]])
    assert np.array_equal(m, gl)
    # the box's corners map to the NDC cube's
    for (x, y, z), want in (((l, b, -n), (-1, -1, -1)), ((r, t, -f), (1, 1, 1))):
        v = m @ np.array([x, y, z, 1.0])
        assert np.array_equal(v[:3], np.array(want, dtype=float))


def test_hex_color_against_hand_computed_values():
    d = 255.0
    cases = {
        "#468966": (0x46 / d, 0x89 / d, 0x66 / d, 1.0),
        "#FFF8E3": (1.0, 0xF8 / d, 0xE3 / d, 1.0),
        "abc": (0xAA / d, 0xBB / d, 0xCC / d, 1.0),
        "#1234": (0x11 / d, 0x22 / d, 0x33 / d, 0x44 / d),
        "#11223380": (0x11 / d, 0x22 / d, 0x33 / d, 0x80 / d),
        "##00ff00##": (0.0, 1.0, 0.0, 1.0),  # strings.Trim strips both ends
        "#12345": (0.0, 0.0, 0.0, 1.0),       # no case for length 5
        "": (0.0, 0.0, 0.0, 1.0),
        "#12zz00": (0x12 / d, 0.0, 0.0, 1.0),  # Sscanf stops at the first bad field
        "#1g3": (0x11 / d, 0.0, 0.0, 1.0),
    }
    for s, want in cases.items():
        assert tuple(A.HexColor(s)) == want, s
    # the NRGBA of the clear color config 1 uses (color.go:56-63, truncating)
    assert A.HexColor("#FFF8E3").NRGBA() == (255, 248, 227, 255)
    nrgba = np.zeros(4, dtype=np.uint8)
    for s in ("#468966", "#FFF8E3", "#11223380"):
        assert tuple(O.color_nrgba(tuple(A.HexColor(s)))) == A.HexColor(s).NRGBA()


def test_new_phong_shader_defaults_are_shader_go():
    m = A.Identity()
    sh = A.NewPhongShader(m, A.V(0, 0, 1), A.V(1, 2, 3))
    assert tuple(sh.AmbientColor) == (0.2, 0.2, 0.2, 1.0)
    assert tuple(sh.DiffuseColor) == (0.8, 0.8, 0.8, 1.0)
    assert tuple(sh.SpecularColor) == (1.0, 1.0, 1.0, 1.0)
    assert tuple(sh.ObjectColor) == (0.0, 0.0, 0.0, 0.0) and sh.Texture is None
    assert sh.SpecularPower == 32.0
    # the fgl_shader it becomes carries the same values as the oracle's
    # default Phong shader
    s = A.shader_struct(sh)
    o = O.make_shader(3, list(m), light_dir=(0, 0, 1), camera_pos=(1, 2, 3))
    for f in ("ambient", "diffuse", "specular", "object_color", "light_dir", "camera_pos", "matrix"):
        assert list(getattr(s, f)) == list(getattr(o, f)), f
    assert s.specular_power == o.specular_power and s.texture == -1


def test_context_defaults_are_new_context():
    # NewContext (context.go:60-81) field defaults, as the state struct the
    # draw receives (no device needed: the state is built on the host)
    ctx = object.__new__(A.Context)
    ctx.ReadDepth = ctx.WriteDepth = ctx.WriteColor = ctx.AlphaBlend = True
    ctx.Wireframe = False
    ctx.FrontFace, ctx.Cull, ctx.LineWidth, ctx.DepthBias = A.FaceCCW, A.CullBack, 2.0, 0.0
    st = A.Context._state(ctx)
    assert (st.front_face, st.cull) == (2, 3)  # FaceCCW, CullBack in Go's iota order
    assert (A.FaceCW, A.FaceCCW, A.CullNone, A.CullFront, A.CullBack) == (1, 2, 1, 2, 3)


def test_host_mesh_transform_bits_equal_the_oracle():
    rng = np.random.Generator(np.random.PCG64(24))
    pos = rng.normal(size=(50, 3, 3))
    nrm = rng.normal(size=(50, 3, 3))
    nrm[3, 1] = 0.0  # a zero normal normalises to NaN, as in Go
    lpos = rng.normal(size=(7, 2, 3))
    lnrm = rng.normal(size=(7, 2, 3))
    mesh = A.Mesh(A.Prims(pos.copy(), nrm.copy()), A.Prims(lpos.copy(), lnrm.copy(), nv=2))
    m = A.Translate(A.V(0.5, -1, 2)).Mul(A.Scale(A.V(2, 3, 0.5))).Mul(
        A.LookAt(A.V(1, 2, 3), A.V(0, 0, 0), A.V(0, 1, 0)))
    g0 = mesh.generation
    mesh.Transform(m)
    assert mesh.generation == g0 + 1  # Transform calls dirty (mesh.go:174)
    P, Nn = O.mesh_transform(pos.reshape(-1, 3), nrm.reshape(-1, 3), np.array(m))
    assert same_bits(mesh.Triangles.position, P) and same_bits(mesh.Triangles.normal, Nn)
    P, Nn = O.mesh_transform(lpos.reshape(-1, 3), lnrm.reshape(-1, 3), np.array(m))
    assert same_bits(mesh.Lines.position, P) and same_bits(mesh.Lines.normal, Nn)
    assert math.isnan(mesh.Triangles.normal[3, 1, 0])
