"""Pins the oracle (the C restatement of fauxgl's draw path) to facts that
do not depend on it: closed forms (gluLookAt / glFrustum / viewport),
exact rational arithmetic (coverage, clipping, barycentrics), hand-computed
integer cases (NRGBA, premultiply, blend), and invariants.  Each check is
chosen so that a plausible slip in the restatement (a dropped term, a wrong
sign or index, a transposed operand) fails it; test_oracle_mutants.py
plants such slips and shows these checks catch them.
"""
import math
import struct
from fractions import Fraction as F

import numpy as np
import pytest

import oracle as O

MAXF = 1.7976931348623157e308


def bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def ulps(a, b):
    ia, ib = bits(a), bits(b)
    if ia >> 63:
        ia = (1 << 63) - ia
    if ib >> 63:
        ib = (1 << 63) - ib
    return abs(ia - ib)


# ---------------------------------------------------------------------------
# Go semantics


def test_go_min_max_special_cases():
    nan, inf = float("nan"), float("inf")
    assert math.isnan(O.go_min(nan, 1.0)) and math.isnan(O.go_min(1.0, nan))
    assert O.go_min(nan, -inf) == -inf and O.go_min(-inf, nan) == -inf  # -Inf wins over NaN
    assert O.go_max(nan, inf) == inf and math.isnan(O.go_max(nan, -inf))
    assert math.copysign(1, O.go_min(0.0, -0.0)) < 0 and math.copysign(1, O.go_min(-0.0, 0.0)) < 0
    assert math.copysign(1, O.go_max(0.0, -0.0)) > 0 and math.copysign(1, O.go_max(-0.0, 0.0)) > 0
    assert O.go_min(3.0, 2.0) == 2.0 and O.go_max(3.0, 2.0) == 3.0


def test_go_float_to_int():
    i64min = -(1 << 63)
    assert O.go_f2i(2.9) == 2 and O.go_f2i(-2.9) == -2 and O.go_f2i(-0.0) == 0
    for v in (float("nan"), float("inf"), -float("inf"), 9.3e18, -9.3e18, 2.0 ** 63):
        assert O.go_f2i(v) == i64min
    assert O.go_f2i(-(2.0 ** 63)) == i64min
    assert O.go_f2i(2.0 ** 62) == 1 << 62
    # uint8(x) goes through int32: truncation, out of range / NaN -> low byte of 0x80000000
    assert O.go_u8(254.99) == 254 and O.go_u8(255.0) == 255 and O.go_u8(256.7) == 0
    assert O.go_u8(float("nan")) == 0 and O.go_u8(3e9) == 0 and O.go_u8(-1.5) == 0xFF


def test_color_nrgba_truncates_and_clamps():
    assert O.color_nrgba((1.0, 0.5, 0.0, 1.0)) == (255, 127, 0, 255)      # 127.5 truncates
    assert O.color_nrgba((2.0, -1.0, float("nan"), 0.999)) == (255, 0, 0, 254)
    # every 8-bit value v/255 round-trips
    for v in range(256):
        assert O.color_nrgba((v / 255, 0, 0, 1))[0] == v


def test_nrgba_premultiply_hand_cases():
    # color.NRGBA.RGBA(): x = v | v<<8; x = x*A/0xff; alpha = A | A<<8
    assert O.nrgba_rgba((255, 255, 255, 255)) == (65535, 65535, 65535, 65535)
    assert O.nrgba_rgba((255, 128, 0, 128)) == (32896, 16512, 0, 32896)
    assert O.nrgba_rgba((200, 100, 50, 0)) == (0, 0, 0, 0)
    assert O.nrgba_rgba((1, 2, 3, 1)) == (1, 2, 3, 257)


@pytest.mark.parametrize("alpha", [0.0, 1 / 255, 0.5, 254 / 255, 0.2, 0.8])
def test_blend_against_real_over(alpha):
    rng = np.random.Generator(np.random.PCG64(3))
    for _ in range(200):
        dst = tuple(int(x) for x in rng.integers(0, 256, 4))
        col = (float(rng.uniform()), float(rng.uniform()), float(rng.uniform()), alpha)
        out = O.blend(dst, col)
        n = O.color_nrgba(col)
        a = n[3] / 255.0
        exp = [n[k] * a + dst[k] * (1 - a) for k in range(3)] + [n[3] + dst[3] * (1 - a)]
        for k in range(4):
            assert abs(out[k] - exp[k]) <= 1.5, (dst, col, out, exp)
    # alpha 0 leaves the destination untouched exactly
    if alpha == 0.0:
        for d in range(256):
            assert O.blend((d, 255 - d, d // 2, d), (0.7, 0.2, 0.9, 0.0)) == (d, 255 - d, d // 2, d)


def test_go_pow_integer_exponents_within_the_product_error_bound():
    # x**n by binary powering is a chain of n-1 rounded products at most:
    # relative error <= (n-1) * 2**-53 to first order, i.e. <= n ulps.
    # A wrong bit test, a dropped squaring or a lost exponent is far off.
    rng = np.random.Generator(np.random.PCG64(5))
    xs = list(rng.uniform(1e-3, 1.0, 300)) + [0.5, 0.999999, 1e-5, 0.3]
    worst = 0
    for x in xs:
        for y in (2, 3, 5, 8, 16, 31, 32, 33, 64, 100):
            ref = float(F(float(x)) ** y)
            got = O.go_pow(float(x), float(y))
            if ref < 2.2250738585072014e-308:
                continue
            worst = max(worst, ulps(got, ref))
            assert ulps(got, ref) <= max(2, y), (x, y, got, ref)
            if y <= 3:
                assert ulps(got, ref) <= 1


def test_go_pow_fractional_exponents_against_exact_powers():
    # Non-integer y (shader.go:91, Go's math.Pow): y = yi + yf with yf folded
    # into (-1/2, 1/2], a1 = exp(yf * log x), times x**yi by binary powering.
    # Reference value: x**y for the exact binary x and y at 60 decimal
    # digits.  Error bound, u = 2**-53: log x and the product yf*log x are
    # each within 1 ulp, so the argument t of exp is off by at most
    # |t| * 2**-51 absolute, i.e. a relative error of |t| * 2**-51 in exp(t);
    # exp adds 1 ulp, the integer chain yi ulps (as above) and the final
    # reciprocal for y < 0 half an ulp.  In ulps of the result (an ulp is at
    # least 2**-53 relative): 4|t| + 2*yi + 4.  log10 for log, or a fold
    # that drops the yi + 1, is off by orders of magnitude more.  (The fold
    # itself changes the value only by rounding: x**y = x**(yf-1) * x**(yi+1);
    # which rounding Go's own Exp / Log give is not pinned here.)
    import decimal
    ctx = decimal.Context(prec=60)
    rng = np.random.Generator(np.random.PCG64(11))
    xs = [float(v) for v in rng.uniform(1e-3, 1.0, 120)] + [float(v) for v in rng.uniform(1.0, 40.0, 40)]
    xs += [0.5, 0.999999, 1e-5, 0.3, 3.0, 1.0000001]
    for y in (0.75, 2.5, 12.5, 33.3, -4.5, 1.25, -0.75, 7.6):
        ay = abs(y)
        yi, yf = math.floor(ay), ay - math.floor(ay)
        if yf > 0.5:
            yf, yi = yf - 1, yi + 1
        for x in xs:
            got = O.go_pow(x, y)
            ref = ctx.power(decimal.Decimal(x), decimal.Decimal(y))
            if not (2.2250738585072014e-308 < float(ref) < MAXF):
                continue
            t = abs(yf * math.log(x))
            bound = 4 * t + 2 * yi + 4
            err = abs(decimal.Decimal(got) - ref)
            ulp = decimal.Decimal(math.ulp(float(ref)))
            assert float(err / ulp) <= bound, (x, y, got, float(ref), float(err / ulp), bound)


def test_go_pow_special_cases():
    assert O.go_pow(0.3, 0.0) == 1.0 and O.go_pow(1.0, 1e300) == 1.0 and O.go_pow(0.37, 1.0) == 0.37
    assert O.go_pow(0.25, 0.5) == 0.5 and O.go_pow(4.0, -0.5) == 0.5
    assert O.go_pow(0.0, 3.0) == 0.0 and O.go_pow(0.5, float("inf")) == 0.0
    assert math.isnan(O.go_pow(-0.5, 0.5))
    assert O.go_pow(2.0, 10.0) == 1024.0 and O.go_pow(2.0, -2.0) == 0.25


# ---------------------------------------------------------------------------
# matrices (matrix.go:69-128) against the OpenGL closed forms


def test_look_at_is_glu_look_at():
    rng = np.random.Generator(np.random.PCG64(7))
    for _ in range(50):
        eye, center = rng.uniform(-5, 5, 3), rng.uniform(-1, 1, 3)
        up = np.array([0.0, 1.0, 0.0])
        m = O.look_at(eye, center, up).reshape(4, 4)
        f = (center - eye) / np.linalg.norm(center - eye)
        s = np.cross(f, up)
        s /= np.linalg.norm(s)
        u = np.cross(s, f)
        glu = np.eye(4)
        glu[0, :3], glu[1, :3], glu[2, :3] = s, u, -f
        glu[:3, 3] = -glu[:3, :3] @ eye
        assert np.allclose(m, glu, rtol=0, atol=1e-12)


def test_frustum_perspective_screen_closed_forms():
    l, r, b, t, n, f = -0.7, 0.9, -0.4, 0.6, 1.0, 10.0
    m = O.frustum(l, r, b, t, n, f).reshape(4, 4)
    gl = np.array([[2 * n / (r - l), 0, (r + l) / (r - l), 0], [0, 2 * n / (t - b), (t + b) / (t - b), 0],
                   [0, 0, -(f + n) / (f - n), -2 * f * n / (f - n)], [0, 0, -1, 0]])
    assert np.allclose(m, gl, rtol=1e-15, atol=0)
    p = O.perspective(30, 16 / 9, 1, 10).reshape(4, 4)
    ymax = math.tan(math.radians(15))
    assert math.isclose(p[1, 1], 1 / ymax, rel_tol=1e-15) and math.isclose(p[0, 0], 1 / (ymax * 16 / 9),
                                                                           rel_tol=1e-15)
    # near plane maps to z = -1, far plane to z = +1 in NDC
    for zeye, want in ((-1.0, -1.0), (-10.0, 1.0)):
        v = O.mul_position_w(p.ravel(), (0.0, 0.0, zeye))
        assert math.isclose(v[2] / v[3], want, abs_tol=1e-15)
    s = O.screen(640, 480).reshape(4, 4)
    for ndc, scr in (((-1, -1, -1), (0, 480, 0)), ((1, 1, 1), (640, 0, 1)), ((0, 0, 0), (320, 240, 0.5))):
        v = s @ np.array([*ndc, 1.0])
        assert np.allclose(v[:3], scr)


def test_matrix_mul_and_position_exact_on_integers():
    rng = np.random.Generator(np.random.PCG64(8))
    a = rng.integers(-9, 9, 16).astype(float)
    b = rng.integers(-9, 9, 16).astype(float)
    assert np.array_equal(O.matrix_mul(a, b), (a.reshape(4, 4) @ b.reshape(4, 4)).ravel())
    p = rng.integers(-9, 9, 3).astype(float)
    assert np.array_equal(O.mul_position_w(a, p), a.reshape(4, 4) @ np.array([*p, 1.0]))


# ---------------------------------------------------------------------------
# Barycentric (vertex.go:81-95)


def test_barycentric_vertices_and_reconstruction():
    rng = np.random.Generator(np.random.PCG64(9))
    for _ in range(200):
        p1, p2, p3 = rng.uniform(-3, 3, (3, 3))
        assert np.array_equal(O.barycentric(p1, p2, p3, p1), [1, 0, 0, 1])
        assert np.array_equal(O.barycentric(p1, p2, p3, p2)[1:], [1, 0, 1])
        u, v = rng.uniform(0, 1, 2)
        w = 1 - u - v
        p = u * p1 + v * p2 + w * p3
        b = O.barycentric(p1, p2, p3, p)
        assert b[3] == 1
        assert np.allclose(b[0] * p1 + b[1] * p2 + b[2] * p3, p, atol=1e-9)
        assert np.allclose(b[:3], [u, v, w], atol=1e-7)


# ---------------------------------------------------------------------------
# clipping (clipping.go) against exact rational Sutherland-Hodgman

PLANES = [((1, 0, 0, 1), (-1, 0, 0, 1)), ((-1, 0, 0, 1), (1, 0, 0, 1)), ((0, 1, 0, 1), (0, -1, 0, 1)),
          ((0, -1, 0, 1), (0, 1, 0, 1)), ((0, 0, 1, 1), (0, 0, -1, 1)), ((0, 0, -1, 1), (0, 0, 1, 1))]


def exact_clip(tri):
    pts = [tuple(F(float(c)) for c in p) for p in tri]
    for P, N in PLANES:
        if not pts:
            return []
        out = []
        front = lambda v: sum((v[k] - P[k]) * N[k] for k in range(4)) > 0  # noqa: E731
        s = pts[-1]
        for e in pts:
            if front(e):
                if not front(s):
                    out.append(isect(s, e, P, N))
                out.append(e)
            elif front(s):
                out.append(isect(s, e, P, N))
            s = e
        pts = out
    return pts


def isect(v0, v1, P, N):
    u = [v1[k] - v0[k] for k in range(4)]
    w = [v0[k] - P[k] for k in range(4)]
    d = sum(N[k] * u[k] for k in range(4))
    n = -sum(N[k] * w[k] for k in range(4))
    return tuple(v0[k] + u[k] * (n / d) for k in range(4))


def test_clip_polygon_matches_exact_rational():
    rng = np.random.Generator(np.random.PCG64(10))
    checked = 0
    for _ in range(400):
        tri = rng.uniform(-2.5, 2.5, (3, 4))
        tri[:, 3] = rng.uniform(0.5, 2.0, 3)
        got = O.clip_polygon(tri)
        want = exact_clip(tri)
        assert len(got) == len(want), (tri, got, want)
        for g, w in zip(got, want):
            assert np.allclose(g, [float(x) for x in w], rtol=0, atol=1e-12)
        for p in got:  # every output point satisfies every plane within rounding
            for P, N in PLANES:
                assert sum((p[k] - P[k]) * N[k] for k in range(4)) >= -1e-12
        checked += len(got) >= 3
    assert checked > 100


def test_clip_inside_triangle_unchanged_and_fan_count():
    tri = np.array([[0.1, 0.2, 0.3, 1.0], [-0.5, 0.1, 0.0, 1.0], [0.3, -0.4, -0.2, 1.0]])
    out = O.clip_polygon(tri)
    assert len(out) == 3 and np.array_equal(out, tri)
    # one vertex behind the near plane: the polygon gains one vertex
    tri2 = tri.copy()
    tri2[0] = [0.1, 0.2, -3.0, 1.0]
    assert len(O.clip_polygon(tri2)) == 4
    # fully outside: nothing
    tri3 = tri + np.array([5.0, 0, 0, 0])
    assert len(O.clip_polygon(tri3)) == 0


def test_clip_line_exact():
    w1, w2 = np.array([-3.0, 0.2, 0.1, 1.0]), np.array([0.5, -0.1, 0.2, 1.0])
    out = O.clip_line(w1, w2)
    assert out is not None and math.isclose(out[0][0], -1.0, abs_tol=1e-15) and np.array_equal(out[1], w2)
    assert O.clip_line(np.array([2.0, 0, 0, 1]), np.array([3.0, 0, 0, 1])) is None


# ---------------------------------------------------------------------------
# rasterize() coverage and depth against exact arithmetic


def exact_cover(s0, s1, s2, W, H):
    """Pixels whose centre lies inside or on the triangle, exactly, and the
    minimum |edge| margin of each (in squared-pixel units).  Like the
    reference, a centre (x, y) outside the screen still lands on pixel
    i = y*W + x when 0 <= i < W*H (context.go:223-228)."""
    S = [tuple(F(c) for c in s) for s in (s0, s1, s2)]

    def edge(a, b, c):
        return (b[0] - c[0]) * (a[1] - c[1]) - (b[1] - c[1]) * (a[0] - c[0])

    area = edge(S[0], S[1], S[2])
    cov = set()
    margin = {(x, y): F(10 ** 9) for y in range(H) for x in range(W)}
    if area == 0:
        return cov, margin, area
    xs = [float(c[0]) for c in S]
    ys = [float(c[1]) for c in S]
    for y in range(math.floor(min(ys)), math.ceil(max(ys)) + 1):
        for x in range(math.floor(min(xs)), math.ceil(max(xs)) + 1):
            i = y * W + x
            if i < 0 or i >= W * H:
                continue
            key = (i % W, i // W)
            p = (F(2 * x + 1, 2), F(2 * y + 1, 2))
            w = [edge(S[1], S[2], p) / area, edge(S[2], S[0], p) / area, edge(S[0], S[1], p) / area]
            margin[key] = min(margin[key], min(abs(v) for v in w) * abs(area))
            if all(v >= 0 for v in w):
                cov.add(key)
    return cov, margin, area


def rasterize_solid(W, H, s0, s1, s2, wts=(1.0, 1.0, 1.0)):
    ctx = O.Context(W, H)
    ctx.set_state(O.make_state(cull=O.CULL_NONE))
    ctx.set_shader(O.make_shader(O.SHADER_SOLID, np.eye(4).ravel(), color=(1, 1, 1, 1)))
    info = ctx.rasterize_screen(s0, s1, s2, wts)
    return ctx, info


def positive(s):
    """Orders the vertices so that edge(s0, s1, s2) > 0, the orientation
    every rasterize() call of the reference has (drawClippedTriangle swaps
    v0/v2 when the NDC area is negative and the screen map flips y; line
    quads are built with this orientation).  rasterize()'s skip of the
    pixels left of the triangle assumes it."""
    e = (s[1, 0] - s[2, 0]) * (s[0, 1] - s[2, 1]) - (s[1, 1] - s[2, 1]) * (s[0, 0] - s[2, 0])
    return s[[2, 1, 0]] if e < 0 else s


def test_coverage_matches_exact_rational_small_triangles():
    rng = np.random.Generator(np.random.PCG64(11))
    W = H = 12
    for _ in range(250):
        s = rng.uniform(-1, 13, (3, 3))
        s[:, 2] = rng.uniform(0, 1, 3)
        s = positive(s)
        ctx, info = rasterize_solid(W, H, *s)
        cov, margin, area = exact_cover(s[0], s[1], s[2], W, H)
        got = set(zip(*np.nonzero(ctx.depth().T < MAXF)))
        got = {(int(x), int(y)) for x, y in got}
        diff = got ^ cov
        # only pixels whose centre is within rounding of an edge may differ
        assert all(margin[p] < 1e-9 for p in diff), (s, sorted(diff))
        if not diff:
            assert info[0] == len(cov) and info[1] == len(cov)


def test_depth_is_the_plane_through_the_vertices():
    rng = np.random.Generator(np.random.PCG64(12))
    W = H = 16
    for _ in range(60):
        s = rng.uniform(0, 16, (3, 3))
        s[:, 2] = rng.uniform(-1, 1, 3)
        s = positive(s)
        ctx, _ = rasterize_solid(W, H, *s)
        d = ctx.depth()
        S = [tuple(F(c) for c in v) for v in s]
        area = (S[1][0] - S[2][0]) * (S[0][1] - S[2][1]) - (S[1][1] - S[2][1]) * (S[0][0] - S[2][0])
        if area == 0:
            continue
        for y, x in zip(*np.nonzero(d < MAXF)):
            p = (F(2 * int(x) + 1, 2), F(2 * int(y) + 1, 2))
            b0 = ((S[2][0] - p[0]) * (S[1][1] - p[1]) - (S[2][1] - p[1]) * (S[1][0] - p[0])) / area
            b1 = ((S[0][0] - p[0]) * (S[2][1] - p[1]) - (S[0][1] - p[1]) * (S[2][0] - p[0])) / area
            b2 = 1 - b0 - b1
            z = float(b0 * S[0][2] + b1 * S[1][2] + b2 * S[2][2])
            assert abs(d[y, x] - z) < 1e-9, (s, x, y, d[y, x], z)


def test_both_windings_cover_the_exact_triangle():
    # DrawTriangle with CullNone: a triangle and its reversed-winding twin
    # cover exactly the pixels whose centres lie in the screen triangle
    # (drawClippedTriangle's v0/v2 swap makes both reach rasterize() with
    # the same orientation).
    W, H = 24, 20
    rng = np.random.Generator(np.random.PCG64(15))
    S = O.screen(W, H).reshape(4, 4)
    for _ in range(30):
        ndc = rng.uniform(-0.95, 0.95, (3, 3))
        for order in ([0, 1, 2], [2, 1, 0]):
            ctx = O.Context(W, H)
            ctx.set_state(O.make_state(cull=O.CULL_NONE))
            ctx.set_shader(O.make_shader(O.SHADER_SOLID, np.eye(4).ravel(), color=(1, 1, 1, 1)))
            ctx.draw_triangles(O.Mesh(tris={"position": ndc[order][None]}))
            got = {(int(x), int(y)) for y, x in zip(*np.nonzero(ctx.depth() < MAXF))}
            scr = np.array([(S @ np.array([*p, 1.0]))[:3] for p in ndc])
            cov, margin, _ = exact_cover(*positive(scr), W, H)
            assert all(margin[p] < 1e-9 for p in got ^ cov), (ndc, order, sorted(got ^ cov))
            assert len(cov) == 0 or len(got) > 0


def test_perspective_correct_interpolation_of_a_constant():
    # constant vertex color under wildly different W: every fragment keeps it
    W = H = 24
    ctx = O.Context(W, H)
    ctx.set_state(O.make_state(cull=O.CULL_NONE))
    ctx.set_shader(O.make_shader(O.SHADER_PHONG, np.eye(4).ravel(), light_dir=(0, 0, 1), ambient=(1, 1, 1, 1),
                                 diffuse=(0, 0, 0, 0), specular=(0, 0, 0, 0), specular_power=0))
    pos = np.array([[[-0.9, -0.9, 0.1], [0.9, -0.8, 0.2], [0.1, 0.9, 0.3]]])
    col = np.full((1, 3, 4), 0.25)
    col[..., 3] = 1.0
    m = O.Mesh(tris={"position": pos, "normal": np.tile([0, 0, 1.0], (1, 3, 1)), "color": col})
    ctx.draw_triangles(m)
    img = ctx.color()
    cov = ctx.depth() < MAXF
    assert cov.sum() > 100
    assert np.all(img[cov][:, :3] == 63) and np.all(img[cov][:, 3] == 255)


def test_counters_total_and_updated():
    # two identical triangles: the second passes the <= test everywhere
    W = H = 20
    s = positive(np.array([[1.0, 1.0, 0.5], [18.5, 2.0, 0.5], [5.0, 17.0, 0.5]]))
    ctx, info1 = rasterize_solid(W, H, *s)
    cov, _, _ = exact_cover(*s, W, H)
    assert info1 == (len(cov), len(cov))
    info2 = ctx.rasterize_screen(*s, (1, 1, 1))
    assert info2 == (len(cov), len(cov))
    # a farther copy: counted in TotalPixels, never updated
    s2 = s.copy()
    s2[:, 2] = 0.75
    assert ctx.rasterize_screen(*s2, (1, 1, 1)) == (len(cov), 0)


# ---------------------------------------------------------------------------
# shaders (shader.go:75-96) and textures (texture.go:41-63)


def phong(**kw):
    base = dict(light_dir=(0, 0, 1), camera_pos=(0, 0, 5), object_color=(0.4, 0.6, 0.8, 1.0))
    base.update(kw)
    return O.make_shader(O.SHADER_PHONG, np.eye(4).ravel(), **base)


def test_phong_special_cases():
    # N == L, power 0: ambient + diffuse = 1 -> the object color itself
    c = O.phong_fragment(phong(specular_power=0), (0, 0, 0), (0, 0, 1))
    assert np.allclose(c, [0.4, 0.6, 0.8, 1.0], atol=1e-15)
    # N . L < 0: ambient only
    c = O.phong_fragment(phong(), (0, 0, 0), (0, 0, -1))
    assert np.allclose(c, [0.08, 0.12, 0.16, 1.0], atol=1e-15)
    # camera on the reflection ray: light = 0.2 + 0.8 + 1 = 2 per channel,
    # then Min(White) clamps
    c = O.phong_fragment(phong(), (0, 0, 0), (0, 0, 1))
    assert np.array_equal(c, [0.8, 1.0, 1.0, 1.0])
    # Discard object color -> the interpolated vertex color (alpha kept)
    c = O.phong_fragment(phong(object_color=(0, 0, 0, 0), specular_power=0), (0, 0, 0), (0, 0, 1),
                         col=(0.5, 0.25, 0.125, 0.5))
    assert np.allclose(c, [0.5, 0.25, 0.125, 0.5])
    # diffuse at 60 degrees is cos = 0.5
    n = (math.sin(math.radians(60)), 0, math.cos(math.radians(60)))
    c = O.phong_fragment(phong(specular_power=0, object_color=(1, 1, 1, 1)), (0, 0, 0), n)
    assert np.allclose(c[:3], 0.2 + 0.8 * 0.5)


def test_bilinear_constant_and_hand_cases():
    const = np.zeros((5, 7, 4), np.uint8)
    const[...] = (40, 80, 120, 255)
    t = O.Texture(const, O.TEX_NRGBA8)
    for uv in ((0.3, 0.7), (0.0, 0.0), (-2.25, 3.5), (0.999, 0.001)):
        c = O.bilinear_sample(t, *uv)
        assert np.allclose(c, np.array([40, 80, 120, 255]) * 257 / 65535, atol=1e-15)
    tex = np.zeros((2, 2, 4), np.uint8)
    tex[0, 0] = (255, 0, 0, 255)
    tex[0, 1] = (0, 255, 0, 255)
    tex[1, 0] = (0, 0, 255, 255)
    tex[1, 1] = (255, 255, 255, 255)
    t = O.Texture(tex, O.TEX_NRGBA8)
    assert np.array_equal(O.bilinear_sample(t, 0.0, 1.0), [1, 0, 0, 1])  # v = 1 is image row 0
    # v = 0 wraps to image row 0 as well (1 - 0 = 1, minus floor(1))
    assert np.array_equal(O.bilinear_sample(t, 0.0, 0.0), [1, 0, 0, 1])
    assert np.array_equal(O.bilinear_sample(t, 0.0, 1e-300), [1, 0, 0, 1])  # 1 - 1e-300 rounds to 1
    assert np.allclose(O.bilinear_sample(t, 0.5, 0.5), [0.5, 0.5, 0.5, 1.0])
    # x = y = 0.25: weights 9/16 red, 3/16 green, 3/16 blue, 1/16 white
    assert np.allclose(O.bilinear_sample(t, 0.25, 0.75), [0.625, 0.25, 0.25, 1.0])
    # x = 0.25, y = 0.5: 3/8 red, 1/8 green (x1, y0), 3/8 blue (x0, y1), 1/8 white
    assert np.allclose(O.bilinear_sample(t, 0.25, 0.5), [0.5, 0.25, 0.5, 1.0])
    # u wraps: 1.0 samples column 0 again
    assert np.array_equal(O.bilinear_sample(t, 1.0, 1.0), O.bilinear_sample(t, 0.0, 1.0))
    # NRGBA premultiplies at sampling; RGBA does not
    tex2 = np.array([[[200, 100, 50, 128]]], np.uint8)
    cn = O.bilinear_sample(O.Texture(tex2, O.TEX_NRGBA8), 0.5, 0.5)
    cr = O.bilinear_sample(O.Texture(tex2, O.TEX_RGBA8), 0.5, 0.5)
    assert np.allclose(cn[:3] * 65535, [(200 * 257) * 128 // 255, (100 * 257) * 128 // 255,
                                        (50 * 257) * 128 // 255])
    assert np.allclose(cr * 65535, [200 * 257, 100 * 257, 50 * 257, 128 * 257])


# ---------------------------------------------------------------------------
# DepthImage (context.go:87-117)


def test_depth_image_closed_form():
    W, H = 5, 3
    ctx = O.Context(W, H)
    s = positive(np.array([[0.0, 0.0, 0.2], [5.0, 0.0, 0.6], [0.0, 3.0, 0.9]]))
    ctx.set_state(O.make_state(cull=O.CULL_NONE))
    ctx.set_shader(O.make_shader(O.SHADER_SOLID, np.eye(4).ravel(), color=(1, 1, 1, 1)))
    ctx.rasterize_screen(*s, (1, 1, 1))
    d = ctx.depth()
    img = ctx.depth_image()
    lo, hi = d[d < MAXF].min(), d[d < MAXF].max()
    for y in range(H):
        for x in range(W):
            want = 0xFFFF if d[y, x] == MAXF else int((d[y, x] - lo) / (hi - lo) * 0xFFFF)
            assert img[y, x] == want


# ---------------------------------------------------------------------------
# scene invariants


def _scene(W, H, pos, col, blend, order):
    ctx = O.Context(W, H)
    ctx.clear_color((0.1, 0.1, 0.1, 1))
    ctx.set_state(O.make_state(cull=O.CULL_NONE))
    ctx.set_shader(O.make_shader(O.SHADER_PHONG, np.eye(4).ravel(), light_dir=(0, 0, 1), ambient=(1, 1, 1, 1),
                                 diffuse=(0, 0, 0, 0), specular=(0, 0, 0, 0), specular_power=0))
    c = col.copy()
    c[..., 3] = 0.5 if blend else 1.0
    m = O.Mesh(tris={"position": pos[order], "normal": np.tile([0, 0, 1.0], (len(order), 3, 1)),
                     "color": c[order]})
    ctx.draw_triangles(m)
    return ctx.color(), ctx.depth()


def test_opaque_order_independent_blend_order_dependent():
    rng = np.random.Generator(np.random.PCG64(13))
    n = 60
    pos = rng.uniform(-1, 1, (n, 3, 3))
    pos[..., 2] = np.repeat(rng.uniform(-0.9, 0.9, (n, 1)), 3, axis=1)  # flat, distinct depths
    col = rng.uniform(0, 1, (n, 3, 4))
    col[...] = col[:, :1]
    order = np.arange(n)
    perm = rng.permutation(n)
    c1, d1 = _scene(40, 40, pos, col, False, order)
    c2, d2 = _scene(40, 40, pos, col, False, perm)
    assert np.array_equal(d1, d2) and np.array_equal(c1, c2)
    c3, _ = _scene(40, 40, pos, col, True, order)
    c4, _ = _scene(40, 40, pos, col, True, perm)
    assert not np.array_equal(c3, c4)


def test_threaded_mirror_matches_on_opaque_unique_depths():
    rng = np.random.Generator(np.random.PCG64(14))
    n = 300
    pos = rng.uniform(-1, 1, (n, 3, 3))
    pos[..., 2] = np.repeat(rng.uniform(-0.9, 0.9, (n, 1)), 3, axis=1)
    m = O.Mesh(tris={"position": pos})
    infos, depths = [], []
    for threads in (1, 4):
        ctx = O.Context(64, 64)
        ctx.set_state(O.make_state(cull=O.CULL_NONE))
        ctx.set_shader(O.make_shader(O.SHADER_SOLID, np.eye(4).ravel(), color=(1, 0, 0, 1)))
        infos.append(ctx.draw_triangles(m, threads))
        depths.append(ctx.depth())
    assert infos[0][0] == infos[1][0]
    assert np.array_equal(depths[0], depths[1])


# ---------------------------------------------------------------------------
# SSAA resolve (the build's own filter; pinned by its closed forms)

def _tent_exact(src, dst):
    """The tent table in exact rationals: centre c = s*(o+1/2) - 1/2,
    start = trunc(c) - ceil(s) + 1, weight floor(max(0, 1 - |c - start - t|/s) * 256)."""
    s = F(src, dst)
    half = max(math.ceil(s), 1)
    w = np.zeros((dst, 2 * half), dtype=np.int64)
    st = np.zeros(dst, dtype=np.int64)
    for o in range(dst):
        c = s * (o + F(1, 2)) - F(1, 2)
        st[o] = int(c) - half + 1  # int() truncates toward zero like Go/C
        for t in range(2 * half):
            x = abs(c - st[o] - t) / s
            w[o, t] = math.floor(max(F(0), 1 - x) * 256)
    return w, st


@pytest.mark.parametrize("f,dst", [(1, 5), (2, 7), (3, 4), (4, 6), (8, 3)])
def test_resolve_tent_weights_closed_form(f, dst):
    w, st = O.tent_weights(f * dst, dst)
    we, ste = _tent_exact(f * dst, dst)
    assert np.array_equal(st, ste)
    assert np.array_equal(w, we)
    if f == 4:
        assert w[0].tolist() == [32, 96, 160, 224, 224, 160, 96, 32]
        assert st.tolist() == [4 * o - 2 for o in range(dst)]


@pytest.mark.parametrize("filt", [1, 2])
@pytest.mark.parametrize("a", [255, 128, 7, 0])
def test_resolve_constant_image(filt, a):
    c = np.array([200, 31, 97, a], np.uint8)
    img = np.broadcast_to(c, (12, 20, 4)).copy()
    out = O.resolve(img, 5, 3, filt)
    if filt == 2:
        exp = c  # the mean of a constant
    else:
        # tent output is premultiplied: c * a / 255 with integer division
        exp = np.array([200 * a // 255, 31 * a // 255, 97 * a // 255, a], np.uint8)
    assert (out == exp).all()


def test_resolve_box_is_the_rounded_block_mean():
    rng = np.random.Generator(np.random.PCG64(21))
    for f in (2, 3, 4):
        img = rng.integers(0, 256, (3 * f, 5 * f, 4), dtype=np.uint8)
        out = O.resolve(img, 5, 3, 2)
        blocks = img.reshape(3, f, 5, f, 4).astype(np.int64).sum(axis=(1, 3))
        exp = (blocks + (f * f) // 2) // (f * f)
        assert np.array_equal(out, exp.astype(np.uint8))


def test_resolve_tent_separable_hand_case():
    # opaque image varying along x only: every output row is the 1-D
    # horizontal filter; 4x with taps [32,96,160,224,224,160,96,32] and
    # edge clamping, computed by hand for the first output pixel:
    # sources -2..5 clamp to 0,0,0,1,2,3,4,5: 16960 // 1024 = 16
    f, dw, dh = 4, 3, 2
    row = np.array([0, 10, 20, 30, 40, 50, 60, 70, 80, 90, 100, 110], np.int64)
    img = np.zeros((dh * f, dw * f, 4), np.uint8)
    img[..., 0] = row[None, :]
    img[..., 3] = 255
    out = O.resolve(img, dw, dh, 1)
    taps = [32, 96, 160, 224, 224, 160, 96, 32]
    first = sum(t * row[max(0, min(11, -2 + k))] for k, t in enumerate(taps)) // 1024
    assert first == (32 * 0 + 96 * 0 + 160 * 0 + 224 * 10 + 224 * 20 + 160 * 30 + 96 * 40 + 32 * 50) // 1024 == 16
    assert (out[:, 0, 0] == first).all()
    mid = sum(t * row[max(0, min(11, 2 + k))] for k, t in enumerate(taps)) // 1024
    assert (out[:, 1, 0] == mid).all() and mid == 55
    # the same image transposed: the vertical pass gives the same numbers
    outT = O.resolve(np.ascontiguousarray(img.transpose(1, 0, 2)), dh, dw, 1)
    assert np.array_equal(outT[:, :, 0].T, out[:, :, 0])
