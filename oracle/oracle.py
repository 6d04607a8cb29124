"""ctypes binding of the CPU oracle (liboracle.so).

Test infrastructure: imported only by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's CPU-baseline leg.  The product library never
uses it.  See fauxgl_oracle.h for what the oracle restates and what pins it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB points the binding at another build of the oracle (the mutation
# tests use it to load builds with planted mistakes).
LIB = os.environ.get("ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

SHADER_SOLID, SHADER_TEXTURE, SHADER_PHONG = 1, 2, 3
TEX_NRGBA8, TEX_RGBA8, TEX_RGBA16 = 1, 2, 3
FACE_CW, FACE_CCW = 1, 2
CULL_NONE, CULL_FRONT, CULL_BACK = 1, 2, 3
CNT_NAMES = ["input", "raster", "rows", "scanned", "inside", "shaded", "clipped", "skipped"]

_d = ctypes.c_double
_P = ctypes.POINTER


class OrcTexture(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("kind", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("pix", ctypes.c_void_p)]


class OrcShader(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("pad", ctypes.c_int32), ("matrix", _d * 16), ("color", _d * 4),
                ("light_dir", _d * 3), ("camera_pos", _d * 3), ("object_color", _d * 4), ("ambient", _d * 4),
                ("diffuse", _d * 4), ("specular", _d * 4), ("specular_power", _d),
                ("texture", _P(OrcTexture))]


class OrcState(ctypes.Structure):
    _fields_ = [("read_depth", ctypes.c_int32), ("write_depth", ctypes.c_int32), ("write_color", ctypes.c_int32),
                ("alpha_blend", ctypes.c_int32), ("wireframe", ctypes.c_int32), ("front_face", ctypes.c_int32),
                ("cull", ctypes.c_int32), ("pad", ctypes.c_int32), ("line_width", _d), ("depth_bias", _d)]


def build():
    """Compiles liboracle.so with the committed Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    L = ctypes.CDLL(LIB)
    vp = ctypes.c_void_p
    dp = _P(_d)
    u64p = _P(ctypes.c_uint64)
    sigs = {
        "orc_ctx_new": (vp, [ctypes.c_int32, ctypes.c_int32]),
        "orc_ctx_free": (None, [vp]),
        "orc_set_state": (None, [vp, _P(OrcState)]),
        "orc_set_shader": (None, [vp, _P(OrcShader)]),
        "orc_clear_color": (None, [vp, dp]),
        "orc_clear_depth": (None, [vp, _d]),
        "orc_color": (_P(ctypes.c_uint8), [vp]),
        "orc_depth": (dp, [vp]),
        "orc_depth_image": (None, [vp, _P(ctypes.c_uint16)]),
        "orc_resolve": (None, [_P(ctypes.c_uint8), ctypes.c_int32, ctypes.c_int32, _P(ctypes.c_uint8), ctypes.c_int32,
                               ctypes.c_int32, ctypes.c_int32]),
        "orc_tent_weights": (ctypes.c_int32, [ctypes.c_int32, ctypes.c_int32, _P(ctypes.c_int16),
                                              _P(ctypes.c_int32)]),
        "orc_counters": (None, [vp, u64p]),
        "orc_reset_counters": (None, [vp]),
        "orc_mesh_new": (vp, [ctypes.c_int64, dp, dp, dp, dp, ctypes.c_int64, dp, dp, dp, dp]),
        "orc_mesh_free": (None, [vp]),
        "orc_draw_triangles": (None, [vp, vp, ctypes.c_int32, u64p]),
        "orc_draw_lines": (None, [vp, vp, ctypes.c_int32, u64p]),
        "orc_draw_mesh": (None, [vp, vp, ctypes.c_int32, u64p]),
        "orc_go_min": (_d, [_d, _d]),
        "orc_go_max": (_d, [_d, _d]),
        "orc_go_f2i": (ctypes.c_int64, [_d]),
        "orc_go_pow": (_d, [_d, _d]),
        "orc_go_u8": (ctypes.c_uint8, [_d]),
        "orc_color_nrgba": (None, [dp, _P(ctypes.c_uint8)]),
        "orc_nrgba_rgba": (None, [_P(ctypes.c_uint8), _P(ctypes.c_uint32)]),
        "orc_blend": (None, [_P(ctypes.c_uint8), dp]),
        "orc_barycentric": (None, [dp, dp, dp, dp, dp]),
        "orc_clip_polygon": (ctypes.c_int32, [dp, dp]),
        "orc_clip_line": (ctypes.c_int32, [dp, dp, dp]),
        "orc_matrix_mul": (None, [dp, dp, dp]),
        "orc_mul_position_w": (None, [dp, dp, dp]),
        "orc_look_at": (None, [dp, dp, dp, dp]),
        "orc_frustum": (None, [_d, _d, _d, _d, _d, _d, dp]),
        "orc_perspective": (None, [_d, _d, _d, _d, dp]),
        "orc_screen": (None, [ctypes.c_int32, ctypes.c_int32, dp]),
        "orc_bilinear_sample": (None, [_P(OrcTexture), _d, _d, dp]),
        "orc_phong_fragment": (None, [_P(OrcShader), dp, dp, dp, dp, dp]),
        "orc_rasterize_screen": (None, [vp, dp, dp, dp, dp, u64p]),
        "orc_mesh_bounding_box": (None, [dp, ctypes.c_int64, dp, ctypes.c_int64, dp]),
        "orc_mesh_transform": (None, [dp, dp, ctypes.c_int64, dp]),
        "orc_mesh_fit_inside": (None, [dp, dp, ctypes.c_int64, dp, dp, ctypes.c_int64, dp, dp, dp]),
        "orc_mesh_smooth_normals": (None, [dp, dp, ctypes.c_int64]),
        "orc_mesh_smooth_normals_threshold": (None, [dp, dp, ctypes.c_int64, _d]),
        "orc_load_stl": (ctypes.c_int, [_P(ctypes.c_uint8), ctypes.c_int64, _P(ctypes.c_int64), dp, dp]),
        "orc_load_obj": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int64, _P(ctypes.c_int64), dp, dp, dp]),
        "orc_png_filter": (ctypes.c_int32, [_P(ctypes.c_uint8), ctypes.c_int32, ctypes.c_int32, _P(ctypes.c_uint8)]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _dptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_P(_d))


def _arr(a, cols):
    if a is None:
        return None
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64)).reshape(-1, cols)


def darr(vals, n=None):
    a = np.ascontiguousarray(np.asarray(vals, dtype=np.float64))
    if n is not None:
        assert a.size == n
    return a


class Texture:
    """Keeps the texel array alive while the oracle samples it."""

    def __init__(self, texels, kind):
        dtype = np.uint16 if kind == TEX_RGBA16 else np.uint8
        self.pix = np.ascontiguousarray(np.asarray(texels, dtype=dtype))
        h, w = self.pix.shape[:2]
        self.s = OrcTexture(w, h, kind, 0, self.pix.ctypes.data_as(ctypes.c_void_p))


def make_shader(kind, matrix, color=(1, 0, 1, 1), light_dir=(0, 0, 0), camera_pos=(0, 0, 0),
                object_color=(0, 0, 0, 0), ambient=(0.2, 0.2, 0.2, 1), diffuse=(0.8, 0.8, 0.8, 1),
                specular=(1, 1, 1, 1), specular_power=32.0, texture=None):
    s = OrcShader()
    s.kind = kind
    s.matrix[:] = [float(x) for x in matrix]
    s.color[:] = [float(x) for x in color]
    s.light_dir[:] = [float(x) for x in light_dir]
    s.camera_pos[:] = [float(x) for x in camera_pos]
    s.object_color[:] = [float(x) for x in object_color]
    s.ambient[:] = [float(x) for x in ambient]
    s.diffuse[:] = [float(x) for x in diffuse]
    s.specular[:] = [float(x) for x in specular]
    s.specular_power = float(specular_power)
    s.texture = ctypes.pointer(texture.s) if texture is not None else None
    s._keep = texture
    return s


def make_state(read_depth=True, write_depth=True, write_color=True, alpha_blend=True, wireframe=False,
               front_face=FACE_CCW, cull=CULL_BACK, line_width=2.0, depth_bias=0.0):
    return OrcState(int(read_depth), int(write_depth), int(write_color), int(alpha_blend), int(wireframe),
                    int(front_face), int(cull), 0, float(line_width), float(depth_bias))


class Mesh:
    """Triangle and line soups built into the oracle's Vertex structs."""

    def __init__(self, tris=None, lines=None):
        L = lib()
        t = tris or {}
        l = lines or {}
        self._keep = []

        def prep(d, key, cols):
            a = d.get(key)
            if a is None:
                return None
            a = _arr(a, cols)
            self._keep.append(a)
            return a

        tp, tn, tu, tc = (prep(t, "position", 3), prep(t, "normal", 3), prep(t, "texture", 2),
                          prep(t, "color", 4))
        lp, ln, lu, lc = (prep(l, "position", 3), prep(l, "normal", 3), prep(l, "texture", 2),
                          prep(l, "color", 4))
        self.ntri = 0 if tp is None else tp.shape[0] // 3
        self.nline = 0 if lp is None else lp.shape[0] // 2
        self.h = L.orc_mesh_new(self.ntri, _dptr(tp), _dptr(tn), _dptr(tu), _dptr(tc),
                                self.nline, _dptr(lp), _dptr(ln), _dptr(lu), _dptr(lc))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mesh_free(self.h)
            self.h = None


class Context:
    """A W x H oracle framebuffer (NewContext, context.go:60-81)."""

    def __init__(self, w, h):
        self.w, self.h = w, h
        self.c = lib().orc_ctx_new(w, h)

    def __del__(self):
        if getattr(self, "c", None):
            lib().orc_ctx_free(self.c)
            self.c = None

    def set_state(self, st):
        lib().orc_set_state(self.c, ctypes.byref(st))

    def set_shader(self, sh):
        self._shader = sh
        lib().orc_set_shader(self.c, ctypes.byref(sh))

    def clear_color(self, rgba):
        a = darr(rgba, 4)
        lib().orc_clear_color(self.c, _dptr(a))

    def clear_depth(self, v):
        lib().orc_clear_depth(self.c, float(v))

    def draw_triangles(self, mesh, nthreads=1):
        info = (ctypes.c_uint64 * 2)()
        lib().orc_draw_triangles(self.c, mesh.h, nthreads, info)
        return int(info[0]), int(info[1])

    def draw_lines(self, mesh, nthreads=1):
        info = (ctypes.c_uint64 * 2)()
        lib().orc_draw_lines(self.c, mesh.h, nthreads, info)
        return int(info[0]), int(info[1])

    def draw_mesh(self, mesh, nthreads=1):
        info = (ctypes.c_uint64 * 2)()
        lib().orc_draw_mesh(self.c, mesh.h, nthreads, info)
        return int(info[0]), int(info[1])

    def color(self):
        p = lib().orc_color(self.c)
        return np.ctypeslib.as_array(p, shape=(self.h, self.w, 4)).copy()

    def depth(self):
        p = lib().orc_depth(self.c)
        return np.ctypeslib.as_array(p, shape=(self.h, self.w)).copy()

    def depth_image(self):
        out = np.zeros((self.h, self.w), dtype=np.uint16)
        lib().orc_depth_image(self.c, out.ctypes.data_as(_P(ctypes.c_uint16)))
        return out

    def counters(self):
        out = (ctypes.c_uint64 * 8)()
        lib().orc_counters(self.c, out)
        return dict(zip(CNT_NAMES, [int(x) for x in out]))

    def reset_counters(self):
        lib().orc_reset_counters(self.c)

    def rasterize_screen(self, s0, s1, s2, w=(1.0, 1.0, 1.0)):
        info = (ctypes.c_uint64 * 2)()
        a, b, c, ww = darr(s0, 3), darr(s1, 3), darr(s2, 3), darr(w, 3)
        lib().orc_rasterize_screen(self.c, _dptr(a), _dptr(b), _dptr(c), _dptr(ww), info)
        return int(info[0]), int(info[1])


# ---- single functions ----

def go_min(a, b):
    return lib().orc_go_min(a, b)


def go_max(a, b):
    return lib().orc_go_max(a, b)


def go_f2i(x):
    return lib().orc_go_f2i(x)


def go_pow(x, y):
    return lib().orc_go_pow(x, y)


def go_u8(x):
    return lib().orc_go_u8(x)


def color_nrgba(c):
    a = darr(c, 4)
    out = (ctypes.c_uint8 * 4)()
    lib().orc_color_nrgba(_dptr(a), out)
    return tuple(out)


def nrgba_rgba(c):
    a = (ctypes.c_uint8 * 4)(*c)
    out = (ctypes.c_uint32 * 4)()
    lib().orc_nrgba_rgba(a, out)
    return tuple(out)


def blend(dst, color):
    d = (ctypes.c_uint8 * 4)(*dst)
    a = darr(color, 4)
    lib().orc_blend(d, _dptr(a))
    return tuple(d)


def barycentric(p1, p2, p3, p):
    out = np.zeros(4)
    lib().orc_barycentric(_dptr(darr(p1, 3)), _dptr(darr(p2, 3)), _dptr(darr(p3, 3)), _dptr(darr(p, 3)),
                          _dptr(out))
    return out


def clip_polygon(tri):
    inp = darr(tri, 12)
    out = np.zeros(36)
    n = lib().orc_clip_polygon(_dptr(inp), _dptr(out))
    return out.reshape(9, 4)[:n].copy()


def clip_line(w1, w2):
    out = np.zeros(8)
    ok = lib().orc_clip_line(_dptr(darr(w1, 4)), _dptr(darr(w2, 4)), _dptr(out))
    return out.reshape(2, 4).copy() if ok else None


def matrix_mul(a, b):
    out = np.zeros(16)
    lib().orc_matrix_mul(_dptr(darr(a, 16)), _dptr(darr(b, 16)), _dptr(out))
    return out


def mul_position_w(m, p):
    out = np.zeros(4)
    lib().orc_mul_position_w(_dptr(darr(m, 16)), _dptr(darr(p, 3)), _dptr(out))
    return out


def look_at(eye, center, up):
    out = np.zeros(16)
    lib().orc_look_at(_dptr(darr(eye, 3)), _dptr(darr(center, 3)), _dptr(darr(up, 3)), _dptr(out))
    return out


def frustum(l, r, b, t, n, f):
    out = np.zeros(16)
    lib().orc_frustum(l, r, b, t, n, f, _dptr(out))
    return out


def perspective(fovy, aspect, near, far):
    out = np.zeros(16)
    lib().orc_perspective(fovy, aspect, near, far, _dptr(out))
    return out


def screen(w, h):
    out = np.zeros(16)
    lib().orc_screen(w, h, _dptr(out))
    return out


def bilinear_sample(tex, u, v):
    out = np.zeros(4)
    lib().orc_bilinear_sample(ctypes.byref(tex.s), u, v, _dptr(out))
    return out


def phong_fragment(shader, pos, nrm, uv=(0, 0), col=(0, 0, 0, 0)):
    out = np.zeros(4)
    lib().orc_phong_fragment(ctypes.byref(shader), _dptr(darr(pos, 3)), _dptr(darr(nrm, 3)), _dptr(darr(uv, 2)),
                             _dptr(darr(col, 4)), _dptr(out))
    return out


def resolve(src, dw, dh, filter=1):
    """src: (sh, sw, 4) uint8 NRGBA -> (dh, dw, 4) uint8 (filter 1 tent, 2 box)."""
    src = np.ascontiguousarray(src, dtype=np.uint8)
    sh, sw = src.shape[:2]
    out = np.zeros((dh, dw, 4), dtype=np.uint8)
    lib().orc_resolve(src.ctypes.data_as(_P(ctypes.c_uint8)), sw, sh, out.ctypes.data_as(_P(ctypes.c_uint8)), dw, dh,
                      filter)
    return out


def tent_weights(src, dst):
    """(weights (dst, taps) int16, start (dst,) int32) of one axis."""
    taps = lib().orc_tent_weights(src, dst, None, None)
    w = np.zeros((dst, taps), dtype=np.int16)
    st = np.zeros(dst, dtype=np.int32)
    lib().orc_tent_weights(src, dst, w.ctypes.data_as(_P(ctypes.c_int16)), st.ctypes.data_as(_P(ctypes.c_int32)))
    return w, st


# ---------------------------------------------------------------------------
# mesh preparation (mesh.go); arrays are (n, nv, 3) float64, returned as copies


def _c(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64).copy()


def _p(a):
    return None if a is None else a.ctypes.data_as(_P(_d))


def mesh_bounding_box(tpos, lpos=None):
    """Mesh.BoundingBox -> (min (3,), max (3,))."""
    t, l = _c(tpos), _c(lpos)
    box = np.zeros(6)
    lib().orc_mesh_bounding_box(_p(t), 0 if t is None else t.shape[0], _p(l), 0 if l is None else l.shape[0], _p(box))
    return box[:3], box[3:]


def mesh_transform(pos, nrm, m):
    """Triangle/Line.Transform of every vertex -> (pos, nrm)."""
    P, Nn = _c(pos), _c(nrm)
    M = np.ascontiguousarray(m, dtype=np.float64)
    lib().orc_mesh_transform(_p(P), _p(Nn), P.size // 3, _p(M))
    return P, Nn


def mesh_fit_inside(tpos, tnrm, box_min, box_max, anchor, lpos=None, lnrm=None):
    """Mesh.FitInside -> (tpos, tnrm, lpos, lnrm, matrix (16,))."""
    tp, tn, lp, ln = _c(tpos), _c(tnrm), _c(lpos), _c(lnrm)
    box = np.array(list(box_min) + list(box_max), dtype=np.float64)
    a = np.array(anchor, dtype=np.float64)
    m = np.zeros(16)
    lib().orc_mesh_fit_inside(_p(tp), _p(tn), 0 if tp is None else tp.shape[0], _p(lp), _p(ln),
                              0 if lp is None else lp.shape[0], _p(box), _p(a), _p(m))
    return tp, tn, lp, ln, m


def mesh_smooth_normals(tpos, tnrm, threshold=None):
    """Mesh.SmoothNormals (threshold None) or SmoothNormalsThreshold with
    threshold = cos(radians) -> new normals."""
    tp, tn = _c(tpos), _c(tnrm)
    if threshold is None:
        lib().orc_mesh_smooth_normals(_p(tp), _p(tn), tp.shape[0])
    else:
        lib().orc_mesh_smooth_normals_threshold(_p(tp), _p(tn), tp.shape[0], float(threshold))
    return tn


def load_stl(data):
    """LoadSTL from the file's bytes -> (pos (n,3,3), nrm (n,3,3)); raises on
    the inputs the reference rejects or panics on."""
    buf = np.frombuffer(bytes(data), dtype=np.uint8)
    n = ctypes.c_int64(0)
    bp = buf.ctypes.data_as(_P(ctypes.c_uint8))
    if lib().orc_load_stl(bp, buf.size, ctypes.byref(n), None, None):
        raise ValueError("LoadSTL fails on this input")
    pos = np.zeros((n.value, 3, 3))
    nrm = np.zeros((n.value, 3, 3))
    lib().orc_load_stl(bp, buf.size, ctypes.byref(n), _p(pos), _p(nrm))
    return pos, nrm


def load_obj(text):
    """LoadOBJ from the file's text -> (pos, nrm (n,3,3), uv (n,3,2))."""
    raw = text.encode() if isinstance(text, str) else bytes(text)
    n = ctypes.c_int64(0)
    if lib().orc_load_obj(raw, len(raw), ctypes.byref(n), None, None, None):
        raise ValueError("LoadOBJ fails on this input")
    pos = np.zeros((n.value, 3, 3))
    nrm = np.zeros((n.value, 3, 3))
    uv = np.zeros((n.value, 3, 2))
    lib().orc_load_obj(raw, len(raw), ctypes.byref(n), _p(pos), _p(nrm), _p(uv))
    return pos, nrm, uv


def png_filter(nrgba):
    """(bpp, filtered stream) of Go image/png's row filtering for an (h, w, 4)
    NRGBA image."""
    img = np.ascontiguousarray(nrgba, dtype=np.uint8)
    h, w = img.shape[:2]
    ip = img.ctypes.data_as(_P(ctypes.c_uint8))
    bpp = lib().orc_png_filter(ip, w, h, None)
    out = np.zeros(h * (w * bpp + 1), dtype=np.uint8)
    lib().orc_png_filter(ip, w, h, out.ctypes.data_as(_P(ctypes.c_uint8)))
    return bpp, out
