"""fauxgl's Go API for the draw path, over the gfx950 C ABI.

Same names and argument meaning as the reference (context.go, shader.go,
texture.go, matrix.go, color.go, mesh.go): NewContext, Context.DrawMesh /
DrawTriangles / DrawLines, ClearColorBuffer[With] / ClearDepthBuffer[With],
the SolidColor / Texture / Phong shaders, Image / DepthImage.  Draw calls
return RasterizeInfo(TotalPixels, UpdatedPixels) like the reference.

Meshes are triangle / line soups held as numpy arrays in the C ABI's SoA
layout (fgl_prims): position (n, nv, 3), normal (n, nv, 3), texture
(n, nv, 2), color (n, nv, 4) with nv = 3 for triangles, 2 for lines.

Every draw runs on the GPU; a shader that is not one of the three built-ins
raises (the reference's Go CPU path is what serves custom shaders).
"""
import ctypes
import os
import math
from collections import namedtuple

import numpy as np

from . import _native as N

# ---------------------------------------------------------------------------
# enums (context.go:11-26)

FaceCW, FaceCCW = N.FGL_FACE_CW, N.FGL_FACE_CCW
CullNone, CullFront, CullBack = N.FGL_CULL_NONE, N.FGL_CULL_FRONT, N.FGL_CULL_BACK

RasterizeInfo = namedtuple("RasterizeInfo", ["TotalPixels", "UpdatedPixels"])

MaxFloat64 = 1.7976931348623157e308

# ---------------------------------------------------------------------------
# Vector (vector.go)


class Vector(namedtuple("Vector", ["X", "Y", "Z"])):
    __slots__ = ()

    def Add(self, b):
        return Vector(self.X + b.X, self.Y + b.Y, self.Z + b.Z)

    def Sub(self, b):
        return Vector(self.X - b.X, self.Y - b.Y, self.Z - b.Z)

    def Mul(self, b):
        return Vector(self.X * b.X, self.Y * b.Y, self.Z * b.Z)

    def Div(self, b):
        return Vector(self.X / b.X, self.Y / b.Y, self.Z / b.Z)

    def MulScalar(self, s):
        return Vector(self.X * s, self.Y * s, self.Z * s)

    def DivScalar(self, s):
        return Vector(self.X / s, self.Y / s, self.Z / s)

    def Dot(self, b):
        return self.X * b.X + self.Y * b.Y + self.Z * b.Z

    def Cross(self, b):
        return Vector(self.Y * b.Z - self.Z * b.Y, self.Z * b.X - self.X * b.Z, self.X * b.Y - self.Y * b.X)

    def Length(self):
        return math.sqrt(self.X * self.X + self.Y * self.Y + self.Z * self.Z)

    def Normalize(self):
        r = 1 / math.sqrt(self.X * self.X + self.Y * self.Y + self.Z * self.Z)
        return Vector(self.X * r, self.Y * r, self.Z * r)

    def Negate(self):
        return Vector(-self.X, -self.Y, -self.Z)

    def MinComponent(self):
        return go_min(go_min(self.X, self.Y), self.Z)


def V(x, y, z):
    return Vector(float(x), float(y), float(z))


def go_min(x, y):
    """math.Min"""
    if (math.isinf(x) and x < 0) or (math.isinf(y) and y < 0):
        return -math.inf
    if math.isnan(x) or math.isnan(y):
        return math.nan
    if x == 0 and x == y:
        return x if math.copysign(1, x) < 0 else y
    return x if x < y else y


def go_max(x, y):
    """math.Max"""
    if (math.isinf(x) and x > 0) or (math.isinf(y) and y > 0):
        return math.inf
    if math.isnan(x) or math.isnan(y):
        return math.nan
    if x == 0 and x == y:
        return y if math.copysign(1, x) < 0 else x
    return x if x > y else y


def Radians(degrees):
    return degrees * math.pi / 180


# ---------------------------------------------------------------------------
# Color (color.go)


class Color(namedtuple("Color", ["R", "G", "B", "A"])):
    __slots__ = ()

    def NRGBA(self):
        """Color.NRGBA (color.go:56-63): clamp, *255, truncating uint8()."""
        return tuple(_go_u8(_clamp(c, 0.0, 1.0) * 255.0) for c in self)

    def Alpha(self, a):
        return Color(self.R, self.G, self.B, float(a))

    def Opaque(self):
        return Color(self.R, self.G, self.B, 1.0)


def _clamp(x, lo, hi):
    if x < lo:
        return lo
    if x > hi:
        return hi
    return x


def _go_u8(x):
    # uint8(x) converts through int32 on amd64; NaN / out of range -> 0x80000000
    if not (-2147483649.0 < x < 2147483648.0):
        return 0
    return int(x) & 0xFF


Discard = Color(0.0, 0.0, 0.0, 0.0)
Transparent = Color(0.0, 0.0, 0.0, 0.0)
Black = Color(0.0, 0.0, 0.0, 1.0)
White = Color(1.0, 1.0, 1.0, 1.0)


def Gray(x):
    return Color(float(x), float(x), float(x), 1.0)


def _scan_hex(x, width, count):
    """fmt.Sscanf(x, "%<width>x" * count): each field takes up to `width` hex
    digits; the first field without a digit ends the scan, leaving it and
    the rest unset (None), as the error Sscanf returns does."""
    out, pos = [], 0
    for _ in range(count):
        j = pos
        while j < len(x) and j - pos < width and x[j] in "0123456789abcdefABCDEF":
            j += 1
        if j == pos:
            break
        out.append(int(x[pos:j], 16))
        pos = j
    return out + [None] * (count - len(out))


def HexColor(x):
    """HexColor (color.go:31-54): strings.Trim of '#', then by length (in
    bytes) 3 or 4 one-digit fields each widened to (v << 4) | v, or 6 or 8
    two-digit fields; any other length leaves R = G = B = 0, A = 255."""
    x = x.strip("#")
    n = len(x.encode("utf-8"))
    r = g = b = 0
    a = 255
    if n in (3, 4):
        f = _scan_hex(x, 1, n)
        f = [None if v is None else (v << 4) | v for v in f]
    elif n in (6, 8):
        f = _scan_hex(x, 2, n // 2)
    else:
        f = []
    if f:
        r, g, b = (v if v is not None else 0 for v in f[:3])
        if len(f) == 4 and f[3] is not None:
            a = f[3]
    d = 255.0
    return Color(r / d, g / d, b / d, a / d)


# ---------------------------------------------------------------------------
# Matrix (matrix.go)


class Matrix(tuple):
    """16 doubles X00..X33, row-major."""

    def __new__(cls, vals):
        vals = tuple(float(v) for v in vals)
        assert len(vals) == 16
        return tuple.__new__(cls, vals)

    def Mul(self, b):
        """Matrix.Mul (matrix.go:188-207): ((a0*b0 + a1*b1) + a2*b2) + a3*b3."""
        a = self
        out = [0.0] * 16
        for r in range(4):
            for c in range(4):
                out[r * 4 + c] = a[r * 4] * b[c] + a[r * 4 + 1] * b[4 + c] + a[r * 4 + 2] * b[8 + c] + \
                    a[r * 4 + 3] * b[12 + c]
        return Matrix(out)

    def MulPosition(self, v):
        a = self
        return Vector(a[0] * v.X + a[1] * v.Y + a[2] * v.Z + a[3], a[4] * v.X + a[5] * v.Y + a[6] * v.Z + a[7],
                      a[8] * v.X + a[9] * v.Y + a[10] * v.Z + a[11])

    def MulDirection(self, v):
        a = self
        return Vector(a[0] * v.X + a[1] * v.Y + a[2] * v.Z, a[4] * v.X + a[5] * v.Y + a[6] * v.Z,
                      a[8] * v.X + a[9] * v.Y + a[10] * v.Z).Normalize()

    def Translate(self, v):
        return Translate(v).Mul(self)

    def Scale(self, v):
        return Scale(v).Mul(self)

    def Frustum(self, l, r, b, t, n, f):
        return Frustum(l, r, b, t, n, f).Mul(self)

    def Orthographic(self, l, r, b, t, n, f):
        return Orthographic(l, r, b, t, n, f).Mul(self)

    def Perspective(self, fovy, aspect, near, far):
        return Perspective(fovy, aspect, near, far).Mul(self)

    def LookAt(self, eye, center, up):
        return LookAt(eye, center, up).Mul(self)


def Identity():
    return Matrix([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1])


def Translate(v):
    return Matrix([1, 0, 0, v.X, 0, 1, 0, v.Y, 0, 0, 1, v.Z, 0, 0, 0, 1])


def Scale(v):
    return Matrix([v.X, 0, 0, 0, 0, v.Y, 0, 0, 0, 0, v.Z, 0, 0, 0, 0, 1])


def Frustum(l, r, b, t, n, f):
    t1 = 2 * n
    t2 = r - l
    t3 = t - b
    t4 = f - n
    return Matrix([t1 / t2, 0, (r + l) / t2, 0, 0, t1 / t3, (t + b) / t3, 0, 0, 0, (-f - n) / t4, (-t1 * f) / t4,
                   0, 0, -1, 0])


def Orthographic(l, r, b, t, n, f):
    return Matrix([2 / (r - l), 0, 0, -(r + l) / (r - l), 0, 2 / (t - b), 0, -(t + b) / (t - b), 0, 0,
                   -2 / (f - n), -(f + n) / (f - n), 0, 0, 0, 1])


def Perspective(fovy, aspect, near, far):
    # math.tan here is libm's; Go's math.Tan may differ in the last bit.
    ymax = near * math.tan(fovy * math.pi / 360)
    xmax = ymax * aspect
    return Frustum(-xmax, xmax, -ymax, ymax, near, far)


def LookAt(eye, center, up):
    z = eye.Sub(center).Normalize()
    x = up.Cross(z).Normalize()
    y = z.Cross(x)
    return Matrix([x.X, x.Y, x.Z, -x.Dot(eye), y.X, y.Y, y.Z, -y.Dot(eye), z.X, z.Y, z.Z, -z.Dot(eye), 0, 0, 0, 1])


def Screen(w, h):
    w2 = float(w) / 2
    h2 = float(h) / 2
    return Matrix([w2, 0, 0, w2, 0, -h2, 0, h2, 0, 0, 0.5, 0.5, 0, 0, 0, 1])


# ---------------------------------------------------------------------------
# Textures (texture.go)


class ImageTexture:
    """An image texture.  texels: (H, W, 4) uint8 (kind "NRGBA" for an
    image.NRGBA, "RGBA" for a premultiplied image.RGBA) or (H, W, 4) uint16
    holding At(x, y).RGBA() of any Go image ("RGBA16")."""

    _kinds = {"NRGBA": N.FGL_TEX_NRGBA8, "RGBA": N.FGL_TEX_RGBA8, "RGBA16": N.FGL_TEX_RGBA16}

    def __init__(self, texels, kind="NRGBA"):
        if kind not in self._kinds:
            raise ValueError(f"unknown texture kind {kind}")
        dtype = np.uint16 if kind == "RGBA16" else np.uint8
        self.Image = np.ascontiguousarray(np.asarray(texels, dtype=dtype))
        if self.Image.ndim != 3 or self.Image.shape[2] != 4:
            raise ValueError("texels must be (H, W, 4)")
        self.Height, self.Width = self.Image.shape[:2]
        self.kind = kind
        self._handles = {}

    def handle(self, ctx):
        ptr = ctx._ptr()
        key = ptr.value
        h = self._handles.get(key)
        if h is None:
            out = ctypes.c_int64(0)
            N.check(N.load().fgl_texture_create(ptr, self.Width, self.Height, self._kinds[self.kind],
                                                self.Image.ctypes.data_as(ctypes.c_void_p), ctypes.byref(out)), ptr)
            h = out.value
            self._handles[key] = h
            ctx._textures.append(self)
        return h


def NewImageTexture(texels, kind="NRGBA"):
    return ImageTexture(texels, kind)


# ---------------------------------------------------------------------------
# Shaders (shader.go)


class SolidColorShader:
    def __init__(self, matrix, color):
        self.Matrix = Matrix(matrix)
        self.Color = Color(*color)


class TextureShader:
    def __init__(self, matrix, texture):
        self.Matrix = Matrix(matrix)
        self.Texture = texture


class PhongShader:
    def __init__(self, matrix, light_direction, camera_position, object_color=Discard,
                 ambient=Color(0.2, 0.2, 0.2, 1.0), diffuse=Color(0.8, 0.8, 0.8, 1.0),
                 specular=Color(1.0, 1.0, 1.0, 1.0), texture=None, specular_power=32.0):
        self.Matrix = Matrix(matrix)
        self.LightDirection = Vector(*light_direction)
        self.CameraPosition = Vector(*camera_position)
        self.ObjectColor = Color(*object_color)
        self.AmbientColor = Color(*ambient)
        self.DiffuseColor = Color(*diffuse)
        self.SpecularColor = Color(*specular)
        self.Texture = texture
        self.SpecularPower = float(specular_power)


def NewSolidColorShader(matrix, color):
    return SolidColorShader(matrix, color)


def NewTextureShader(matrix, texture):
    return TextureShader(matrix, texture)


def NewPhongShader(matrix, light_direction, camera_position):
    """NewPhongShader (shader.go:61-68) defaults."""
    return PhongShader(matrix, light_direction, camera_position)


def shader_struct(shader, ctx=None):
    """The fgl_shader of a built-in shader."""
    s = N.FglShader()
    s.texture = -1
    if isinstance(shader, SolidColorShader):
        s.kind = N.FGL_SHADER_SOLID
        s.color[:] = list(shader.Color)
    elif isinstance(shader, TextureShader):
        s.kind = N.FGL_SHADER_TEXTURE
        if shader.Texture is None:
            raise ValueError("TextureShader without a texture")
        s.texture = shader.Texture.handle(ctx) if ctx is not None else -1
    elif isinstance(shader, PhongShader):
        s.kind = N.FGL_SHADER_PHONG
        s.light_dir[:] = list(shader.LightDirection)
        s.camera_pos[:] = list(shader.CameraPosition)
        s.object_color[:] = list(shader.ObjectColor)
        s.ambient[:] = list(shader.AmbientColor)
        s.diffuse[:] = list(shader.DiffuseColor)
        s.specular[:] = list(shader.SpecularColor)
        s.specular_power = shader.SpecularPower
        if shader.Texture is not None:
            s.texture = shader.Texture.handle(ctx) if ctx is not None else -1
    else:
        raise NotImplementedError(
            f"{type(shader).__name__} is not a built-in shader; custom shaders run on the reference's CPU path")
    s.matrix[:] = list(shader.Matrix)
    return s


# ---------------------------------------------------------------------------
# Meshes (mesh.go:9-29)


class Prims:
    """A triangle (nv=3) or line (nv=2) soup in SoA form."""

    def __init__(self, position, normal=None, texture=None, color=None, nv=3):
        self.nv = nv
        self.position = _soa(position, nv, 3)
        n = self.position.shape[0]
        self.normal = None if normal is None else _soa(normal, nv, 3, n)
        self.texture = None if texture is None else _soa(texture, nv, 2, n)
        self.color = None if color is None else _soa(color, nv, 4, n)

    def __len__(self):
        return self.position.shape[0]

    def struct(self):
        def p(a):
            return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        return N.FglPrims(len(self), p(self.position), p(self.normal), p(self.texture), p(self.color))

    def arrays(self):
        return {"position": self.position, "normal": self.normal, "texture": self.texture, "color": self.color}


def _soa(a, nv, k, n=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if a.ndim == 2:
        a = a.reshape(-1, nv, k)
    if a.ndim != 3 or a.shape[1:] != (nv, k) or (n is not None and a.shape[0] != n):
        raise ValueError(f"expected ({n if n is not None else 'n'}, {nv}, {k}) array, got {a.shape}")
    return a


class Mesh:
    """Mesh{Triangles, Lines} with SoA soups, held on the host.

    Context.DrawMesh uploads a host mesh's soups at every draw (the
    reference reads the caller's triangles at every draw too), unless the
    context was asked to keep it resident (Context.KeepResident).  A
    resident copy is tagged with the mesh's generation, which dirty()
    advances (mesh.go:31; Add, Transform and the other in-place methods
    call it), and is uploaded again when the generation moved on.  Code
    that edits the soups' arrays directly calls dirty() itself."""

    def __init__(self, triangles=None, lines=None):
        self.Triangles = triangles
        self.Lines = lines
        self.generation = 0

    def dirty(self):
        """Mesh.dirty (mesh.go:31-33): the mesh changed in place."""
        self.generation += 1

    def Add(self, other):
        """Mesh.Add (mesh.go:50-53): appends other's triangles and lines."""
        def cat(a, b):
            if a is None:
                return b
            if b is None:
                return a
            parts = {}
            for k in ("position", "normal", "texture", "color"):
                x, y = getattr(a, k), getattr(b, k)
                if x is None and y is None:
                    parts[k] = None
                    continue
                if x is None:
                    x = np.zeros((len(a),) + y.shape[1:])
                if y is None:
                    y = np.zeros((len(b),) + x.shape[1:])
                parts[k] = np.concatenate([x, y])
            return Prims(parts["position"], parts["normal"], parts["texture"], parts["color"], nv=a.nv)
        self.Triangles = cat(self.Triangles, other.Triangles)
        self.Lines = cat(self.Lines, other.Lines)
        self.dirty()

    def Transform(self, matrix):
        """Mesh.Transform (mesh.go:167-175): Triangle / Line.Transform of
        every vertex: Position by MulPosition, Normal by MulDirection
        (matrix.go:209-229), in Go's term order."""
        m = Matrix(matrix)
        for p in (self.Triangles, self.Lines):
            if p is None:
                continue
            P = p.position
            x, y, z = P[..., 0].copy(), P[..., 1].copy(), P[..., 2].copy()
            for r in range(3):
                P[..., r] = ((m[4 * r] * x + m[4 * r + 1] * y) + m[4 * r + 2] * z) + m[4 * r + 3]
            if p.normal is None:
                p.normal = np.zeros_like(P)
            Nn = p.normal
            x, y, z = Nn[..., 0].copy(), Nn[..., 1].copy(), Nn[..., 2].copy()
            for r in range(3):
                Nn[..., r] = (m[4 * r] * x + m[4 * r + 1] * y) + m[4 * r + 2] * z
            with np.errstate(divide="ignore", invalid="ignore"):
                rr = 1 / np.sqrt((Nn[..., 0] * Nn[..., 0] + Nn[..., 1] * Nn[..., 1]) + Nn[..., 2] * Nn[..., 2])
                Nn *= rr[..., None]
        self.dirty()


def NewTriangleMesh(triangles):
    return Mesh(triangles, None)


def NewLineMesh(lines):
    return Mesh(None, lines)


class DeviceMesh:
    """A Mesh uploaded once to a context's device (fgl_mesh)."""

    def __init__(self, ctx, mesh, handle=None):
        self.ctx = ctx
        if handle is not None:
            self.h = handle
            return
        t = mesh.Triangles.struct() if mesh.Triangles is not None else None
        l = mesh.Lines.struct() if mesh.Lines is not None else None
        h = ctypes.c_void_p()
        N.check(N.load().fgl_mesh_create(ctx._ptr(), ctypes.byref(t) if t else None,
                                         ctypes.byref(l) if l else None, ctypes.byref(h)), ctx._ptr())
        self.h = h

    def close(self):
        if self.h:
            N.load().fgl_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ptr(self):
        if not self.h:
            raise RuntimeError("mesh is closed")
        return self.h

    # mesh preparation on the device (mesh.go), in place
    def BoundingBox(self):
        """Mesh.BoundingBox (mesh.go:153-165) -> (min Vector, max Vector)."""
        b = (ctypes.c_double * 6)()
        N.check(N.load().fgl_mesh_bounding_box(self._ptr(), b), self.ctx._ptr())
        return Vector(b[0], b[1], b[2]), Vector(b[3], b[4], b[5])

    def Transform(self, matrix):
        """Mesh.Transform (mesh.go:167-175)."""
        m = (ctypes.c_double * 16)(*Matrix(matrix))
        N.check(N.load().fgl_mesh_transform(self._ptr(), m), self.ctx._ptr())

    def FitInside(self, box_min, box_max, anchor):
        """Mesh.FitInside (mesh.go:142-151); returns the matrix applied."""
        b = (ctypes.c_double * 6)(*box_min, *box_max)
        a = (ctypes.c_double * 3)(*anchor)
        m = (ctypes.c_double * 16)()
        N.check(N.load().fgl_mesh_fit_inside(self._ptr(), b, a, m), self.ctx._ptr())
        return Matrix(list(m))

    def BiUnitCube(self):
        """Mesh.BiUnitCube (mesh.go:132-135)."""
        return self.FitInside(V(-1, -1, -1), V(1, 1, 1), V(0.5, 0.5, 0.5))

    def UnitCube(self):
        """Mesh.UnitCube (mesh.go:127-130)."""
        return self.FitInside(V(-0.5, -0.5, -0.5), V(0.5, 0.5, 0.5), V(0.5, 0.5, 0.5))

    def MoveTo(self, position, anchor):
        """Mesh.MoveTo (mesh.go:137-141): Translate(position - box.Anchor(anchor))."""
        lo, hi = self.BoundingBox()
        a = lo.Add(hi.Sub(lo).Mul(Vector(*anchor)))
        m = Translate(Vector(*position).Sub(a))
        self.Transform(m)
        return m

    def Center(self):
        """Mesh.Center (mesh.go:143-145)."""
        return self.MoveTo(V(0, 0, 0), V(0.5, 0.5, 0.5))

    def SmoothNormals(self):
        """Mesh.SmoothNormals (mesh.go:105-125)."""
        N.check(N.load().fgl_mesh_smooth_normals(self._ptr()), self.ctx._ptr())

    def SmoothNormalsThreshold(self, radians):
        """Mesh.SmoothNormalsThreshold (mesh.go:91-103).  The threshold is
        cos(radians) from Python's math.cos (the C library's); a Go caller
        passes math.Cos(radians) to fgl_mesh_smooth_normals_threshold."""
        self.SmoothNormalsThresholdCos(math.cos(radians))

    def SmoothNormalsThresholdCos(self, threshold):
        N.check(N.load().fgl_mesh_smooth_normals_threshold(self._ptr(), float(threshold)), self.ctx._ptr())

    def Download(self):
        """The mesh's current vertex data as a host Mesh (fgl_prims layout)."""
        nt, nl = ctypes.c_int64(), ctypes.c_int64()
        L = N.load()
        N.check(L.fgl_mesh_counts(self._ptr(), ctypes.byref(nt), ctypes.byref(nl)), self.ctx._ptr())

        def attrs(lines, n, nv):
            out = []
            for attr, w in enumerate((3, 3, 2, 4)):
                a = np.empty((n, nv, w), dtype=np.float64)
                N.check(L.fgl_mesh_read(self._ptr(), lines, attr, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                        a.size), self.ctx._ptr())
                out.append(a)
            return Prims(*out, nv=nv)

        tris = attrs(0, nt.value, 3) if nt.value else None
        lines = attrs(1, nl.value, 2) if nl.value else None
        return Mesh(tris, lines)


class Batch:
    """Many views rendered per step (fgl_batch): each view is a context, its
    clear color, its current state and shader, and a resident mesh.  The
    first Run draws them one by one (sizing every context's buffers) and
    prepares the views' draws for replay; later Runs launch each pipeline
    stage once per group of views with one tile shape and kernel variant.
    Run returns [RasterizeInfo] per view."""

    def __init__(self):
        h = ctypes.c_void_p()
        N.check(N.load().fgl_batch_create(ctypes.byref(h)))
        self.h = h
        self.n = 0
        self._keep = []

    def _check(self, rc):
        if rc != N.FGL_OK:
            msg = N.load().fgl_batch_error(self.h)
            raise N.FglError(rc, msg.decode() if msg else "")

    def Add(self, ctx, mesh, clear_color=None, clear_depth=MaxFloat64):
        """A view: ClearColorBufferWith(clear_color or ctx.ClearColor),
        ClearDepthBufferWith(clear_depth), DrawMesh(mesh) with ctx's
        current state and shader."""
        c = (ctypes.c_uint8 * 4)(*Color(*(clear_color if clear_color is not None else ctx.ClearColor)).NRGBA())
        st = ctx._state()
        sh = shader_struct(ctx.Shader, ctx)
        self._check(N.load().fgl_batch_add(self.h, ctx._ptr(), c, float(clear_depth), ctypes.byref(st),
                                           ctypes.byref(sh), mesh._ptr()))
        self._keep.append((ctx, mesh))
        self.n += 1

    def Run(self):
        info = (ctypes.c_uint64 * (2 * max(self.n, 1)))()
        self._check(N.load().fgl_batch_run(self.h, info))
        return [RasterizeInfo(int(info[2 * i]), int(info[2 * i + 1])) for i in range(self.n)]

    def close(self):
        if getattr(self, "h", None):
            N.load().fgl_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _file_bytes(src):
    if isinstance(src, (bytes, bytearray, memoryview)):
        return bytes(src)
    with open(src, "rb") as f:
        return f.read()


def LoadSTL(ctx, src):
    """LoadSTL (stl.go:23-57) into a mesh resident on ctx's device; src is a
    path or the file's bytes."""
    data = _file_bytes(src)
    buf = (ctypes.c_uint8 * max(len(data), 1)).from_buffer_copy(data or b"\0")
    h = ctypes.c_void_p()
    N.check(N.load().fgl_mesh_load_stl(ctx._ptr(), buf, len(data), ctypes.byref(h)), ctx._ptr())
    return DeviceMesh(ctx, None, handle=h)


def LoadOBJ(ctx, src):
    """LoadOBJ (obj.go:19-79) into a mesh resident on ctx's device; src is a
    path or the file's bytes / text."""
    data = src.encode() if isinstance(src, str) and "\n" in src else _file_bytes(src)
    h = ctypes.c_void_p()
    N.check(N.load().fgl_mesh_load_obj(ctx._ptr(), data, len(data), ctypes.byref(h)), ctx._ptr())
    return DeviceMesh(ctx, None, handle=h)


# ---------------------------------------------------------------------------
# Context (context.go)


class Context:
    """NewContext (context.go:60-81): a W x H framebuffer on the GPU plus the
    pipeline state fields read at draw time."""

    def __init__(self, width, height):
        L = N.load()
        h = ctypes.c_void_p()
        rc = L.fgl_ctx_create(int(width), int(height), ctypes.byref(h))
        N.check(rc, None)
        self.h = h
        self.Width = int(width)
        self.Height = int(height)
        self.ClearColor = Transparent
        self.Shader = NewSolidColorShader(Identity(), Color(1.0, 0.0, 1.0, 1.0))
        self.ReadDepth = True
        self.WriteDepth = True
        self.WriteColor = True
        self.AlphaBlend = True
        self.Wireframe = False
        self.FrontFace = FaceCCW
        self.Cull = CullBack
        self.LineWidth = 2.0
        self.DepthBias = 0.0
        self._textures = []
        self._resident = {}  # id(host Mesh) -> [mesh, generation uploaded, DeviceMesh]

    def _ptr(self):
        if not self.h:
            raise RuntimeError("context is closed")
        return self.h

    def close(self):
        if getattr(self, "h", None):
            # the resident copies refer to this context (and it to them), so
            # the collector may reach either first: free them before it
            for ent in getattr(self, "_resident", {}).values():
                if ent[2] is not None:
                    ent[2].close()
            self._resident = {}
            N.load().fgl_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _state(self):
        return N.FglState(int(bool(self.ReadDepth)), int(bool(self.WriteDepth)), int(bool(self.WriteColor)),
                          int(bool(self.AlphaBlend)), int(bool(self.Wireframe)), (ctypes.c_uint8 * 3)(),
                          int(self.FrontFace), int(self.Cull), float(self.LineWidth), float(self.DepthBias))

    # clears (context.go:119-145)
    def ClearColorBufferWith(self, color):
        c = (ctypes.c_uint8 * 4)(*Color(*color).NRGBA())
        N.check(N.load().fgl_clear_color(self._ptr(), c), self._ptr())

    def ClearColorBuffer(self):
        self.ClearColorBufferWith(self.ClearColor)

    def ClearDepthBufferWith(self, value):
        N.check(N.load().fgl_clear_depth(self._ptr(), float(value)), self._ptr())

    def ClearDepthBuffer(self):
        self.ClearDepthBufferWith(MaxFloat64)

    # drawing (context.go:391-439)
    def DrawTriangles(self, triangles):
        info = (ctypes.c_uint64 * 2)()
        st = self._state()
        sh = shader_struct(self.Shader, self)
        p = triangles.struct()
        N.check(N.load().fgl_draw_triangles(self._ptr(), ctypes.byref(st), ctypes.byref(sh), ctypes.byref(p), info),
                self._ptr())
        return RasterizeInfo(int(info[0]), int(info[1]))

    def DrawLines(self, lines):
        info = (ctypes.c_uint64 * 2)()
        st = self._state()
        sh = shader_struct(self.Shader, self)
        p = lines.struct()
        N.check(N.load().fgl_draw_lines(self._ptr(), ctypes.byref(st), ctypes.byref(sh), ctypes.byref(p), info),
                self._ptr())
        return RasterizeInfo(int(info[0]), int(info[1]))

    def DrawMesh(self, mesh):
        if isinstance(mesh, DeviceMesh):
            info = (ctypes.c_uint64 * 2)()
            st = self._state()
            sh = shader_struct(self.Shader, self)
            N.check(N.load().fgl_draw_mesh(self._ptr(), ctypes.byref(st), ctypes.byref(sh), mesh.h, info),
                    self._ptr())
            return RasterizeInfo(int(info[0]), int(info[1]))
        ent = self._resident.get(id(mesh))
        if ent is not None and ent[0] is mesh:
            if ent[1] != mesh.generation or ent[2] is None:
                # the mesh changed in place since its upload (Mesh.dirty)
                if ent[2] is not None:
                    ent[2].close()
                ent[2] = DeviceMesh(self, mesh)
                ent[1] = mesh.generation
            return self.DrawMesh(ent[2])
        i1 = self.DrawTriangles(mesh.Triangles) if mesh.Triangles is not None else RasterizeInfo(0, 0)
        i2 = self.DrawLines(mesh.Lines) if mesh.Lines is not None else RasterizeInfo(0, 0)
        return RasterizeInfo(i1[0] + i2[0], i1[1] + i2[1])

    def KeepResident(self, mesh, keep=True):
        """Opts a host Mesh into residency on this context's device: DrawMesh
        then draws an uploaded copy, uploaded again whenever the mesh's
        generation (Mesh.dirty) moved on since.  keep=False frees the copy;
        DrawMesh goes back to uploading the soups at every draw."""
        ent = self._resident.pop(id(mesh), None)
        if ent is not None and ent[2] is not None:
            ent[2].close()
        if keep:
            self._resident[id(mesh)] = [mesh, None, None]

    def Upload(self, mesh):
        """Keeps mesh resident on this context's device for repeated DrawMesh."""
        return DeviceMesh(self, mesh)

    # tile shape and the autotuner (no reference counterpart: the images do
    # not depend on the shape, only the speed does)
    def SetTileShape(self, width=0, height=0):
        """Fixes this context's tile shape (one of TileShapes()); (0, 0)
        returns to the tuning table / the 32 x 32 default."""
        N.check(N.load().fgl_ctx_set_tile(self._ptr(), int(width), int(height)), self._ptr())

    def TileShape(self, ntri):
        """The (w, h) a draw of ntri triangles on this context would use."""
        out = (ctypes.c_int32 * 2)()
        N.check(N.load().fgl_ctx_get_tile(self._ptr(), int(ntri), out), self._ptr())
        return int(out[0]), int(out[1])

    def Autotune(self, mesh, reps=3):
        """Times DrawMesh(mesh) with the current state and shader for every
        tile shape, records the fastest (and the capacities its draw needed)
        in the process-wide tuning table and returns it as (w, h).  The
        framebuffer is left as it was."""
        if not isinstance(mesh, DeviceMesh):
            mesh = DeviceMesh(self, mesh)
        out = (ctypes.c_int32 * 2)()
        st = self._state()
        sh = shader_struct(self.Shader, self)
        N.check(N.load().fgl_autotune_mesh(self._ptr(), ctypes.byref(st), ctypes.byref(sh), mesh.h, int(reps), out),
                self._ptr())
        return int(out[0]), int(out[1])

    # readback (context.go:83-117)
    def Image(self):
        out = np.empty((self.Height, self.Width, 4), dtype=np.uint8)
        N.check(N.load().fgl_read_color(self._ptr(), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))),
                self._ptr())
        return out

    @property
    def ColorBuffer(self):
        return self.Image()

    @property
    def DepthBuffer(self):
        out = np.empty((self.Height, self.Width), dtype=np.float64)
        N.check(N.load().fgl_read_depth(self._ptr(), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                self._ptr())
        return out

    def SetColorBuffer(self, nrgba):
        a = np.ascontiguousarray(np.asarray(nrgba, dtype=np.uint8).reshape(self.Height, self.Width, 4))
        N.check(N.load().fgl_write_color(self._ptr(), a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))),
                self._ptr())

    def SetDepthBuffer(self, depth):
        a = np.ascontiguousarray(np.asarray(depth, dtype=np.float64).reshape(self.Height, self.Width))
        N.check(N.load().fgl_write_depth(self._ptr(), a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                self._ptr())

    def DepthImage(self):
        out = np.empty((self.Height, self.Width), dtype=np.uint16)
        N.check(N.load().fgl_depth_image(self._ptr(), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))),
                self._ptr())
        return out

    def EncodePNG(self):
        """The color buffer as PNG file bytes (SavePNG's encoding, util.go:56-63)."""
        L = N.load()
        bound = ctypes.c_int64()
        N.check(L.fgl_encode_png_bound(self._ptr(), ctypes.byref(bound)), self._ptr())
        buf = (ctypes.c_uint8 * bound.value)()
        n = ctypes.c_int64()
        N.check(L.fgl_encode_png(self._ptr(), buf, bound.value, ctypes.byref(n)), self._ptr())
        return bytes(buf[:n.value])

    def SavePNG(self, path):
        """SavePNG(path, Image()) (util.go:56-63)."""
        data = self.EncodePNG()
        with open(path, "wb") as f:
            f.write(data)

    def Resolve(self, dst, filter="tent"):
        """Downsamples this context's color into dst (an SSAA resolve)."""
        f = {"tent": N.FGL_RESOLVE_TENT, "box": N.FGL_RESOLVE_BOX}[filter]
        N.check(N.load().fgl_resolve(self._ptr(), dst._ptr(), f), self._ptr())

    # measurement
    def ProfileEnable(self, on=True):
        N.check(N.load().fgl_profile_enable(self._ptr(), 1 if on else 0), self._ptr())

    def ProfileOnly(self, kernel=None):
        """Times only launches of `kernel` (None: every launch)."""
        N.check(N.load().fgl_profile_only(self._ptr(), kernel.encode() if kernel else None), self._ptr())

    def ProfileReset(self):
        N.check(N.load().fgl_profile_reset(self._ptr()), self._ptr())

    def ProfileRead(self):
        cap = 64
        arr = (N.FglKernelTime * cap)()
        n = ctypes.c_int32(0)
        N.check(N.load().fgl_profile_read(self._ptr(), arr, cap, ctypes.byref(n)), self._ptr())
        return {arr[i].name.decode(): (int(arr[i].launches), float(arr[i].total_ms)) for i in range(min(n.value, cap))}

    def Sync(self):
        N.check(N.load().fgl_ctx_sync(self._ptr()), self._ptr())

    def ForceCapacities(self, side_triangles=0, clipped_vertices=0, pairs=0, spans=0, fragments=0):
        """Failure injection: the next draw's first pass runs with at most
        these capacities (0 = as computed), overflows, grows its buffers and
        runs again; DrawStats() then reports the overflow bits and reruns."""
        caps = (ctypes.c_uint32 * 5)(side_triangles, clipped_vertices, pairs, spans, fragments)
        N.check(N.load().fgl_ctx_force_capacities(self._ptr(), caps), self._ptr())

    def DrawStats(self):
        """Work counts of the last draw (pairs, spans, fragments, ...); a
        diagnostic with no reference counterpart."""
        st = N.FglDrawStat()
        N.check(N.load().fgl_draw_stats(self._ptr(), ctypes.byref(st)), self._ptr())
        return {name: int(getattr(st, name)) for name, _ in N.FglDrawStat._fields_}


def TileShapes():
    """The compiled tile shapes as (w, h) pairs."""
    n = ctypes.c_int32()
    buf = (ctypes.c_int32 * 32)()
    N.check(N.load().fgl_tile_shapes(buf, 16, ctypes.byref(n)))
    return [(int(buf[2 * i]), int(buf[2 * i + 1])) for i in range(min(n.value, 16))]


def SaveTuning(path):
    """Writes the tuning table (tile shape and capacities per workload)."""
    N.check(N.load().fgl_tune_save(os.fsencode(path)))


def LoadTuning(path):
    """Merges a tuning file written by SaveTuning into the table."""
    N.check(N.load().fgl_tune_load(os.fsencode(path)))


def ClearTuning():
    N.check(N.load().fgl_tune_clear())


def NewContext(width, height):
    return Context(width, height)
