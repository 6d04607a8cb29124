"""Mesh generators and mesh preparation (host side, numpy).

These build the inputs of the draw path; they are not on it.  Each follows
the reference function it names (shapes.go, mesh.go, triangle.go) with the
same operation order, so shared vertices stay bit-identical and smoothing
groups form exactly as in the reference.  Arrays use the C ABI's SoA
layout: position / normal (n, 3, 3), texture (n, 3, 2), color (n, 3, 4).

The seeded workloads at the bottom are the benchmark and parity scenes.
"""
import math

import numpy as np

from .api import (Color, HexColor, LookAt, Matrix, Mesh, NewPhongShader, Perspective, Prims, Radians, Scale,
                  TextureShader, ImageTexture, Translate, V, Vector, Identity)

# ---------------------------------------------------------------------------
# vector helpers over (..., 3) arrays, in the reference's operation order


def _dot(a, b):
    return a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1] + a[..., 2] * b[..., 2]


def _cross(a, b):
    x = a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1]
    y = a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2]
    z = a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]
    return np.stack([x, y, z], axis=-1)


def _normalize(a):
    """Vector.Normalize (vector.go:83-86): multiply by 1/sqrt(len^2)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        r = 1.0 / np.sqrt(a[..., 0] * a[..., 0] + a[..., 1] * a[..., 1] + a[..., 2] * a[..., 2])
    return a * r[..., None]


def face_normals(pos):
    """Triangle.Normal (triangle.go:33-37) per triangle, shape (n, 3)."""
    e1 = pos[:, 1] - pos[:, 0]
    e2 = pos[:, 2] - pos[:, 0]
    return _normalize(_cross(e1, e2))


def with_face_normals(pos):
    """NewTriangleForPoints -> FixNormals: every vertex gets the face normal."""
    n = face_normals(pos)
    return np.repeat(n[:, None, :], 3, axis=1)


# ---------------------------------------------------------------------------
# shapes.go

ICO_A = 0.8506507174597755
ICO_B = 0.5257312591858783


def new_icosahedron():
    """NewIcosahedron (shapes.go:189-236): positions (20, 3, 3)."""
    a, b = ICO_A, ICO_B
    v = np.array([[-a, -b, 0], [-a, b, 0], [-b, 0, -a], [-b, 0, a], [0, -a, -b], [0, -a, b], [0, a, -b], [0, a, b],
                  [b, 0, -a], [b, 0, a], [a, -b, 0], [a, b, 0]], dtype=np.float64)
    idx = np.array([[0, 3, 1], [1, 3, 7], [2, 0, 1], [2, 1, 6], [4, 0, 2], [4, 5, 0], [5, 3, 0], [6, 1, 7],
                    [6, 7, 11], [7, 3, 9], [8, 2, 6], [8, 4, 2], [8, 6, 11], [8, 10, 4], [8, 11, 10], [9, 3, 5],
                    [10, 5, 4], [10, 9, 5], [11, 7, 9], [11, 9, 10]])
    return v[idx]


_SPHERE_CACHE = {}


def sphere_positions_cached(detail):
    """sphere_positions(detail), computed once per detail (read-only)."""
    if detail not in _SPHERE_CACHE:
        a = sphere_positions(detail)
        a.setflags(write=False)
        _SPHERE_CACHE[detail] = a
    return _SPHERE_CACHE[detail]


def sphere_positions(detail):
    """NewSphere(detail) (shapes.go:107-133) triangle positions in the
    reference's depth-first order: each level replaces a triangle by
    (v1,v12,v13), (v2,v23,v12), (v3,v13,v23), (v12,v23,v13) in place."""
    t = new_icosahedron()
    for _ in range(detail):
        v1, v2, v3 = t[:, 0], t[:, 1], t[:, 2]
        v12 = _normalize((v1 + v2) / 2)
        v13 = _normalize((v1 + v3) / 2)
        v23 = _normalize((v2 + v3) / 2)
        kids = np.stack([np.stack([v1, v12, v13], 1), np.stack([v2, v23, v12], 1), np.stack([v3, v13, v23], 1),
                         np.stack([v12, v23, v13], 1)], axis=1)
        t = kids.reshape(-1, 3, 3)
    return np.ascontiguousarray(t)


def new_sphere(detail):
    pos = sphere_positions(detail)
    return Prims(pos, with_face_normals(pos))


def new_cube():
    """NewCube (shapes.go:16-37): 12 triangles, then Transform(Scale(0.5))."""
    v = np.array([[-1, -1, -1], [-1, -1, 1], [-1, 1, -1], [-1, 1, 1], [1, -1, -1], [1, -1, 1], [1, 1, -1],
                  [1, 1, 1]], dtype=np.float64)
    idx = [[3, 5, 7], [5, 3, 1], [0, 6, 4], [6, 0, 2], [0, 5, 1], [5, 0, 4], [5, 6, 7], [6, 5, 4], [6, 3, 7],
           [3, 6, 2], [0, 3, 2], [3, 0, 1]]
    pos = v[np.array(idx)]
    p = Prims(pos, with_face_normals(pos))
    return transform(p, Scale(V(0.5, 0.5, 0.5)))


def new_plane():
    """NewPlane (shapes.go:3-14)."""
    v1, v2, v3, v4 = [-0.5, -0.5, 0], [0.5, -0.5, 0], [0.5, 0.5, 0], [-0.5, 0.5, 0]
    pos = np.array([[v1, v2, v3], [v1, v3, v4]], dtype=np.float64)
    return Prims(pos, with_face_normals(pos))


# ---------------------------------------------------------------------------
# mesh.go / triangle.go


def transform(p, m):
    """Mesh.Transform (mesh.go:167-175): MulPosition on positions,
    MulDirection (normalized) on normals."""
    m = Matrix(m)
    x, y, z = p.position[..., 0], p.position[..., 1], p.position[..., 2]
    pos = np.stack([m[0] * x + m[1] * y + m[2] * z + m[3], m[4] * x + m[5] * y + m[6] * z + m[7],
                    m[8] * x + m[9] * y + m[10] * z + m[11]], axis=-1)
    nrm = None
    if p.normal is not None:
        x, y, z = p.normal[..., 0], p.normal[..., 1], p.normal[..., 2]
        nrm = _normalize(np.stack([m[0] * x + m[1] * y + m[2] * z, m[4] * x + m[5] * y + m[6] * z,
                                   m[8] * x + m[9] * y + m[10] * z], axis=-1))
    return Prims(pos, nrm, p.texture, p.color, nv=p.nv)


def bounding_box(p):
    pos = p.position.reshape(-1, 3)
    return Vector(*map(float, pos.min(axis=0))), Vector(*map(float, pos.max(axis=0)))


def fit_inside(p, box_min, box_max, anchor):
    """Mesh.FitInside (mesh.go:142-151)."""
    bmin, bmax = bounding_box(p)
    size = bmax.Sub(bmin)
    bsize = box_max.Sub(box_min)
    scale = bsize.Div(size).MinComponent()
    extra = bsize.Sub(size.MulScalar(scale))
    m = Identity()
    m = m.Translate(bmin.Negate())
    m = m.Scale(Vector(scale, scale, scale))
    m = m.Translate(box_min.Add(extra.Mul(anchor)))
    return transform(p, m)


def bi_unit_cube(p):
    """Mesh.BiUnitCube (mesh.go:127-130)."""
    return fit_inside(p, V(-1, -1, -1), V(1, 1, 1), V(0.5, 0.5, 0.5))


def smooth_normals_threshold(p, radians):
    """Mesh.SmoothNormalsThreshold (mesh.go:80-103): each vertex normal
    becomes the normalized sum, in triangle order, of the normals stored at
    the same position whose dot with it is >= cos(radians)."""
    threshold = math.cos(radians)
    pos = p.position.reshape(-1, 3) + 0.0  # Go map keys treat -0 == +0
    nrm = p.normal.reshape(-1, 3)
    n = pos.shape[0]
    order = np.lexsort((pos[:, 2], pos[:, 1], pos[:, 0]))  # stable: keeps occurrence order per key
    sp = pos[order]
    new_group = np.ones(n, dtype=bool)
    new_group[1:] = np.any(sp[1:] != sp[:-1], axis=1)
    gid_sorted = np.cumsum(new_group) - 1
    starts = np.flatnonzero(new_group)
    sizes = np.diff(np.append(starts, n))
    gid = np.empty(n, dtype=np.int64)
    gid[order] = gid_sorted
    rank = np.empty(n, dtype=np.int64)
    rank[order] = np.arange(n) - starts[gid_sorted]
    members = order  # members of group g: order[starts[g] : starts[g] + sizes[g]]
    result = np.zeros_like(nrm)
    own = nrm
    gsize = sizes[gid]
    gstart = starts[gid]
    for k in range(int(sizes.max())):
        has = gsize > k
        j = members[np.minimum(gstart + k, n - 1)]
        x = nrm[j]
        ok = has & (_dot(x, own) >= threshold)
        result = np.where(ok[:, None], result + x, result)
    out = _normalize(result).reshape(p.normal.shape)
    return Prims(p.position, out, p.texture, p.color, nv=p.nv)


def concat(*ps):
    def cat(attr):
        arrs = [getattr(q, attr) for q in ps]
        if all(a is None for a in arrs):
            return None
        return np.concatenate([a if a is not None else np.zeros((len(q), q.nv, {"normal": 3, "texture": 2,
                                                                              "color": 4}[attr]))
                               for a, q in zip(arrs, ps)])
    return Prims(np.concatenate([q.position for q in ps]), cat("normal"), cat("texture"), cat("color"),
                 nv=ps[0].nv)


# ---------------------------------------------------------------------------
# seeded noise (the workload's displacement and texture; not from the reference)


def _hash3(ix, iy, iz, seed):
    """Integer lattice hash -> float in [-1, 1)."""
    h = (ix.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)) ^ \
        (iy.astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F)) ^ \
        (iz.astype(np.uint64) * np.uint64(0x165667B19E3779F9)) ^ np.uint64(seed * 0x27D4EB2F165667C5 & (2**64 - 1))
    h ^= h >> np.uint64(31)
    h *= np.uint64(0xBF58476D1CE4E5B9)
    h ^= h >> np.uint64(27)
    h *= np.uint64(0x94D049BB133111EB)
    h ^= h >> np.uint64(33)
    return (h >> np.uint64(11)).astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0


def value_noise(p, seed):
    """Trilinear value noise with smoothstep fade, p: (n, 3)."""
    c = np.floor(p)
    f = p - c
    u = f * f * (3.0 - 2.0 * f)
    ci = c.astype(np.int64)
    out = np.zeros(p.shape[0])
    for dx in (0, 1):
        wx = u[:, 0] if dx else 1.0 - u[:, 0]
        for dy in (0, 1):
            wy = u[:, 1] if dy else 1.0 - u[:, 1]
            for dz in (0, 1):
                wz = u[:, 2] if dz else 1.0 - u[:, 2]
                out = out + wx * wy * wz * _hash3(ci[:, 0] + dx, ci[:, 1] + dy, ci[:, 2] + dz, seed)
    return out


def fbm(p, octaves=6, seed=1, base_freq=2.0):
    """Fractal sum of value noise, normalized to [-1, 1]."""
    total = np.zeros(p.shape[0])
    amp, freq, norm = 1.0, base_freq, 0.0
    for o in range(octaves):
        total = total + amp * value_noise(p * freq, seed * 1000 + o)
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return total / norm


# ---------------------------------------------------------------------------
# workloads

EYE = V(-3, 1, -0.75)
CENTER = V(0, -0.07, 0)
UP = V(0, 1, 0)
LIGHT = V(-0.75, 1, 0.25).Normalize()
OBJECT_COLOR = HexColor("#468966")
BACKGROUND = HexColor("#FFF8E3")


def camera(width, height, fovy=30.0, near=1.0, far=10.0, eye=EYE, center=CENTER, up=UP):
    """LookAt(eye, center, up).Perspective(fovy, aspect, near, far) (README.md:96-97)."""
    return LookAt(eye, center, up).Perspective(fovy, float(width) / float(height), near, far)


def phong_shader(width, height, eye=EYE, **kw):
    sh = NewPhongShader(camera(width, height, eye=eye, **kw), LIGHT, eye)
    sh.ObjectColor = OBJECT_COLOR
    return sh


def sphere_and_cube_mesh(detail=5):
    """NewSphere(5) (20,480 flat-normal triangles) followed by NewCube()
    scaled 0.5 and moved to (0.8, 0, 0)."""
    sphere = new_sphere(detail)
    cube = transform(new_cube(), Scale(V(0.5, 0.5, 0.5)).Translate(V(0.8, 0, 0)))
    return Mesh(concat(sphere, cube))


def noise_sphere_prims(ntri=871306, detail=8, amplitude=0.15, seed=1):
    """Icosphere NewSphere(detail) truncated to ntri triangles in generation
    order, each vertex radius displaced by 1 + amplitude * fbm6(p, seed)
    (a function of position, so shared vertices stay shared), face normals:
    the noise sphere before its mesh preparation."""
    pos = sphere_positions_cached(detail)[:ntri]
    flat = pos.reshape(-1, 3) + 0.0
    uniq, inv = np.unique(flat, axis=0, return_inverse=True)
    disp = 1.0 + amplitude * fbm(uniq, 6, seed)
    moved = (uniq * disp[:, None])[inv.reshape(-1)].reshape(pos.shape)
    return Prims(moved, with_face_normals(moved))


def noise_sphere_mesh(ntri=871306, detail=8, amplitude=0.15, seed=1, smooth_degrees=30.0):
    """noise_sphere_prims, then BiUnitCube and
    SmoothNormalsThreshold(Radians(smooth_degrees))."""
    p = noise_sphere_prims(ntri, detail, amplitude, seed)
    p = bi_unit_cube(p)
    p = smooth_normals_threshold(p, Radians(smooth_degrees))
    return Mesh(p)


def noise_texture(size=1024, seed=2, alpha_lo=0.2, alpha_hi=0.8):
    """size x size NRGBA8 value-noise texture with alpha in [alpha_lo, alpha_hi]."""
    ys, xs = np.meshgrid(np.arange(size), np.arange(size), indexing="ij")
    p = np.stack([xs.ravel() / 64.0, ys.ravel() / 64.0, np.zeros(size * size)], axis=1)
    chans = []
    for c in range(3):
        n = fbm(p + c * 17.0, 4, seed + c, 1.0)
        chans.append(np.clip((n * 0.5 + 0.5) * 255.0, 0, 255))
    na = fbm(p + 51.0, 4, seed + 7, 1.0) * 0.5 + 0.5
    a = (alpha_lo + (alpha_hi - alpha_lo) * np.clip(na, 0, 1)) * 255.0
    tex = np.stack(chans + [a], axis=1).astype(np.uint8).reshape(size, size, 4)
    return tex


def triangle_soup(n=200000, seed=2, lmin=0.002, lmax=0.2):
    """n triangles: centroid uniform in [-1,1]^3, size L log-uniform in
    [lmin, lmax], vertices at c + L * (random unit directions), UVs uniform."""
    rng = np.random.Generator(np.random.PCG64(seed))
    c = rng.uniform(-1.0, 1.0, size=(n, 1, 3))
    L = np.exp(rng.uniform(math.log(lmin), math.log(lmax), size=(n, 1, 1)))
    d = rng.normal(size=(n, 3, 3))
    d = _normalize(d)
    pos = c + L * d
    uv = rng.uniform(0.0, 1.0, size=(n, 3, 2))
    return Mesh(Prims(pos, with_face_normals(pos), uv))


def soup_scene(width=1920, height=1080, n=200000, seed=2, tex_size=1024):
    mesh = triangle_soup(n, seed)
    tex = ImageTexture(noise_texture(tex_size, seed), "NRGBA")
    shader = TextureShader(camera(width, height), tex)
    return mesh, shader


# ---------------------------------------------------------------------------
# batched multi-view workload

BATCH_RESOLUTIONS = [(256, 256), (512, 512), (1024, 1024), (1920, 1080)]


def batch_workload(nmesh=32, views_per_mesh=8, seed=4, nmin=1e3, nmax=1e6, radius=3.5,
                   resolutions=BATCH_RESOLUTIONS):
    """Noise-displaced icospheres (as the headline mesh) with triangle counts log-uniform in
    [nmin, nmax] (raw: face normals, before BiUnitCube and smoothing), and
    views_per_mesh views of each: eye on a sphere of the given radius at a
    random azimuth and elevation, looking at the origin, resolution drawn
    from resolutions.  Returns ([Prims], [(mesh index, eye, W, H)])."""
    rng = np.random.Generator(np.random.PCG64(seed))
    meshes, views = [], []
    for m in range(nmesh):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        d = 0
        while 20 * 4 ** d < n:
            d += 1
        meshes.append(noise_sphere_prims(ntri=n, detail=d, amplitude=0.15, seed=100 + m))
        for _ in range(views_per_mesh):
            az = rng.uniform(0, 2 * math.pi)
            el = rng.uniform(-1.2, 1.2)
            eye = V(radius * math.cos(el) * math.cos(az), radius * math.sin(el), radius * math.cos(el) * math.sin(az))
            w, h = resolutions[int(rng.integers(len(resolutions)))]
            views.append((m, eye, w, h))
    return meshes, views


def batch_shader(w, h, eye):
    """Phong shader of a batch view: camera at eye looking at the origin."""
    return phong_shader(w, h, eye=eye, center=V(0, 0, 0))
