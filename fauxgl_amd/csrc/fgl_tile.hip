// fgl_tile.hip -- shading and the ordered framebuffer update.
//
// k_shade: one thread per coverage span (fgl_spans.hip).  It fetches the
// raster triangle's record and vertex attributes once, then for each
// covered pixel of the span replays the reference's per-pixel recurrence
// from the span's first edge values (context.go:208-213), computes z
// (context.go:230), the perspective-correct interpolants
// (context.go:236-238, vertex.go:18-79) and the built-in fragment shader
// (shader.go:25-96, texture.go:41-63), and writes one Frag in submission
// order.  Shading has no side effects, so shading a fragment that the
// ordered update later rejects changes nothing observable.
//
// k_rop: one workgroup per TILE_W x TILE_H tile keeps the tile's color and
// depth in LDS and applies the tile's fragments in submission order
// (context.go:232-273): in each round every pixel's lowest pending fragment
// wins and does the early depth test, the Discard check, the locked depth
// test and the depth / color write exactly as the serial loop would at
// that moment.
#include "fgl_internal.h"
#include "fgl_scan.h"
#include "fgl_tiles.h"
#include "fgl_draws.h"

namespace fgl {
namespace FGL_TILE_NS {

#ifndef FGL_ROP_CHUNK
#define FGL_ROP_CHUNK 1024
#endif
constexpr int ROP_CHUNK = FGL_ROP_CHUNK;                 // fragments per round set
constexpr int ROP_PER_THREAD = ROP_CHUNK / TILE_THREADS;
constexpr uint32_t OWNER_NONE = 0xffffffffu;

// ---------------------------------------------------------------------------
// math.Pow (see fgl_device.h)

// x**yf for the fractional part of a non-integer exponent; kept out of line
// (the library exp/log are large and only non-integer powers need them).
__device__ __noinline__ double pow_frac(double x, double yf) { return exp(yf * log(x)); }

// math.Pow(x, sh.power).  The y-only steps (Modf, the yf > 0.5 adjustment)
// were made on the host, so the squaring loop runs over the bits of yi in
// scalar registers; xe and ae stay within +-2^19 here (the loop stops once
// |xe| > 2^12), so 32-bit integers hold Go's int values exactly.
__device__ double go_pow(double x, const DevShader &sh) {
    const double y = sh.power;
    if (y == 0 || x == 1) return 1;
    if (y == 1) return x;
    if (__builtin_isnan(x) || __builtin_isnan(y)) return __builtin_nan("");
    if (x == 0) {
        if (y < 0) {
            if (__builtin_signbit(x) && go_is_odd_int(y)) return -__builtin_inf();
            return __builtin_inf();
        }
        if (__builtin_signbit(x) && go_is_odd_int(y)) return x;
        return 0;
    }
    if (__builtin_isinf(y)) {
        if (x == -1) return 1;
        if ((__builtin_fabs(x) < 1) == (y > 0)) return 0;
        return __builtin_inf();
    }
    if (__builtin_isinf(x)) {
        if (x < 0) {
            // Pow(-Inf, y) = Pow(-0, -y)
            double ny = -y;
            if (ny < 0) return go_is_odd_int(ny) ? -__builtin_inf() : __builtin_inf();
            return go_is_odd_int(ny) ? -0.0 : 0.0;
        }
        if (y < 0) return 0;
        return __builtin_inf();
    }
    if (y == 0.5) return __builtin_sqrt(x);
    if (y == -0.5) return 1.0 / __builtin_sqrt(x);
    if (sh.pw_frac && x < 0) return __builtin_nan("");
    if (sh.pw_big) {
        if (x == -1) return 1;
        if ((__builtin_fabs(x) < 1) == (y > 0)) return 0;
        return __builtin_inf();
    }
    double a1 = sh.pw_frac ? pow_frac(x, sh.pw_yf) : 1.0;
    if (x >= sh.pw_lo && x <= sh.pw_hi) {
        // Every value below stays normal (the host's bound on x), and there
        // scaling by a power of two commutes with rounding: each product is
        // the frexp form's product times its exact power of two, so the same
        // double results without the exponent bookkeeping.
        double x1 = x;
        for (uint64_t i = sh.pw_yi; i != 0;) {
            if (i & 1) a1 *= x1;
            i >>= 1;
            if (i != 0) x1 *= x1;
        }
        return y < 0 ? 1.0 / a1 : a1;
    }
    int32_t ae = 0;
    int xe;
    double x1 = frexp(x, &xe);
    // |xe| at most doubles (plus one) per step, so when (|xe| + 1) shifted
    // by the step count stays within 2^12 the overflow exit cannot be taken
    // and the loop runs without it.
    const int32_t axe = xe < 0 ? -xe : xe;
    if (sh.pw_niter <= 13 && axe + 1 <= ((1 << 12) >> (sh.pw_niter > 0 ? sh.pw_niter - 1 : 0))) {
        for (uint64_t i = sh.pw_yi; i != 0; i >>= 1) {
            if (i & 1) {
                a1 *= x1;
                ae += xe;
            }
            x1 *= x1;
            xe <<= 1;
            if (x1 < .5) {
                x1 += x1;
                xe--;
            }
        }
    } else {
        for (uint64_t i = sh.pw_yi; i != 0; i >>= 1) {
            if (xe < -(1 << 12) || (1 << 12) < xe) {
                ae += xe;
                break;
            }
            if (i & 1) {
                a1 *= x1;
                ae += xe;
            }
            x1 *= x1;
            xe <<= 1;
            if (x1 < .5) {
                x1 += x1;
                xe--;
            }
        }
    }
    if (y < 0) {
        a1 = 1.0 / a1;
        ae = -ae;
    }
    return ldexp(a1, ae);
}

// ---------------------------------------------------------------------------
// textures (texture.go:41-63; image.Image.At(x,y).RGBA() then MakeColor)

FGL_D C4 texel(const DevShader &sh, int64_t x, int64_t y) {
    uint32_t r = 0, g = 0, b = 0, al = 0;
    if (x >= 0 && y >= 0 && x < sh.tex_w && y < sh.tex_h) {
        int64_t i = (y * sh.tex_w + x) * 4;
        if (sh.tex_kind == 3) {  // RGBA16
            const uint16_t *p = (const uint16_t *)sh.tex_pix;
            r = p[i]; g = p[i + 1]; b = p[i + 2]; al = p[i + 3];
        } else {
            uint32_t w = ((const uint32_t *)sh.tex_pix)[i >> 2];
            uint32_t R = w & 0xff, G = (w >> 8) & 0xff, B = (w >> 16) & 0xff, A = w >> 24;
            r = R | (R << 8);
            g = G | (G << 8);
            b = B | (B << 8);
            al = A | (A << 8);
            if (sh.tex_kind == 1) {  // NRGBA: premultiply as color.NRGBA.RGBA()
                r = r * A / 0xff;
                g = g * A / 0xff;
                b = b * A / 0xff;
            }
        }
    }
    const double d = 65535.0;
    return c4((double)r / d, (double)g / d, (double)b / d, (double)al / d);
}

FGL_D C4 cadd_scaled(C4 c, C4 t, double s) { return c4(c.r + t.r * s, c.g + t.g * s, c.b + t.b * s, c.a + t.a * s); }

__device__ __noinline__ C4 bilinear_sample(const DevShader &sh, double u, double v) {
    v = 1 - v;
    u -= __builtin_floor(u);
    v -= __builtin_floor(v);
    double x = u * (double)(sh.tex_w - 1);
    double y = v * (double)(sh.tex_h - 1);
    int64_t x0 = go_f2i(x);
    int64_t y0 = go_f2i(y);
    int64_t x1 = x0 + 1;
    int64_t y1 = y0 + 1;
    x -= (double)x0;
    y -= (double)y0;
    C4 c00 = texel(sh, x0, y0);
    C4 c01 = texel(sh, x0, y1);
    C4 c10 = texel(sh, x1, y0);
    C4 c11 = texel(sh, x1, y1);
    C4 c = c4(0, 0, 0, 0);
    c = cadd_scaled(c, c00, (1 - x) * (1 - y));
    c = cadd_scaled(c, c10, x * (1 - y));
    c = cadd_scaled(c, c01, (1 - x) * y);
    c = cadd_scaled(c, c11, x * y);
    return c;
}

// ---------------------------------------------------------------------------
// vertex attribute fetch for the shaders

// Per-vertex attribute arrays.  A missing array (a mesh without normals,
// texture coordinates or colors) is replaced by a run of zeros read at
// offset 0 (its mask is 0), so every fetch is the same branch-free address
// computation.
struct AttrSrc {
    const double *pos, *nrm, *uv, *col;  // stride 3, 3, 2, 4 doubles
    uint32_t mpos, mnrm, muv, mcol;      // offset masks: ~0 present, 0 missing
    const double *side;                  // side vertices (stride SIDE_VTX_DOUBLES)
};

FGL_D AttrSrc attr_src(const DrawArgs &a) {
    AttrSrc s;
    s.pos = a.in.pos ? a.in.pos : a.zeros;
    s.nrm = a.in.nrm ? a.in.nrm : a.zeros;
    s.uv = a.in.uv ? a.in.uv : a.zeros;
    s.col = a.in.col ? a.in.col : a.zeros;
    s.mpos = a.in.pos ? ~0u : 0u;
    s.mnrm = a.in.nrm ? ~0u : 0u;
    s.muv = a.in.uv ? ~0u : 0u;
    s.mcol = a.in.col ? ~0u : 0u;
    s.side = a.side_vtx;
    return s;
}

// The three vertices of a raster triangle: vertex ids into the input arrays
// (triangle 3i+sel, line 2i+sel) or into the side vertices (i+sel).
// Offsets fit 32 bits: at most 2^26 input primitives (3 * 2^26 vertices
// times 4 doubles) and 2^28 side vertices (times SIDE_VTX_DOUBLES).
struct VtxRef {
    uint32_t vid[3];
    bool side;
};

FGL_D VtxRef vtx_ref(uint64_t attr) {
    const uint32_t kind = attr_kind(attr);
    const uint32_t idx = (uint32_t)attr_index(attr);
    const uint32_t b = kind == ATTR_TRI ? idx + (idx << 1) : kind == ATTR_LINE ? idx << 1 : idx;
    VtxRef r;
#pragma unroll
    for (int k = 0; k < 3; k++) r.vid[k] = b + attr_sel(attr, k);
    r.side = kind == ATTR_SIDE;
    return r;
}

// Address of vertex k's attribute: the input array (scale doubles per
// vertex, offset mask m) or the side vertex at offset side_off; both the
// base and the offset are selected, not branched on.
FGL_D const double *attr_ptr(const AttrSrc &src, const VtxRef &r, int k, const double *arr, uint32_t scale,
                             uint32_t m, uint32_t side_off) {
    const uint32_t vid = r.vid[k];
    const uint32_t in_off = (vid * scale) & m;
    const uint32_t side_at = vid * (uint32_t)SIDE_VTX_DOUBLES + side_off;
    const double *base = r.side ? src.side : arr;
    return base + (r.side ? side_at : in_off);
}

FGL_D V3 fetch_pos(const AttrSrc &src, const VtxRef &r, int k) {
    const double *q = attr_ptr(src, r, k, src.pos, 3u, src.mpos, 0u);
    return v3(q[0], q[1], q[2]);
}
FGL_D V3 fetch_nrm(const AttrSrc &src, const VtxRef &r, int k) {
    const double *q = attr_ptr(src, r, k, src.nrm, 3u, src.mnrm, 3u);
    return v3(q[0], q[1], q[2]);
}
FGL_D void fetch_uv(const AttrSrc &src, const VtxRef &r, int k, double &u, double &v) {
    const double *q = attr_ptr(src, r, k, src.uv, 2u, src.muv, 6u);
    u = q[0];
    v = q[1];
}
FGL_D C4 fetch_col(const AttrSrc &src, const VtxRef &r, int k) {
    const double *q = attr_ptr(src, r, k, src.col, 4u, src.mcol, 8u);
    return c4(q[0], q[1], q[2], q[3]);
}

// InterpolateFloats-style sum: ((((0 + a*bx) + b*by) + c*bz) * bw)
FGL_D double interp(double a, double b, double c, V4 k) {
    double n = 0;
    n = n + a * k.x;
    n = n + b * k.y;
    n = n + c * k.z;
    return n * k.w;
}

// Vertex attributes of one raster triangle, fetched once per span.
struct ShadeIn {
    V3 p[3], n[3];        // PhongShader: Position, Normal
    double u[3], v[3];    // Texture.X, Texture.Y
    C4 c[3];              // Color (PhongShader with ObjectColor == Discard)
};

// SH: 1 SolidColorShader, 2 TextureShader, 3 PhongShader, 4 PhongShader
// with a Texture (its own instantiation keeps the sampler's registers out
// of the untextured Phong kernel).
// COLS: also the vertex colors Phong reads when ObjectColor is Discard.
template <int SH, bool COLS>
FGL_D void load_in(const DevShader &sh, const AttrSrc &src, uint64_t attr, ShadeIn &in) {
    const bool uvs = SH == 2 || SH == 4;
    const bool cols = COLS && SH == 3 && sh.obj_is_discard;
    const VtxRef r = vtx_ref(attr);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        if (SH >= 3) {
            in.p[k] = fetch_pos(src, r, k);
            in.n[k] = fetch_nrm(src, r, k);
        }
        if (uvs) fetch_uv(src, r, k, in.u[k], in.v[k]);
        if (cols) in.c[k] = fetch_col(src, r, k);
    }
}

// The fragment shaders (shader.go) on interpolated attributes.
template <int SH>
FGL_D C4 eval(const DevShader &sh, const ShadeIn &in, V4 b) {
    if (SH == 1) {  // SolidColorShader.Fragment (shader.go:25-27)
        return c4(sh.color[0], sh.color[1], sh.color[2], sh.color[3]);
    } else if (SH == 2) {  // TextureShader.Fragment (shader.go:44-46)
        return bilinear_sample(sh, interp(in.u[0], in.u[1], in.u[2], b), interp(in.v[0], in.v[1], in.v[2], b));
    } else {  // PhongShader.Fragment (shader.go:75-96)
        V3 P = v3(interp(in.p[0].x, in.p[1].x, in.p[2].x, b), interp(in.p[0].y, in.p[1].y, in.p[2].y, b),
                  interp(in.p[0].z, in.p[1].z, in.p[2].z, b));
        V3 N = vnormalize(v3(interp(in.n[0].x, in.n[1].x, in.n[2].x, b), interp(in.n[0].y, in.n[1].y, in.n[2].y, b),
                             interp(in.n[0].z, in.n[1].z, in.n[2].z, b)));
        C4 light = c4(sh.amb[0], sh.amb[1], sh.amb[2], sh.amb[3]);
        C4 color;
        if (SH == 4) {
            color = bilinear_sample(sh, interp(in.u[0], in.u[1], in.u[2], b), interp(in.v[0], in.v[1], in.v[2], b));
        } else if (!sh.obj_is_discard) {
            color = c4(sh.obj[0], sh.obj[1], sh.obj[2], sh.obj[3]);
        } else {
            color = c4(interp(in.c[0].r, in.c[1].r, in.c[2].r, b), interp(in.c[0].g, in.c[1].g, in.c[2].g, b),
                       interp(in.c[0].b, in.c[1].b, in.c[2].b, b), interp(in.c[0].a, in.c[1].a, in.c[2].a, b));
        }
        V3 L = v3(sh.light[0], sh.light[1], sh.light[2]);
        double diffuse = go_max0(vdot(N, L));
        light = cadd_scaled(light, c4(sh.dif[0], sh.dif[1], sh.dif[2], sh.dif[3]), diffuse);
        if (diffuse > 0 && sh.power > 0) {
            V3 camera = vnormalize(vsub(v3(sh.cam[0], sh.cam[1], sh.cam[2]), P));
            V3 i = v3(-L.x, -L.y, -L.z);
            V3 reflected = vsub(i, vmuls(N, 2 * vdot(N, i)));
            double specular = go_max0(vdot(camera, reflected));
            if (specular > 0) {
                specular = go_pow(specular, sh);
                light = cadd_scaled(light, c4(sh.spc[0], sh.spc[1], sh.spc[2], sh.spc[3]), specular);
            }
        }
        return c4(go_min1(color.r * light.r), go_min1(color.g * light.g), go_min1(color.b * light.b), color.a);
    }
}

FGL_D const PrimRec *prim_of(const DrawArgs &a, uint32_t ref) {
    return (ref & REF_SIDE) ? a.side_rec + (ref & ~REF_SIDE) : a.rec_main + ref;
}

// ---------------------------------------------------------------------------
// fragment offsets: exclusive scan of the per-block fragment counts k_spans
// accumulated

// Workgroups of the shading kernel loop over the FRAG_BLOCK-span blocks
// that exist, so the grid (sized from capacity) need not match the count.
constexpr uint32_t SPAN_GRID_MAX = 2048;

// Two levels: each workgroup scans FRAG_GROUP consecutive span-block sums
// (coalesced) into group-local offsets and writes the group's total; the
// shading kernel adds the totals of the groups before its block.
constexpr int FRAG_GROUP = 1024;
// Group offsets the shading kernel keeps in LDS (groups beyond, more than
// FRAG_GTAB * FRAG_GROUP * FRAG_BLOCK spans, add their sums one by one).
constexpr int FRAG_GTAB = 256;

template <class G>
FGL_D void frag_scan_body(const DrawArgs &a, const G &grid) {
    __shared__ uint64_t wsum[FRAG_GROUP / 64];
    if (a.ctr->overflow) return;
    const uint32_t NS = a.ctr->nspans;
    const uint32_t nb = (NS + FRAG_BLOCK - 1) / FRAG_BLOCK;
    const uint32_t g = grid.bid();
    if ((uint64_t)g * FRAG_GROUP >= nb) return;
    const uint32_t b = g * FRAG_GROUP + threadIdx.x;
    const uint64_t v = b < nb ? a.frag_blocksum[b] : 0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t incl = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t y = __shfl_up(incl, off);
        if (lane >= off) incl += y;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint64_t base = 0, total = 0;
    for (int w = 0; w < FRAG_GROUP / 64; w++) {
        const uint64_t x = wsum[w];
        if (w < wave) base += x;
        total += x;
    }
    if (b < nb) a.frag_blockoff[b] = base + incl - v;
    if (threadIdx.x == 0) a.frag_gsum[g] = total;
}
FGL_DRAW_KERNELS(k_frag_scan, frag_scan_body, __launch_bounds__(FRAG_GROUP))

// ---------------------------------------------------------------------------
// shading
//
// Each wave takes 64 consecutive spans (one per lane) and shades their
// fragments 64 at a time, one fragment per lane: a lane finds the span that
// owns its fragment (each span marks the window position of its first
// fragment, a max-scan carries the mark over the span's pixels), takes the
// span's edge values from the owning lane, replays the reference's
// per-pixel additions up to its pixel and runs the shader.

// Waves per SIMD the untextured Phong kernel is compiled for (its register
// budget).
#ifndef FGL_SHADE_WAVES
#define FGL_SHADE_WAVES 4
#endif

// Per-triangle shading inputs a wave shares: the first span of each run of
// spans of one raster triangle (runs are consecutive in the span list: one
// triangle's rows within one tile) loads the triangle's record fields and
// vertex attributes once into an LDS slot, and the fragments of every span
// of the run read them from there.  Fragment lanes of one triangle read the
// same LDS address (a broadcast), so the per-fragment vector-memory return
// traffic that bounded the one-load-per-fragment form is gone.  Triangles
// past the wave's SHADE_SLOTS slots load from global memory per fragment.
#ifndef FGL_SHADE_SLOTS
#define FGL_SHADE_SLOTS 32
#endif
constexpr int SHADE_SLOTS = FGL_SHADE_SLOTS;

// Slot layout in doubles: a12, a20, a01 (the per-pixel steps of the edge
// values), sz[3], r[3], ra, the attr word, one pad, then the attributes the
// shader reads (Phong: positions, normals; textures: u, v).
template <int SH>
struct SlotLayout {
    static constexpr bool kPN = SH >= 3;
    static constexpr bool kUV = SH == 2 || SH == 4;
    static constexpr int kAttr = 12;
    static constexpr int kUVat = kAttr + (kPN ? 18 : 0);
    static constexpr int N = kUVat + (kUV ? 6 : 0);
};

// The per-fragment inputs taken from a raster triangle's record.
struct TriIn {
    double a12, a20, a01;
    double sz[3], r[3], ra;
    uint64_t attr;
};

FGL_D void tri_from_rec(const PrimRec *pr, TriIn &t) {
    t.a12 = pr->sy[2] - pr->sy[1];
    t.a20 = pr->sy[0] - pr->sy[2];
    t.a01 = pr->sy[1] - pr->sy[0];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        t.sz[k] = pr->sz[k];
        t.r[k] = pr->r[k];
    }
    t.ra = pr->ra;
    t.attr = pr->attr;
}

template <int SH>
FGL_D void slot_store(double *d, const TriIn &t, const ShadeIn &in) {
    using L = SlotLayout<SH>;
    d[0] = t.a12;
    d[1] = t.a20;
    d[2] = t.a01;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        d[3 + k] = t.sz[k];
        d[6 + k] = t.r[k];
    }
    d[9] = t.ra;
    d[10] = __builtin_bit_cast(double, t.attr);
    d[11] = 0.0;
    if (L::kPN) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            d[L::kAttr + 3 * k + 0] = in.p[k].x;
            d[L::kAttr + 3 * k + 1] = in.p[k].y;
            d[L::kAttr + 3 * k + 2] = in.p[k].z;
            d[L::kAttr + 9 + 3 * k + 0] = in.n[k].x;
            d[L::kAttr + 9 + 3 * k + 1] = in.n[k].y;
            d[L::kAttr + 9 + 3 * k + 2] = in.n[k].z;
        }
    }
    if (L::kUV) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            d[L::kUVat + k] = in.u[k];
            d[L::kUVat + 3 + k] = in.v[k];
        }
    }
}

template <int SH>
FGL_D void slot_load(const double *d, TriIn &t, ShadeIn &in) {
    using L = SlotLayout<SH>;
    t.a12 = d[0];
    t.a20 = d[1];
    t.a01 = d[2];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        t.sz[k] = d[3 + k];
        t.r[k] = d[6 + k];
    }
    t.ra = d[9];
    t.attr = __builtin_bit_cast(uint64_t, d[10]);
    if (L::kPN) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            in.p[k] = v3(d[L::kAttr + 3 * k], d[L::kAttr + 3 * k + 1], d[L::kAttr + 3 * k + 2]);
            in.n[k] = v3(d[L::kAttr + 9 + 3 * k], d[L::kAttr + 9 + 3 * k + 1], d[L::kAttr + 9 + 3 * k + 2]);
        }
    }
    if (L::kUV) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            in.u[k] = d[L::kUVat + k];
            in.v[k] = d[L::kUVat + 3 + k];
        }
    }
}

// Vertex attributes that the slot does not hold (Phong's vertex colors,
// read when ObjectColor is Discard), fetched per fragment.
template <int SH>
FGL_D void load_extra(const DevShader &sh, const AttrSrc &src, uint64_t attr, ShadeIn &in) {
    if (SH == 3 && sh.obj_is_discard) {
        const VtxRef r = vtx_ref(attr);
#pragma unroll
        for (int k = 0; k < 3; k++) in.c[k] = fetch_col(src, r, k);
    }
}

template <int SH, class G>
FGL_D void shade_body(const DrawArgs &a, const G &grid) {
    using L = SlotLayout<SH>;
    __shared__ uint32_t ostart[SPAN_THREADS / 64][64];
    __shared__ double slots[SPAN_THREADS / 64][SHADE_SLOTS][L::N];
    __shared__ uint64_t gpre[FRAG_GTAB];  // exclusive prefix of the group sums
    __shared__ uint64_t gw[SPAN_THREADS / 64];
    if (a.ctr->overflow) return;
    const uint32_t NS = a.ctr->nspans;
    const uint32_t nb = (NS + FRAG_BLOCK - 1) / FRAG_BLOCK;
    const uint32_t ngroups = (nb + FRAG_GROUP - 1) / FRAG_GROUP;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // fragment total and the groups' offsets from the group sums, loaded in
    // parallel and scanned (every workgroup sees the same)
    uint64_t total = 0;
    for (uint32_t g0 = 0; g0 < ngroups; g0 += SPAN_THREADS) {
        const uint32_t g = g0 + threadIdx.x;
        const uint64_t gv = g < ngroups ? a.frag_gsum[g] : 0;
        uint64_t gi = gv;
        for (int off = 1; off < 64; off <<= 1) {
            const uint64_t y = __shfl_up(gi, off);
            if (lane >= off) gi += y;
        }
        if (lane == 63) gw[wave] = gi;
        __syncthreads();
        uint64_t ge = total + gi - gv, gt = 0;
        for (int w = 0; w < SPAN_THREADS / 64; w++) {
            if (w < wave) ge += gw[w];
            gt += gw[w];
        }
        if (g < (uint32_t)FRAG_GTAB) gpre[g] = ge;
        total += gt;
        __syncthreads();
    }
    if (grid.bid() == 0 && threadIdx.x == 0) {
        if (total > a.frag_cap) atomicOr(&a.ctr->overflow, 32u);
        a.ctr->nfrags = total > a.frag_cap ? a.frag_cap : (uint32_t)total;
        a.ctr->frags_lo = (uint32_t)total;
        a.ctr->frags_hi = (uint32_t)(total >> 32);
        a.foff[NS] = total > a.frag_cap ? a.frag_cap : (uint32_t)total;
        // TotalPixels: every fragment is an inside pixel with a valid index
        a.ctr->total = total;
    }
    if (total > a.frag_cap) return;  // the host grows the fragment buffer and reruns
    // each wave takes its own FRAG_BLOCK-span blocks: no workgroup barrier
    // in the loop, so a wave with few fragments does not wait for one with
    // many
    constexpr uint32_t WAVES = SPAN_THREADS / 64;
    for (uint32_t blk = grid.bid() * WAVES + (uint32_t)wave; blk < nb; blk += grid.nblk() * WAVES) {
        const uint32_t i = blk * FRAG_BLOCK + (uint32_t)lane;
        const Span sp = i < NS ? a.spans[i] : Span{{0, 0, 0}, 0, 0, 0};
        const uint32_t n = sp.n & SPAN_N;
        uint32_t incl = n;  // at most 64 * TILE_W
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(incl, off);
            if (lane >= off) incl += y;
        }
        const uint32_t grp = blk / FRAG_GROUP;
        uint64_t fo = a.frag_blockoff[blk];
        if (grp < (uint32_t)FRAG_GTAB) {
            fo += gpre[grp];
        } else {
            fo += gpre[FRAG_GTAB - 1] + a.frag_gsum[FRAG_GTAB - 1];
            for (uint32_t g = FRAG_GTAB; g < grp; g++) fo += a.frag_gsum[g];
        }
        const uint32_t excl = incl - n;
        if (i < NS) a.foff[i] = (uint32_t)(fo + excl);
        const uint32_t T = __shfl(incl, 63);
        const uint32_t pn = (uint32_t)sp.pix | ((uint32_t)sp.n << 16);
        Frag *out = a.frags + fo;

        // the wave's triangles: a span whose predecessor lane holds another
        // triangle starts a run; slot = runs started up to this lane, - 1
        const uint32_t prev_ref = __shfl_up(sp.ref, 1);
        const bool lead = i < NS && (lane == 0 || prev_ref != sp.ref);
        const uint64_t lm = __ballot(lead);
        const uint32_t before = __builtin_amdgcn_mbcnt_hi((uint32_t)(lm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)lm, 0u));
        // staging pays when the wave's triangles average two fragments or
        // more; otherwise every fragment loads its own inputs
        const bool staged = T >= 2u * (uint32_t)__popcll(lm);
        const uint32_t myslot = staged ? before + (lead ? 1u : 0u) - 1u : ~0u;  // lanes >= NS own no fragments
        if (staged && lead && before < (uint32_t)SHADE_SLOTS) {
            const DrawArgs &ka = grid.args();
            const PrimRec *pr = prim_of(ka, sp.ref);
            TriIn t;
            tri_from_rec(pr, t);
            ShadeIn in;
            if (L::kPN || L::kUV) load_in<SH, false>(ka.sh, attr_src(ka), t.attr, in);
            slot_store<SH>(&slots[wave][before][0], t, in);
        }
        __builtin_amdgcn_wave_barrier();

        uint32_t carry = 0;  // owner lane + 1 of the last fragment of the previous window
        for (uint32_t base = 0; base < T; base += 64) {
            const uint32_t p = base + (uint32_t)lane;
            // owner: each span starting in this window marks its start with
            // its lane (+1); a max-scan carries it over the span's pixels
            ostart[wave][lane] = 0;
            if (n && excl >= base && excl - base < 64) ostart[wave][excl - base] = (uint32_t)lane + 1;
            __builtin_amdgcn_wave_barrier();
            uint32_t om = wave_incl_max(ostart[wave][lane]);
            om = om ? om : carry;
            carry = __builtin_amdgcn_readlane(om, 63);
            __builtin_amdgcn_wave_barrier();
            const int ow = (int)om - 1 < 0 ? 0 : (int)om - 1;
            const uint32_t oexcl = __shfl(excl, ow);
            const uint32_t oref = __shfl(sp.ref, ow);
            const uint32_t opn = __shfl(pn, ow);
            const uint32_t oslot = __shfl(myslot, ow);
            double w0 = __shfl(sp.w[0], ow);
            double w1 = __shfl(sp.w[1], ow);
            double w2 = __shfl(sp.w[2], ow);
            if (p < T) {
                const DrawArgs &ka = grid.args();
                const AttrSrc src = attr_src(ka);
                const uint32_t j = p - oexcl;
                TriIn t;
                ShadeIn in;
                if (oslot < (uint32_t)SHADE_SLOTS) {
                    slot_load<SH>(&slots[wave][oslot][0], t, in);
                } else {
                    const PrimRec *pr = prim_of(ka, oref);
                    tri_from_rec(pr, t);
                    if (L::kPN || L::kUV) load_in<SH, false>(ka.sh, src, t.attr, in);
                }
                load_extra<SH>(ka.sh, src, t.attr, in);
                // the reference's per-pixel step, j times (context.go:208-213)
#pragma unroll 4
                for (uint32_t k = 0; k < j; k++) {
                    w0 += t.a12;
                    w1 += t.a20;
                    w2 += t.a01;
                }
                const double b0 = w0 * t.ra, b1 = w1 * t.ra, b2 = w2 * t.ra;
                Frag f;
                f.z = b0 * t.sz[0] + b1 * t.sz[1] + b2 * t.sz[2];
                // perspective-correct interpolation weights (context.go:236-237)
                V4 bw = v4(b0 * t.r[0], b1 * t.r[1], b2 * t.r[2], 0);
                bw.w = 1.0 / (bw.x + bw.y + bw.z);
                const C4 c = eval<SH>(ka.sh, in, bw);
                f.color = nrgba_pack(c);
                f.pix = (uint16_t)((opn & 0xffffu) + j);
                f.flags = ((opn >> 16) & SPAN_INB ? FRAG_INB : (uint16_t)0) | (c.a < 1 ? FRAG_BLEND : (uint16_t)0) |
                          ((c.r == 0 && c.g == 0 && c.b == 0 && c.a == 0) ? FRAG_DISCARD : (uint16_t)0);
                out[p] = f;
            }
        }
    }
}
template <int SH>
__global__ __launch_bounds__(SPAN_THREADS, SH == 3 ? FGL_SHADE_WAVES : 1) void k_shade(DrawArgs a) {
    shade_body<SH>(a, OneDraw{});
}
template <int SH>
__global__ __launch_bounds__(SPAN_THREADS, SH == 3 ? FGL_SHADE_WAVES : 1) void k_shade_m(MultiArgs m) {
    const ManyDraws g = many_draws(m);
    shade_body<SH>(draw_of(g), g);
}

// ---------------------------------------------------------------------------
// the ordered framebuffer update, one workgroup per tile

#ifndef FGL_ROP_WAVES
#define FGL_ROP_WAVES 1
#endif

template <class G>
FGL_D void rop_body(const DrawArgs &a, const G &g) {
    __shared__ uint32_t col[TILE_PIX];
    __shared__ double dep[TILE_PIX];
    __shared__ uint32_t owner[TILE_PIX];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int t = (int)g.bid();
    if (a.ctr->overflow) return;  // the host reruns the whole draw
    __shared__ uint32_t upd[TILE_THREADS / 64];
    const int tx = t % a.ntx, ty = t / a.ntx;
    const int cx0 = tx * TILE_W, cy0 = ty * TILE_H;
    const int cw = min(TILE_W, a.W - cx0), chh = min(TILE_H, a.H - cy0);
    // the tile's fragment range [foff[poff[tstart]], foff[poff[tend]])
    const uint32_t ps = a.tstart[t], pe = a.tend[t];
    const uint32_t f0 = pe > ps ? a.foff[a.poff[ps]] : 0u;
    const uint32_t f1 = pe > ps ? a.foff[a.poff[pe]] : 0u;
    // the tile's blocks of deferred clears (CB_W x CB_H pixels each)
    constexpr int TBW = TILE_W / CB_W, TBH = TILE_H / CB_H;
    static_assert(TILE_W % CB_W == 0 && TILE_H % CB_H == 0, "tiles are whole clear blocks");
    const int bx0 = cx0 / CB_W, by0 = cy0 / CB_H;
    const int nbx = (cw + CB_W - 1) / CB_W, nby = (chh + CB_H - 1) / CB_H;
    if (f1 <= f0) {
        // no fragment: a pending clear stays deferred, marked on the blocks
        const uint8_t bits = (a.clear_color_pending ? CB_COLOR : 0) | (a.clear_depth_pending ? CB_DEPTH : 0);
        if (bits && tid < TBW * TBH) {
            const int bx = tid % TBW, by = tid / TBW;
            if (bx < nbx && by < nby) a.cblk[(by0 + by) * a.cbw + bx0 + bx] |= bits;
        }
        return;
    }
    {
        // the tile's current color and depth: the pending clear value, the
        // clear value of a block still marked, or memory; loads are
        // unconditional (clamped to the tile origin outside the framebuffer)
        // so they issue back to back
        constexpr int PPT = TILE_PIX / TILE_THREADS;
        uint32_t cv[PPT];
        double dv[PPT];
        uint8_t mk[PPT];
        const int64_t g0 = (int64_t)cy0 * a.W + cx0;
        const bool both_pending = a.clear_color_pending && a.clear_depth_pending;
#pragma unroll
        for (int k = 0; k < PPT; k++) {
            const int p = k * TILE_THREADS + tid;
            const int lx = p % TILE_W, ly = p / TILE_W;
            const bool in = lx < cw && ly < chh;
            mk[k] = in && !both_pending ? a.cblk[(by0 + ly / CB_H) * a.cbw + bx0 + lx / CB_W] : (uint8_t)0;
        }
#pragma unroll
        for (int k = 0; k < PPT; k++) {
            const int p = k * TILE_THREADS + tid;
            const int lx = p % TILE_W, ly = p / TILE_W;
            const int64_t gi = (lx < cw && ly < chh) ? (int64_t)(cy0 + ly) * a.W + cx0 + lx : g0;
            cv[k] = a.clear_color;
            dv[k] = a.clear_depth;
            if (!a.clear_color_pending && !(mk[k] & CB_COLOR)) cv[k] = a.color[gi];
            if (!a.clear_depth_pending && !(mk[k] & CB_DEPTH)) dv[k] = a.depth[gi];
        }
#pragma unroll
        for (int k = 0; k < PPT; k++) {
            const int p = k * TILE_THREADS + tid;
            col[p] = cv[k];
            dep[p] = dv[k];
            owner[p] = OWNER_NONE;
        }
    }
    const double bias = a.st.depth_bias;
    const bool rd = a.st.read_depth != 0, wd = a.st.write_depth != 0, wc = a.st.write_color != 0;
    const bool ab = a.st.alpha_blend != 0;
    uint32_t my_updated = 0;
    __syncthreads();

    for (uint32_t fc = f0; fc < f1; fc += ROP_CHUNK) {
        Frag fr[ROP_PER_THREAD];
        uint32_t pending = 0;
        // unconditional loads (clamped to the tile's last fragment) so they
        // issue back to back; the pending mask drops the clamped copies
#pragma unroll
        for (int q = 0; q < ROP_PER_THREAD; q++) {
            const uint32_t f = fc + (uint32_t)(q * TILE_THREADS + tid);
            fr[q] = a.frags[f < f1 ? f : f1 - 1];
            if (f < f1) pending |= 1u << q;
        }
        while (true) {
#pragma unroll
            for (int q = 0; q < ROP_PER_THREAD; q++)
                if (pending & (1u << q)) atomicMin(&owner[fr[q].pix], fc + (uint32_t)(q * TILE_THREADS + tid));
            __syncthreads();
#pragma unroll
            for (int q = 0; q < ROP_PER_THREAD; q++) {
                if (!(pending & (1u << q))) continue;
                const uint32_t p = fr[q].pix;
                if (owner[p] != fc + (uint32_t)(q * TILE_THREADS + tid)) continue;
                // context.go:230-273 at this pixel, in order
                const double z = fr[q].z;
                const double bz = z + bias;
                const uint16_t fl = fr[q].flags;
                if (!(rd && bz > dep[p]) && !(fl & FRAG_DISCARD) && (bz <= dep[p] || !rd)) {
                    my_updated++;
                    if (wd) dep[p] = z;
                    if (wc) {
                        if (ab && (fl & FRAG_BLEND)) col[p] = blend_over(col[p], fr[q].color);
                        else if (fl & FRAG_INB) col[p] = fr[q].color;
                    }
                }
                pending &= ~(1u << q);
                owner[p] = OWNER_NONE;
            }
            if (!__syncthreads_or(pending != 0)) break;
        }
    }

    for (int p = tid; p < TILE_PIX; p += TILE_THREADS) {
        int lx = p % TILE_W, ly = p / TILE_W;
        if (lx < cw && ly < chh) {
            int64_t gi = (int64_t)(cy0 + ly) * a.W + cx0 + lx;
            a.color[gi] = col[p];
            a.depth[gi] = dep[p];
        }
    }
    // the whole tile is written: its blocks hold no deferred clear (every
    // thread read its marks before the first barrier of the fragment loop)
    if (tid < TBW * TBH) {
        const int bx = tid % TBW, by = tid / TBW;
        if (bx < nbx && by < nby) a.cblk[(by0 + by) * a.cbw + bx0 + bx] = 0;
    }
    // UpdatedPixels of this tile, into one of the partial sums
    uint32_t uu = my_updated;
    for (int off = 32; off > 0; off >>= 1) uu += __shfl_down(uu, off);
    if (lane == 0) upd[tid >> 6] = uu;
    __syncthreads();
    if (tid == 0) {
        uint32_t s = 0;
        for (int w = 0; w < TILE_THREADS / 64; w++) s += upd[w];
        if (s) atomicAdd(&a.ctr->upd[t % UPD_PARTS], s);
    }
}
FGL_DRAW_KERNELS(k_rop, rop_body, __launch_bounds__(TILE_THREADS, FGL_ROP_WAVES))


// ---------------------------------------------------------------------------

static unsigned span_grid(const DrawArgs &a) {
    const uint64_t nb = ((uint64_t)a.span_cap + SPAN_THREADS - 1) / SPAN_THREADS;  // workgroups of 4 waves
    return (unsigned)(nb < SPAN_GRID_MAX ? (nb ? nb : 1) : SPAN_GRID_MAX);
}
static unsigned frag_groups(const DrawArgs &a) {
    const uint64_t nb = ((uint64_t)a.span_cap + FRAG_BLOCK - 1) / FRAG_BLOCK;
    const unsigned groups = (unsigned)((nb + FRAG_GROUP - 1) / FRAG_GROUP);
    return groups ? groups : 1;
}

hipError_t launch_frag_scan(const DrawArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_frag_scan, dim3(frag_groups(a)), dim3(FRAG_GROUP), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_shade(const DrawArgs &a, hipStream_t s) {
    dim3 grid(span_grid(a)), block(SPAN_THREADS);
    switch (a.sh.kind) {
        case 1: hipLaunchKernelGGL(k_shade<1>, grid, block, 0, s, a); break;
        case 2: hipLaunchKernelGGL(k_shade<2>, grid, block, 0, s, a); break;
        default:
            if (a.sh.has_tex) hipLaunchKernelGGL(k_shade<4>, grid, block, 0, s, a);
            else hipLaunchKernelGGL(k_shade<3>, grid, block, 0, s, a);
            break;
    }
    return hipGetLastError();
}
hipError_t launch_tile(const DrawArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_rop, dim3(a.ntiles), dim3(TILE_THREADS), 0, s, a);
    return hipGetLastError();
}

static hipError_t launch_frag_scan_multi(const MultiArgs &m, uint32_t nb, const DrawArgs &, hipStream_t s) {
    hipLaunchKernelGGL(k_frag_scan_m, dim3(nb), dim3(FRAG_GROUP), 0, s, m);
    return hipGetLastError();
}
static hipError_t launch_shade_multi(const MultiArgs &m, uint32_t nb, const DrawArgs &rep, hipStream_t s) {
    dim3 grid(nb), block(SPAN_THREADS);
    switch (rep.sh.kind) {
        case 1: hipLaunchKernelGGL(k_shade_m<1>, grid, block, 0, s, m); break;
        case 2: hipLaunchKernelGGL(k_shade_m<2>, grid, block, 0, s, m); break;
        default:
            if (rep.sh.has_tex) hipLaunchKernelGGL(k_shade_m<4>, grid, block, 0, s, m);
            else hipLaunchKernelGGL(k_shade_m<3>, grid, block, 0, s, m);
            break;
    }
    return hipGetLastError();
}
static hipError_t launch_rop_multi(const MultiArgs &m, uint32_t nb, const DrawArgs &, hipStream_t s) {
    hipLaunchKernelGGL(k_rop_m, dim3(nb), dim3(TILE_THREADS), 0, s, m);
    return hipGetLastError();
}

uint32_t stage_blocks(int stage, const DrawArgs &a, bool many) {
    switch (stage) {
        case ST_ZERO: return blocks_zero(a);
        case ST_VS: return (uint32_t)a.nvsblocks;
        case ST_VS_DEFERRED: return blocks_vs_deferred(a, many);
        case ST_SCAN_TOP: return 1;
        case ST_EMIT: return (uint32_t)a.nvsblocks;
        case ST_RADIX_HIST: return (uint32_t)a.nrsblocks;
        case ST_RADIX_SCAN: return 256;
        case ST_RADIX_SCATTER: return (uint32_t)a.nrsblocks;
        case ST_SPAN_COUNT: return span_blocks(a);
        case ST_SPAN_SCAN: return 1;
        case ST_SPANS: return span_blocks(a);
        case ST_FRAG_SCAN: return frag_groups(a);
        case ST_SHADE: return span_grid(a);
        case ST_ROP: return (uint32_t)a.ntiles;
        default: return 0;
    }
}

hipError_t launch_multi(int stage, const MultiArgs &m, uint32_t nblocks, const DrawArgs &rep, hipStream_t s) {
    static const MultiLaunch table[ST_COUNT] = {
        launch_zero_multi,       launch_vs_multi,         launch_vs_deferred_multi, launch_scan_top_multi,
        launch_emit_multi,       launch_radix_hist_multi, launch_radix_scan_multi,  launch_radix_scatter_multi,
        launch_span_count_multi, launch_span_scan_multi,  launch_spans_multi,       launch_frag_scan_multi,
        launch_shade_multi,      launch_rop_multi,
    };
    if (stage < 0 || stage >= ST_COUNT) return hipErrorInvalidValue;
    if (nblocks == 0) return hipSuccess;
    return table[stage](m, nblocks, rep, s);
}

}  // namespace FGL_TILE_NS

// This shape's table (a host function: nothing of it is compiled for the
// device).
#define FGL_OPS2(w, h) tile_ops_##w##x##h
#define FGL_OPS1(w, h) FGL_OPS2(w, h)
const TileOps &FGL_OPS1(FGL_TILE_W, FGL_TILE_H)() {
    static const TileOps ops = {
        TILE_W,
        TILE_H,
        FGL_TILE_NS::launch_zero,
        FGL_TILE_NS::launch_vs,
        FGL_TILE_NS::launch_vs_deferred,
        FGL_TILE_NS::launch_scan_top,
        FGL_TILE_NS::launch_emit,
        FGL_TILE_NS::launch_radix_hist,
        FGL_TILE_NS::launch_radix_scan,
        FGL_TILE_NS::launch_radix_scatter,
        FGL_TILE_NS::launch_span_count,
        FGL_TILE_NS::launch_span_scan,
        FGL_TILE_NS::launch_spans,
        FGL_TILE_NS::launch_frag_scan,
        FGL_TILE_NS::launch_shade,
        FGL_TILE_NS::launch_tile,
        FGL_TILE_NS::stage_blocks,
        FGL_TILE_NS::launch_multi,
    };
    return ops;
}
#undef FGL_OPS1
#undef FGL_OPS2

}  // namespace fgl
