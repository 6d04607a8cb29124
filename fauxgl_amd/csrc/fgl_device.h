// fgl_device.h -- device structures and Go floating-point semantics for the
// gfx950 draw path.  Compiled with -ffp-contract=off so that every a*b+c is
// two rounded operations, as Go evaluates it on amd64; FP64 division and
// sqrt lower to the correctly rounded sequences.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define FGL_D __device__ __forceinline__

namespace fgl {

// ---------------------------------------------------------------------------
// Tiling constants.  A tile is the unit a workgroup owns in LDS.

#ifndef FGL_TILE_W
#define FGL_TILE_W 32
#endif
#ifndef FGL_TILE_H
#define FGL_TILE_H 32
#endif
constexpr int TILE_W = FGL_TILE_W;
constexpr int TILE_H = FGL_TILE_H;
// The draw kernels (fgl_vs, fgl_bin, fgl_spans, fgl_tile) are compiled once
// per tile shape, each compilation in its own namespace (t32x32, t64x32, ...);
// the host picks a shape per draw.
#define FGL_TNS2(w, h) t##w##x##h
#define FGL_TNS1(w, h) FGL_TNS2(w, h)
#define FGL_TILE_NS FGL_TNS1(FGL_TILE_W, FGL_TILE_H)
constexpr int TILE_PIX = TILE_W * TILE_H;
// Threads of the framebuffer-update workgroup (one per tile): a 128 x 64
// tile holds 128 KB of LDS, one workgroup per CU, so it gets 16 waves to hide
// its loads; small tiles share a CU among several workgroups of 4 waves.
constexpr int TILE_THREADS = TILE_PIX >= 8192 ? 1024 : (TILE_PIX >= 4096 ? 512 : 256);
// Raster triangles whose integer bounding box leaves [-BBOX_LIM, BBOX_LIM]
// are dropped: the reference only produces such boxes from NaN/Inf screen
// coordinates, where its loops never end or wrap int64.
constexpr int32_t BBOX_LIM = 1 << 26;

// ---------------------------------------------------------------------------
// Raster triangle record: one per call the reference makes to rasterize()
// (context.go:151): screen-space vertices after the cull swap, reciprocal
// clip W, where to fetch the interpolated attributes, and the integer box.

enum AttrKind : uint32_t { ATTR_TRI = 0, ATTR_SIDE = 1, ATTR_LINE = 2 };

struct PrimRec {
    double sx[3], sy[3], sz[3];  // s0, s1, s2 (context.go:339-341)
    double r[3];                 // 1 / v_k.Output.W (context.go:176-178)
    uint64_t attr;               // kind:2 | sel:6 (3 x 2-bit vertex selectors) | index:56
    int32_t x0, y0, x1, y1;      // int(floor(min)), int(ceil(max)) (context.go:155-160)
    double ra;                   // 1 / edge(s0, s1, s2) (context.go:175)
};
static_assert(sizeof(PrimRec) == 128, "PrimRec layout");

FGL_D uint64_t make_attr(uint32_t kind, uint32_t s0, uint32_t s1, uint32_t s2, uint64_t index) {
    return ((uint64_t)kind << 62) | ((uint64_t)(s0 | (s1 << 2) | (s2 << 4)) << 56) | index;
}
FGL_D uint32_t attr_kind(uint64_t a) { return (uint32_t)(a >> 62); }
FGL_D uint32_t attr_sel(uint64_t a, int k) { return (uint32_t)(a >> (56 + 2 * k)) & 3u; }
FGL_D uint64_t attr_index(uint64_t a) { return a & ((1ull << 56) - 1); }

// Side vertex: the attributes of a vertex produced by clipping (clipping.go:68-71).
constexpr int SIDE_VTX_DOUBLES = 12;  // position 3, normal 3, texture 2, color 4

// Reference bit of a raster triangle in the sorted bin lists: bit 31 set =
// side record index, clear = index of the input triangle's main record.
constexpr uint32_t REF_SIDE = 0x80000000u;

// ---------------------------------------------------------------------------
// Shader parameters (shader.go:11-59) as the kernels see them.

struct DevShader {
    int32_t kind;
    int32_t has_tex;
    int32_t tex_kind;
    int32_t tex_w, tex_h;
    int32_t obj_is_discard;  // ObjectColor == Discard
    const void *tex_pix;
    double M[16];
    double color[4];
    double light[3], cam[3];
    double obj[4], amb[4], dif[4], spc[4];
    double power;
    // math.Pow's case analysis of y = power, made once per draw
    // (math.Pow's Modf / fraction adjustment, shader.go:91): yi and yf
    // after the "yf > 0.5" step, pw_frac = the fractional part is nonzero,
    // pw_big = yi >= 2^63.
    uint64_t pw_yi;
    double pw_yf;
    int32_t pw_frac, pw_big;
    int32_t pw_niter;  // bit length of pw_yi (squaring steps)
    int32_t pw_pad;
    // For x in [pw_lo, pw_hi] every intermediate of the squaring loop, taken
    // without the frexp scaling, stays within 2^-1000 .. 2^1000.
    double pw_lo, pw_hi;
};

struct DevState {
    int32_t read_depth, write_depth, write_color, alpha_blend, wireframe;
    int32_t front_face, cull;
    int32_t early_z;  // early depth reject before shading is exact (see tile kernel)
    double line_width, depth_bias;
};

// Per-draw counters and flags kept in device memory.
struct DrawCounters {
    unsigned long long total;    // TotalPixels
    uint32_t side_rec;           // side raster records allocated
    uint32_t side_vtx;           // side vertices allocated
    uint32_t ndeferred;          // input triangles deferred to the clipping kernel
    uint32_t npairs;             // (tile, triangle) pairs emitted
    uint32_t overflow;           // bit 0 side records, 1 side vertices, 2 pairs, 4 spans
    uint32_t pairs_lo, pairs_hi; // unclamped pair count
    uint32_t nspans;             // coverage spans (clamped to capacity)
    uint32_t spans_lo, spans_hi; // unclamped span count
    uint32_t nfrags;             // fragments (clamped to capacity)
    uint32_t frags_lo, frags_hi; // unclamped fragment count
    uint32_t pad2;
    // UpdatedPixels in UPD_PARTS partial sums (tile t adds into t % UPD_PARTS,
    // spreading the atomics over as many addresses); the host adds them up
    uint32_t upd[64];
};
constexpr int UPD_PARTS = 64;

// Coverage span: the covered pixels of one row of one raster triangle in
// one tile, with the edge values the reference holds at the first of them.
struct Span {
    double w[3];   // w0, w1, w2 at the first covered pixel (context.go:208-213)
    uint32_t ref;  // raster triangle
    uint16_t pix;  // tile-local index of the first covered pixel
    uint16_t n;    // covered pixels | SPAN_INB
};
static_assert(sizeof(Span) == 32, "Span layout");
constexpr uint16_t SPAN_INB = 0x8000;  // row y and x in bounds: SetNRGBA writes (context.go:269)
constexpr uint16_t SPAN_N = 0x7fff;

// A shaded fragment, in submission order: what the framebuffer update of
// context.go:230-273 needs at one pixel.
struct Frag {
    double z;        // b0*s0.Z + b1*s1.Z + b2*s2.Z (context.go:230)
    uint32_t color;  // Color.NRGBA() of the shader's color, bytes R,G,B,A
    uint16_t pix;    // tile-local pixel index
    uint16_t flags;  // FRAG_*
};
static_assert(sizeof(Frag) == 16, "Frag layout");
constexpr uint16_t FRAG_INB = 1;      // SetNRGBA writes (x, y in bounds)
constexpr uint16_t FRAG_BLEND = 2;    // color.A < 1 (blended when AlphaBlend)
constexpr uint16_t FRAG_DISCARD = 4;  // color == Discard (context.go:241)

// ---------------------------------------------------------------------------
// Go semantics

// math.Min / math.Max (NaN and signed-zero rules).
FGL_D double go_min(double x, double y) {
    if ((__builtin_isinf(x) && x < 0) || (__builtin_isinf(y) && y < 0)) return -__builtin_inf();
    if (__builtin_isnan(x) || __builtin_isnan(y)) return __builtin_nan("");
    if (x == 0 && x == y) return __builtin_signbit(x) ? x : y;
    return x < y ? x : y;
}
FGL_D double go_max(double x, double y) {
    if ((__builtin_isinf(x) && x > 0) || (__builtin_isinf(y) && y > 0)) return __builtin_inf();
    if (__builtin_isnan(x) || __builtin_isnan(y)) return __builtin_nan("");
    if (x == 0 && x == y) return __builtin_signbit(x) ? y : x;
    return x > y ? x : y;
}
// math.Max(x, 0) and math.Min(x, 1) as selects: with a constant second
// operand the special cases reduce to one comparison (NaN passes through,
// -0 gives +0 for Max, +-Inf fall on the right side of the comparison).
FGL_D double go_max0(double x) { return !(x <= 0.0) ? x : 0.0; }
FGL_D double go_min1(double x) { return !(x >= 1.0) ? x : 1.0; }
// int(x) on amd64: truncation, NaN and out-of-range give INT64_MIN.
FGL_D int64_t go_f2i(double x) {
    if (!(x >= -9223372036854775808.0 && x < 9223372036854775808.0)) return INT64_MIN;
    return (int64_t)x;
}
// uint8(x) goes through a 32-bit conversion: out of range gives 0x80000000.
FGL_D uint32_t go_u8(double x) {
    if (!(x > -2147483649.0 && x < 2147483648.0)) return 0u;
    return ((uint32_t)(int32_t)x) & 0xffu;
}
FGL_D double clampd(double x, double lo, double hi) {
    if (x < lo) return lo;
    if (x > hi) return hi;
    return x;
}

FGL_D int go_is_odd_int(double x) {
    if (__builtin_fabs(x) >= 9007199254740992.0) return 0;
    double xi = __builtin_trunc(x);
    return (x - xi) == 0 && (((int64_t)xi) & 1) == 1;
}

// math.Pow, written from the method Go documents: special cases, then the
// integer part of y by binary exponentiation on the frexp mantissa (the
// exponent carried separately), the fractional part by exp(yf*log(x)).
__device__ double go_pow(double x, double y);

// ---------------------------------------------------------------------------
// Small vector helpers: each is one reference method, evaluated left to
// right exactly as written there.

struct V3 { double x, y, z; };
struct V4 { double x, y, z, w; };
struct C4 { double r, g, b, a; };

FGL_D V3 v3(double x, double y, double z) { V3 o; o.x = x; o.y = y; o.z = z; return o; }
FGL_D V3 vadd(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
FGL_D V3 vsub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
FGL_D V3 vmuls(V3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
FGL_D double vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
FGL_D V3 vcross(V3 a, V3 b) {
    double x = a.y * b.z - a.z * b.y;
    double y = a.z * b.x - a.x * b.z;
    double z = a.x * b.y - a.y * b.x;
    return v3(x, y, z);
}
FGL_D V3 vnormalize(V3 a) {  // vector.go:83-86
    double r = 1.0 / __builtin_sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return v3(a.x * r, a.y * r, a.z * r);
}
FGL_D V4 v4(double x, double y, double z, double w) { V4 o; o.x = x; o.y = y; o.z = z; o.w = w; return o; }
FGL_D V4 wadd(V4 a, V4 b) { return v4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
FGL_D V4 wsub(V4 a, V4 b) { return v4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
FGL_D V4 wmuls(V4 a, double s) { return v4(a.x * s, a.y * s, a.z * s, a.w * s); }
FGL_D double wdot(V4 a, V4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
FGL_D bool woutside(V4 a) {  // vector.go:212-215
    double x = a.x, y = a.y, z = a.z, w = a.w;
    return x < -w || x > w || y < -w || y > w || z < -w || z > w;
}
FGL_D C4 c4(double r, double g, double b, double a) { C4 o; o.r = r; o.g = g; o.b = b; o.a = a; return o; }

// Matrix.MulPositionW / MulPosition (matrix.go:209-222): ((a0*x + a1*y) + a2*z) + a3.
FGL_D V4 mul_position_w(const double *M, V3 p) {
    double x = M[0] * p.x + M[1] * p.y + M[2] * p.z + M[3];
    double y = M[4] * p.x + M[5] * p.y + M[6] * p.z + M[7];
    double z = M[8] * p.x + M[9] * p.y + M[10] * p.z + M[11];
    double w = M[12] * p.x + M[13] * p.y + M[14] * p.z + M[15];
    return v4(x, y, z, w);
}
FGL_D V3 mul_position(const double *M, V3 p) {
    double x = M[0] * p.x + M[1] * p.y + M[2] * p.z + M[3];
    double y = M[4] * p.x + M[5] * p.y + M[6] * p.z + M[7];
    double z = M[8] * p.x + M[9] * p.y + M[10] * p.z + M[11];
    return v3(x, y, z);
}

// Color.NRGBA (color.go:56-63) packed as bytes R,G,B,A (little endian u32).
FGL_D uint32_t nrgba_pack(C4 c) {
    const double d = 255.0;
    uint32_t r = go_u8(clampd(c.r, 0, 1) * d);
    uint32_t g = go_u8(clampd(c.g, 0, 1) * d);
    uint32_t b = go_u8(clampd(c.b, 0, 1) * d);
    uint32_t a = go_u8(clampd(c.a, 0, 1) * d);
    return r | (g << 8) | (b << 16) | (a << 24);
}

// Integer "Over" of context.go:257-267 with color.NRGBA.RGBA()'s 16-bit
// premultiply: src is the fragment's NRGBA, dst the buffer's NRGBA bytes.
FGL_D uint32_t blend_over(uint32_t dst, uint32_t src) {
    uint32_t A = src >> 24;
    uint32_t sa = A | (A << 8);
    uint32_t a = (0xffffu - sa) * 0x101u;
    uint32_t out = 0;
#pragma unroll
    for (int ch = 0; ch < 3; ch++) {
        uint32_t v = (src >> (8 * ch)) & 0xffu;
        uint32_t s = ((v | (v << 8)) * A) / 0xffu;
        uint32_t dv = (dst >> (8 * ch)) & 0xffu;
        out |= (((dv * a / 0xffffu + s) >> 8) & 0xffu) << (8 * ch);
    }
    uint32_t dv = dst >> 24;
    out |= (((dv * a / 0xffffu + sa) >> 8) & 0xffu) << 24;
    return out;
}

}  // namespace fgl
