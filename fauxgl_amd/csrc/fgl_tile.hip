// fgl_tile.hip -- the per-tile raster kernel: one workgroup owns one
// TILE_W x TILE_H tile of the color and depth buffers in LDS and draws the
// tile's raster triangles in submission order.  Restates rasterize()
// (context.go:151-281), InterpolateVertexes (vertex.go:18-79), the built-in
// fragment shaders (shader.go:25-96), BilinearSample (texture.go:41-63) and
// the framebuffer write (context.go:240-273).
//
// Triangles are consumed in batches.  For each batch:
//   A. coverage: one lane per (triangle, row) item replays the reference's
//      forward-differenced edge recurrence for that row (the row recurrence
//      w00 += b12 from y0, the skip d, then w += a per pixel), so every
//      barycentric is bit-identical to the serial loop; covered pixels are
//      appended to an LDS fragment buffer;
//   B. shading: one lane per fragment computes z, the perspective-correct
//      interpolants and the shader color;
//   C. ordered write: fragments are applied in rounds; in each round the
//      lowest (triangle, row) key per pixel wins, so writes to a pixel
//      happen in submission order exactly as in the reference.
// The early depth reject in B is used only when it is provably equal to the
// reference's test at write time (depth non-increasing: ReadDepth and
// DepthBias >= 0, or WriteDepth off).
#include "fgl_internal.h"
#include "fgl_tiles.h"

namespace fgl {

#ifndef FGL_FRAG_CAP
#define FGL_FRAG_CAP 1024
#endif
#ifndef FGL_MAX_ITEMS
#define FGL_MAX_ITEMS 512
#endif
constexpr int FRAG_CAP = FGL_FRAG_CAP;
constexpr int MAX_ITEMS = FGL_MAX_ITEMS;

// Phase timers for profiling builds (-DFGL_TILE_PHASE_TIMERS): thread 0 of
// each workgroup accumulates shader-clock cycles per phase into the draw
// counters.  They add no work to the default build.
#ifdef FGL_TILE_PHASE_TIMERS
#define PT_DECL uint64_t pt_last = 0, pt_acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#define PT_START if (tid == 0) pt_last = clock64();
#define PT_MARK(k)                           \
    if (tid == 0) {                          \
        uint64_t now_ = clock64();           \
        pt_acc[k] += now_ - pt_last;         \
        pt_last = now_;                      \
    }
#define PT_FLUSH \
    if (tid == 0) for (int k_ = 0; k_ < 8; k_++) atomicAdd(&a.ctr->phase[k_], (unsigned long long)pt_acc[k_]);
#else
#define PT_DECL
#define PT_START
#define PT_MARK(k)
#define PT_FLUSH
#endif
constexpr int BATCH = TILE_THREADS;
constexpr uint32_t OWNER_NONE = 0xffffffffu;

constexpr uint16_t F_PIX = 0x0fff;
constexpr uint16_t F_INB = 0x1000;    // x, y in bounds (SetNRGBA writes)
constexpr uint16_t F_BLEND = 0x2000;  // AlphaBlend && color.A < 1
constexpr uint16_t F_DEAD = 0x4000;   // discarded or rejected early
constexpr uint16_t F_DONE = 0x8000;   // applied

struct TileLDS {
    uint32_t col[TILE_PIX];
    double dep[TILE_PIX];
    uint32_t owner[TILE_PIX];
    double fb0[FRAG_CAP], fb1[FRAG_CAP], fb2[FRAG_CAP];
    uint32_t fkey[FRAG_CAP];
    uint32_t fcol[FRAG_CAP];
    uint16_t fpix[FRAG_CAP];
    uint32_t pref[BATCH];
    uint32_t istart[BATCH + 1];
    uint8_t iprim[MAX_ITEMS];
    uint32_t wsum[2][TILE_THREADS / 64];
    uint32_t nfrag;
    uint32_t nitems;
    uint32_t next_lp, next_io;
    unsigned long long total, updated;
};

// ---------------------------------------------------------------------------
// math.Pow (see fgl_device.h)

__device__ double go_pow(double x, double y) {
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
    double ay = __builtin_fabs(y);
    double yi = __builtin_trunc(ay);
    double yf = ay - yi;
    if (yf != 0 && x < 0) return __builtin_nan("");
    if (yi >= 9223372036854775808.0) {
        if (x == -1) return 1;
        if ((__builtin_fabs(x) < 1) == (y > 0)) return 0;
        return __builtin_inf();
    }
    double a1 = 1.0;
    int64_t ae = 0;
    if (yf != 0) {
        if (yf > 0.5) {
            yf = yf - 1;
            yi = yi + 1;
        }
        a1 = exp(yf * log(x));
    }
    int xe_i;
    double x1 = frexp(x, &xe_i);
    int64_t xe = xe_i;
    for (int64_t i = (int64_t)yi; i != 0; i >>= 1) {
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
    if (y < 0) {
        a1 = 1.0 / a1;
        ae = -ae;
    }
    if (ae > 100000) ae = 100000;
    if (ae < -100000) ae = -100000;
    return ldexp(a1, (int)ae);
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

struct AttrSrc {
    const double *pos, *nrm, *uv, *col;  // per-vertex arrays (stride 3, 3, 2, 4)
    const double *side;                  // side vertices (stride SIDE_VTX_DOUBLES)
};

FGL_D void vtx_addr(const AttrSrc &src, uint64_t attr, int k, const double *&base, int64_t &vid, bool &side) {
    uint32_t kind = attr_kind(attr);
    uint64_t idx = attr_index(attr);
    uint32_t sel = attr_sel(attr, k);
    side = kind == ATTR_SIDE;
    vid = kind == ATTR_TRI ? (int64_t)(idx * 3 + sel) : kind == ATTR_LINE ? (int64_t)(idx * 2 + sel) : (int64_t)(idx + sel);
    base = side ? src.side + vid * SIDE_VTX_DOUBLES : nullptr;
}

FGL_D V3 fetch_pos(const AttrSrc &src, uint64_t attr, int k) {
    const double *b; int64_t vid; bool side;
    vtx_addr(src, attr, k, b, vid, side);
    if (side) return v3(b[0], b[1], b[2]);
    if (!src.pos) return v3(0, 0, 0);
    return v3(src.pos[3 * vid], src.pos[3 * vid + 1], src.pos[3 * vid + 2]);
}
FGL_D V3 fetch_nrm(const AttrSrc &src, uint64_t attr, int k) {
    const double *b; int64_t vid; bool side;
    vtx_addr(src, attr, k, b, vid, side);
    if (side) return v3(b[3], b[4], b[5]);
    if (!src.nrm) return v3(0, 0, 0);
    return v3(src.nrm[3 * vid], src.nrm[3 * vid + 1], src.nrm[3 * vid + 2]);
}
FGL_D void fetch_uv(const AttrSrc &src, uint64_t attr, int k, double &u, double &v) {
    const double *b; int64_t vid; bool side;
    vtx_addr(src, attr, k, b, vid, side);
    if (side) { u = b[6]; v = b[7]; return; }
    if (!src.uv) { u = 0; v = 0; return; }
    u = src.uv[2 * vid];
    v = src.uv[2 * vid + 1];
}
FGL_D C4 fetch_col(const AttrSrc &src, uint64_t attr, int k) {
    const double *b; int64_t vid; bool side;
    vtx_addr(src, attr, k, b, vid, side);
    if (side) return c4(b[8], b[9], b[10], b[11]);
    if (!src.col) return c4(0, 0, 0, 0);
    return c4(src.col[4 * vid], src.col[4 * vid + 1], src.col[4 * vid + 2], src.col[4 * vid + 3]);
}

// InterpolateFloats-style sum: ((((0 + a*bx) + b*by) + c*bz) * bw)
FGL_D double interp(double a, double b, double c, V4 k) {
    double n = 0;
    n = n + a * k.x;
    n = n + b * k.y;
    n = n + c * k.z;
    return n * k.w;
}

// Fragment shaders (shader.go).  Returns the shader's Color.
template <int SH>
FGL_D C4 shade(const DevShader &sh, const AttrSrc &src, uint64_t attr, V4 b) {
    if (SH == 1) {  // SolidColorShader.Fragment (shader.go:25-27)
        return c4(sh.color[0], sh.color[1], sh.color[2], sh.color[3]);
    } else if (SH == 2) {  // TextureShader.Fragment (shader.go:44-46)
        double u0, v0, u1, v1, u2, v2;
        fetch_uv(src, attr, 0, u0, v0);
        fetch_uv(src, attr, 1, u1, v1);
        fetch_uv(src, attr, 2, u2, v2);
        return bilinear_sample(sh, interp(u0, u1, u2, b), interp(v0, v1, v2, b));
    } else {  // PhongShader.Fragment (shader.go:75-96)
        V3 p0 = fetch_pos(src, attr, 0), p1 = fetch_pos(src, attr, 1), p2 = fetch_pos(src, attr, 2);
        V3 n0 = fetch_nrm(src, attr, 0), n1 = fetch_nrm(src, attr, 1), n2 = fetch_nrm(src, attr, 2);
        V3 P = v3(interp(p0.x, p1.x, p2.x, b), interp(p0.y, p1.y, p2.y, b), interp(p0.z, p1.z, p2.z, b));
        V3 N = vnormalize(v3(interp(n0.x, n1.x, n2.x, b), interp(n0.y, n1.y, n2.y, b), interp(n0.z, n1.z, n2.z, b)));
        C4 light = c4(sh.amb[0], sh.amb[1], sh.amb[2], sh.amb[3]);
        C4 color;
        if (sh.has_tex) {
            double u0, v0, u1, v1, u2, v2;
            fetch_uv(src, attr, 0, u0, v0);
            fetch_uv(src, attr, 1, u1, v1);
            fetch_uv(src, attr, 2, u2, v2);
            color = bilinear_sample(sh, interp(u0, u1, u2, b), interp(v0, v1, v2, b));
        } else if (!sh.obj_is_discard) {
            color = c4(sh.obj[0], sh.obj[1], sh.obj[2], sh.obj[3]);
        } else {
            C4 c0 = fetch_col(src, attr, 0), c1 = fetch_col(src, attr, 1), c2 = fetch_col(src, attr, 2);
            color = c4(interp(c0.r, c1.r, c2.r, b), interp(c0.g, c1.g, c2.g, b), interp(c0.b, c1.b, c2.b, b),
                       interp(c0.a, c1.a, c2.a, b));
        }
        V3 L = v3(sh.light[0], sh.light[1], sh.light[2]);
        double diffuse = go_max(vdot(N, L), 0);
        light = cadd_scaled(light, c4(sh.dif[0], sh.dif[1], sh.dif[2], sh.dif[3]), diffuse);
        if (diffuse > 0 && sh.power > 0) {
            V3 camera = vnormalize(vsub(v3(sh.cam[0], sh.cam[1], sh.cam[2]), P));
            V3 i = v3(-L.x, -L.y, -L.z);
            V3 reflected = vsub(i, vmuls(N, 2 * vdot(N, i)));
            double specular = go_max(vdot(camera, reflected), 0);
            if (specular > 0) {
                specular = go_pow(specular, sh.power);
                light = cadd_scaled(light, c4(sh.spc[0], sh.spc[1], sh.spc[2], sh.spc[3]), specular);
            }
        }
        return c4(go_min(color.r * light.r, 1), go_min(color.g * light.g, 1), go_min(color.b * light.b, 1),
                  color.a);
    }
}

// ---------------------------------------------------------------------------

FGL_D const PrimRec *prim_of(const DrawArgs &a, uint32_t ref) {
    return (ref & REF_SIDE) ? a.side_rec + (ref & ~REF_SIDE) : a.rec_main + ref;
}

// edge(a, b, c) (context.go:147-149)
FGL_D double edge(double ax, double ay, double bx, double by, double cx, double cy) {
    return (bx - cx) * (ay - cy) - (by - cy) * (ax - cx);
}

// Appends a fragment to the tile's buffer (one LDS atomic per wave).
FGL_D void emit_frag(TileLDS &s, bool pred, uint16_t pix, uint32_t key, double b0, double b1, double b2) {
    uint64_t m = __ballot(pred);
    if (pred) {
        uint32_t lane = __lane_id();
        uint32_t rank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        int leader = __ffsll((long long)m) - 1;
        uint32_t base = 0;
        if ((int)lane == leader) base = atomicAdd(&s.nfrag, (uint32_t)__popcll(m));
        base = __shfl(base, leader);
        uint32_t slot = base + rank;
        if (slot < FRAG_CAP) {
            s.fpix[slot] = pix;
            s.fkey[slot] = key;
            s.fb0[slot] = b0;
            s.fb1[slot] = b1;
            s.fb2[slot] = b2;
        }
    }
}

template <int SH>
__global__ __launch_bounds__(TILE_THREADS) void k_tile(DrawArgs a) {
    __shared__ TileLDS s;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int t = blockIdx.x;
    if (a.ctr->overflow) return;  // the host reruns the whole draw
    const int tx = t % a.ntx, ty = t / a.ntx;
    const int cx0 = tx * TILE_W, cy0 = ty * TILE_H;
    const int cw = min(TILE_W, a.W - cx0), chh = min(TILE_H, a.H - cy0);
    const int cx1 = cx0 + cw - 1, cy1 = cy0 + chh - 1;
    PT_DECL
    PT_START
    const uint32_t ls = a.tstart[t], le = a.tend[t];
    const bool has = le > ls;
    if (!has) {
        if (a.clear_color_pending || a.clear_depth_pending) {
            for (int p = tid; p < TILE_PIX; p += TILE_THREADS) {
                int lx = p % TILE_W, ly = p / TILE_W;
                if (lx < cw && ly < chh) {
                    int64_t gi = (int64_t)(cy0 + ly) * a.W + cx0 + lx;
                    if (a.clear_color_pending) a.color[gi] = a.clear_color;
                    if (a.clear_depth_pending) a.depth[gi] = a.clear_depth;
                }
            }
        }
        return;
    }

    for (int p = tid; p < TILE_PIX; p += TILE_THREADS) {
        int lx = p % TILE_W, ly = p / TILE_W;
        if (lx < cw && ly < chh) {
            int64_t gi = (int64_t)(cy0 + ly) * a.W + cx0 + lx;
            s.col[p] = a.clear_color_pending ? a.clear_color : a.color[gi];
            s.dep[p] = a.clear_depth_pending ? a.clear_depth : a.depth[gi];
        }
        s.owner[p] = OWNER_NONE;
    }
    if (tid == 0) {
        s.total = 0;
        s.updated = 0;
    }

    AttrSrc src;
    src.pos = a.in.pos;
    src.nrm = a.in.nrm;
    src.uv = a.in.uv;
    src.col = a.in.col;
    src.side = a.side_vtx;
    const uint32_t *vals = a.vals[a.sorted_buf];
    const double bias = a.st.depth_bias;
    uint32_t my_total = 0, my_updated = 0;

    uint32_t lp = ls, io = 0;
    __syncthreads();
    PT_MARK(0)
    while (lp < le) {
        // ---- batch formation: items of triangles lp, lp+1, ... that fit ----
        uint32_t idx = lp + (uint32_t)tid;
        bool valid = idx < le;
        uint32_t ref = 0, ni = 0, wb = 0;
        if (valid) {
            ref = vals[idx];
            const PrimRec *pr = prim_of(a, ref);
            ItemGeo g = item_geo(pr->x0, pr->y0, pr->x1, pr->y1, a.W, cx0, cx1, cy0, cy1);
            uint64_t n64 = (uint64_t)g.nY * (uint64_t)g.nk;
            if (tid == 0) n64 -= io;
            ni = n64 > (1u << 20) ? (1u << 20) : (uint32_t)n64;
            wb = (uint32_t)g.wb;
        }
        uint32_t bnd = ni * wb;
        // inclusive scans of bnd and ni over the block
        uint32_t sb = bnd, si = ni;
        for (int off = 1; off < 64; off <<= 1) {
            uint32_t yb = __shfl_up(sb, off), yi = __shfl_up(si, off);
            if (lane >= off) {
                sb += yb;
                si += yi;
            }
        }
        if (lane == 63) {
            s.wsum[0][wave] = sb;
            s.wsum[1][wave] = si;
        }
        __syncthreads();
        for (int w = 0; w < wave; w++) {
            sb += s.wsum[0][w];
            si += s.wsum[1][w];
        }
        uint32_t be = sb - bnd, ie = si - ni;  // exclusive
        bool full = valid && sb <= (uint32_t)FRAG_CAP && si <= (uint32_t)MAX_ITEMS;
        int nfull = __syncthreads_count(full);
        uint32_t take = full ? ni : 0u;
        if (tid == nfull && valid) {
            uint32_t byb = wb ? ((uint32_t)FRAG_CAP - be) / wb : ni;
            uint32_t byi = (uint32_t)MAX_ITEMS - ie;
            take = byb < byi ? byb : byi;
            if (take > ni) take = ni;
            s.nitems = ie + take;
            if (nfull == 0) {
                s.next_lp = lp;
                s.next_io = io + take;
            } else {
                s.next_lp = lp + (uint32_t)nfull;
                s.next_io = take;
            }
        }
        if (tid == nfull - 1 && (nfull == TILE_THREADS || lp + (uint32_t)nfull >= le)) {
            s.nitems = si;
            s.next_lp = lp + (uint32_t)nfull;
            s.next_io = 0;
        }
        if (tid <= nfull && valid) {
            s.pref[tid] = ref;
            s.istart[tid] = ie;
            for (uint32_t m = 0; m < take; m++) s.iprim[ie + m] = (uint8_t)tid;
        }
        if (tid == 0) s.nfrag = 0;
        __syncthreads();
        const uint32_t nitems = s.nitems;
        const uint32_t io0 = io;
        PT_MARK(1)

        // ---- A: coverage walk per (triangle, row) item ----
        for (uint32_t m = (uint32_t)tid; m < nitems; m += TILE_THREADS) {
            const bool act = true;
            uint32_t j = act ? s.iprim[m] : 0u;
            int32_t y = 0, k = 0, xa = 1, xb = 0, x0 = 0, x1 = 0, y0 = 0, ya = 0;
            double w00 = 0, w01 = 0, w02 = 0, a12 = 0, a20 = 0, a01 = 0, b12 = 0, b20 = 0, b01 = 0;
            double ra = 0, ra12 = 0, ra20 = 0, ra01 = 0;
            if (act) {
                uint32_t local = m - s.istart[j] + (j == 0 ? io0 : 0u);
                const PrimRec *pr = prim_of(a, s.pref[j]);
                x0 = pr->x0; y0 = pr->y0; x1 = pr->x1;
                ItemGeo g = item_geo(x0, y0, x1, pr->y1, a.W, cx0, cx1, cy0, cy1);
                ya = g.ya;
                y = g.ya + (int32_t)(local / (uint32_t)g.nk);
                k = g.kmin + (int32_t)(local % (uint32_t)g.nk);
                int32_t py = y + k;
                int32_t kw = k * a.W;
                xa = x0 > kw + cx0 ? x0 : kw + cx0;
                xb = x1 < kw + cx1 ? x1 : kw + cx1;
                if (py < cy0 || py > cy1) xb = xa - 1;
                if (xa <= xb) {
                    double s0x = pr->sx[0], s0y = pr->sy[0];
                    double s1x = pr->sx[1], s1y = pr->sy[1];
                    double s2x = pr->sx[2], s2y = pr->sy[2];
                    // forward differencing variables (context.go:162-181)
                    double px = (double)x0 + 0.5, pyy = (double)y0 + 0.5;
                    w00 = edge(s1x, s1y, s2x, s2y, px, pyy);
                    w01 = edge(s2x, s2y, s0x, s0y, px, pyy);
                    w02 = edge(s0x, s0y, s1x, s1y, px, pyy);
                    a01 = s1y - s0y;
                    b01 = s0x - s1x;
                    a12 = s2y - s1y;
                    b12 = s1x - s2x;
                    a20 = s0y - s2y;
                    b20 = s2x - s0x;
                    ra = 1.0 / edge(s0x, s0y, s1x, s1y, s2x, s2y);
                    ra12 = 1.0 / a12;
                    ra20 = 1.0 / a20;
                    ra01 = 1.0 / a01;
                    // rows y0 .. y-1 of the row recurrence (context.go:275-277)
                    for (int32_t yy = y0; yy < y; yy++) {
                        w00 += b12;
                        w01 += b20;
                        w02 += b01;
                    }
                }
            }
            bool walking = act && xa <= xb;
            double w0 = 0, w1 = 0, w2 = 0;
            int32_t x = 0;
            if (walking) {
                // skip to the first pixel that can be inside (context.go:185-205)
                double d = 0;
                double d0 = -w00 * ra12;
                double d1 = -w01 * ra20;
                double d2 = -w02 * ra01;
                if (w00 < 0 && d0 > d) d = d0;
                if (w01 < 0 && d1 > d) d = d1;
                if (w02 < 0 && d2 > d) d = d2;
                d = (double)go_f2i(d);
                if (d < 0) d = 0;
                w0 = w00 + a12 * d;
                w1 = w01 + a20 * d;
                w2 = w02 + a01 * d;
                int64_t id = go_f2i(d);
                x = (id > (int64_t)x1 - x0) ? x1 + 1 : x0 + (int32_t)id;
            }
            const uint32_t keyb = (j << 23) | (uint32_t)((y - ya) & 0x7fffff);
            const int32_t kw = k * a.W;
            const uint16_t inb = k == 0 ? F_INB : (uint16_t)0;
            bool was_inside = false;
            // pixel loop (context.go:207-221); one ballot per step
            while (__any(walking)) {
                bool em = false;
                double b0 = 0, b1 = 0, b2 = 0;
                if (walking) {
                    if (x > xb) {
                        walking = false;
                    } else {
                        b0 = w0 * ra;
                        b1 = w1 * ra;
                        b2 = w2 * ra;
                        w0 += a12;
                        w1 += a20;
                        w2 += a01;
                        bool outside = b0 < 0 || b1 < 0 || b2 < 0;
                        if (outside) {
                            if (was_inside) walking = false;
                        } else {
                            was_inside = true;
                            em = x >= xa;
                        }
                    }
                }
                uint16_t pix = 0;
                if (em) pix = (uint16_t)((y + k - cy0) * TILE_W + (x - kw - cx0));
                emit_frag(s, em, (uint16_t)(pix | inb), keyb, b0, b1, b2);
                my_total += em ? 1u : 0u;
                x++;
            }
        }
        __syncthreads();
        PT_MARK(2)

        // ---- B: shading ----
        uint32_t nf = s.nfrag;
        if (nf > (uint32_t)FRAG_CAP) {  // cannot happen: the batch bound caps it
            if (tid == 0) atomicOr(&a.ctr->overflow, 8u);
            nf = FRAG_CAP;
        }
        for (uint32_t f = (uint32_t)tid; f < nf; f += TILE_THREADS) {
            uint32_t key = s.fkey[f];
            uint16_t fp = s.fpix[f];
            const PrimRec *pr = prim_of(a, s.pref[key >> 23]);
            double b0 = s.fb0[f], b1 = s.fb1[f], b2 = s.fb2[f];
            double z = b0 * pr->sz[0] + b1 * pr->sz[1] + b2 * pr->sz[2];
            double bz = z + bias;
            if (a.st.early_z && bz > s.dep[fp & F_PIX]) {
                s.fpix[f] = fp | F_DEAD;
                continue;
            }
            // perspective-correct interpolation (context.go:236-238)
            V4 bw = v4(b0 * pr->r[0], b1 * pr->r[1], b2 * pr->r[2], 0);
            bw.w = 1.0 / (bw.x + bw.y + bw.z);
            C4 c = shade<SH>(a.sh, src, pr->attr, bw);
            if (c.r == 0 && c.g == 0 && c.b == 0 && c.a == 0) {  // Discard (context.go:241)
                s.fpix[f] = fp | F_DEAD;
                continue;
            }
            s.fb0[f] = z;
            s.fb1[f] = bz;
            s.fcol[f] = nrgba_pack(c);
            if (a.st.alpha_blend && c.a < 1) fp |= F_BLEND;
            s.fpix[f] = fp;
        }
        __syncthreads();
        PT_MARK(3)

        // ---- C: ordered write, lowest key per pixel first ----
        const bool rd = a.st.read_depth != 0, wd = a.st.write_depth != 0, wc = a.st.write_color != 0;
        while (true) {
            for (uint32_t f = (uint32_t)tid; f < nf; f += TILE_THREADS) {
                uint16_t fp = s.fpix[f];
                if (!(fp & (F_DEAD | F_DONE))) atomicMin(&s.owner[fp & F_PIX], s.fkey[f]);
            }
            __syncthreads();
            int left = 0;
            for (uint32_t f = (uint32_t)tid; f < nf; f += TILE_THREADS) {
                uint16_t fp = s.fpix[f];
                if (fp & (F_DEAD | F_DONE)) continue;
                uint32_t p = fp & F_PIX;
                uint32_t key = s.fkey[f];
                if (s.owner[p] != key) {
                    left = 1;
                    continue;
                }
                double z = s.fb0[f], bz = s.fb1[f];
                if (bz <= s.dep[p] || !rd) {
                    my_updated++;
                    if (wd) s.dep[p] = z;
                    if (wc) {
                        if (fp & F_BLEND) s.col[p] = blend_over(s.col[p], s.fcol[f]);
                        else if (fp & F_INB) s.col[p] = s.fcol[f];
                    }
                }
                s.fpix[f] = fp | F_DONE;
                s.owner[p] = OWNER_NONE;
            }
            if (!__syncthreads_or(left)) break;
        }

        lp = s.next_lp;
        io = s.next_io;
        __syncthreads();
        PT_MARK(4)
    }

    // write the tile back
    for (int p = tid; p < TILE_PIX; p += TILE_THREADS) {
        int lx = p % TILE_W, ly = p / TILE_W;
        if (lx < cw && ly < chh) {
            int64_t gi = (int64_t)(cy0 + ly) * a.W + cx0 + lx;
            a.color[gi] = s.col[p];
            a.depth[gi] = s.dep[p];
        }
    }
    atomicAdd(&s.total, (unsigned long long)my_total);
    atomicAdd(&s.updated, (unsigned long long)my_updated);
    __syncthreads();
    PT_MARK(5)
    PT_FLUSH
    if (tid == 0) {
        atomicAdd(&a.ctr->total, s.total);
        atomicAdd(&a.ctr->updated, s.updated);
    }
}

hipError_t launch_tile(const DrawArgs &a, hipStream_t s) {
    dim3 grid(a.ntiles), block(TILE_THREADS);
    switch (a.sh.kind) {
        case 1: hipLaunchKernelGGL(k_tile<1>, grid, block, 0, s, a); break;
        case 2: hipLaunchKernelGGL(k_tile<2>, grid, block, 0, s, a); break;
        default: hipLaunchKernelGGL(k_tile<3>, grid, block, 0, s, a); break;
    }
    return hipGetLastError();
}

}  // namespace fgl
