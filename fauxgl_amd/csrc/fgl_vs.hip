// fgl_vs.hip -- vertex stage: vertex shader, view-volume clipping, culling,
// screen mapping and raster-triangle records, one thread per input
// triangle or line.  Restates DrawTriangle / DrawLine (context.go:351-389),
// ClipTriangle / ClipLine (clipping.go:54-98), drawClippedTriangle /
// drawClippedLine (context.go:303-349) and line / wireframe
// (context.go:283-301).  Every expression keeps the reference's operation
// order; the file is compiled with -ffp-contract=off.
#include "fgl_internal.h"
#include "fgl_tiles.h"
#include "fgl_draws.h"

namespace fgl {
namespace FGL_TILE_NS {

// A vertex as the pipeline carries it (vertex.go:3-12; Texture.Z dropped,
// no built-in shader reads it).
struct Vtx {
    V3 pos, nrm;
    double u, v;
    C4 col;
    V4 out;
};

FGL_D Vtx load_vertex(const DevPrims &in, int64_t vid) {
    Vtx o;
    o.pos = in.pos ? v3(in.pos[3 * vid], in.pos[3 * vid + 1], in.pos[3 * vid + 2]) : v3(0, 0, 0);
    o.nrm = in.nrm ? v3(in.nrm[3 * vid], in.nrm[3 * vid + 1], in.nrm[3 * vid + 2]) : v3(0, 0, 0);
    if (in.uv) {
        o.u = in.uv[2 * vid];
        o.v = in.uv[2 * vid + 1];
    } else {
        o.u = 0;
        o.v = 0;
    }
    o.col = in.col ? c4(in.col[4 * vid], in.col[4 * vid + 1], in.col[4 * vid + 2], in.col[4 * vid + 3])
                   : c4(0, 0, 0, 0);
    o.out = v4(0, 0, 0, 0);
    return o;
}

FGL_D V3 load_pos(const DevPrims &in, int64_t vid) {
    return in.pos ? v3(in.pos[3 * vid], in.pos[3 * vid + 1], in.pos[3 * vid + 2]) : v3(0, 0, 0);
}

// ---------------------------------------------------------------------------
// Raster triangle records

// True for a small box none of whose pixel centers the reference's inside
// test can accept: every center of [x0, x1] x [y0, y1] is surely outside
// (edge_margin).  Such a rasterize() call touches no pixel.
FGL_D bool surely_empty(V3 s0, V3 s1, V3 s2, int64_t x0, int64_t y0, int64_t x1, int64_t y1) {
    const EdgeMargin m = edge_margin(s0.x, s0.y, s1.x, s1.y, s2.x, s2.y, x0, y0, x1, y1, 3);
    if (!m.ok) return false;
    for (int64_t y = y0; y <= y1; y++)
        for (int64_t x = x0; x <= x1; x++)
            if (!surely_outside(m, s0.x, s0.y, s1.x, s1.y, s2.x, s2.y, x, y)) return false;
    return true;
}

// Writes (when out != nullptr) the record of one rasterize() call and
// reports how many tiles it is binned to.  Returns false when no pixel can
// result (the box is pathological, no index y*W+x of it lands in the
// framebuffer, or a small box has no center the inside test can accept);
// such a call has no effect in the reference either.
FGL_D bool make_raster(const DrawArgs &a, V3 s0, V3 s1, V3 s2, double W0, double W1, double W2, uint64_t attr,
                       PrimRec *out, uint32_t &ntiles, int4 *box = nullptr) {
    // integer bounding box (context.go:155-160)
    double mnx = __builtin_floor(go_min(s0.x, go_min(s1.x, s2.x)));
    double mny = __builtin_floor(go_min(s0.y, go_min(s1.y, s2.y)));
    double mxx = __builtin_ceil(go_max(s0.x, go_max(s1.x, s2.x)));
    double mxy = __builtin_ceil(go_max(s0.y, go_max(s1.y, s2.y)));
    int64_t x0 = go_f2i(mnx), x1 = go_f2i(mxx), y0 = go_f2i(mny), y1 = go_f2i(mxy);
    if (x0 < -BBOX_LIM || x0 > BBOX_LIM || x1 < -BBOX_LIM || x1 > BBOX_LIM || y0 < -BBOX_LIM || y0 > BBOX_LIM ||
        y1 < -BBOX_LIM || y1 > BBOX_LIM)
        return false;
    if (surely_empty(s0, s1, s2, x0, y0, x1, y1)) return false;
    TileRect r;
    if (!prim_tile_rect((int32_t)x0, (int32_t)y0, (int32_t)x1, (int32_t)y1, a.W, a.H, r)) return false;
    ntiles = (uint32_t)(r.tx1 - r.tx0 + 1) * (uint32_t)(r.ty1 - r.ty0 + 1);
    if (out) {
        out->sx[0] = s0.x; out->sx[1] = s1.x; out->sx[2] = s2.x;
        out->sy[0] = s0.y; out->sy[1] = s1.y; out->sy[2] = s2.y;
        out->sz[0] = s0.z; out->sz[1] = s1.z; out->sz[2] = s2.z;
        out->r[0] = 1.0 / W0;
        out->r[1] = 1.0 / W1;
        out->r[2] = 1.0 / W2;
        out->attr = attr;
        out->x0 = (int32_t)x0; out->y0 = (int32_t)y0;
        out->x1 = (int32_t)x1; out->y1 = (int32_t)y1;
        // ra = 1 / edge(s0, s1, s2) (context.go:175)
        out->ra = 1.0 / ((s1.x - s2.x) * (s0.y - s2.y) - (s1.y - s2.y) * (s0.x - s2.x));
        if (box) *box = make_int4((int32_t)x0, (int32_t)y0, (int32_t)x1, (int32_t)y1);
    }
    return true;
}

// Collects the raster triangles of one input primitive, counting on the
// first pass and writing on the second; box (main records only: one raster
// triangle) also receives the record's box, compact for the pair emitter.
struct Emitter {
    PrimRec *base;  // nullptr on the counting pass
    uint32_t n;
    uint32_t ntiles;
    int4 *box;
    FGL_D void add(const DrawArgs &a, V3 s0, V3 s1, V3 s2, double W0, double W1, double W2, uint64_t attr) {
        uint32_t t = 0;
        if (make_raster(a, s0, s1, s2, W0, W1, W2, attr, base ? base + n : nullptr, t, box)) {
            n++;
            ntiles += t;
        }
    }
};

// Vector.Perpendicular (vector.go:179-187)
FGL_D V3 perpendicular(V3 a) {
    if (a.x == 0 && a.y == 0) {
        if (a.z == 0) return v3(0, 0, 0);
        return v3(0, 1, 0);
    }
    return vnormalize(v3(-a.y, a.x, 0));
}

// line() (context.go:283-294) between vertices with selectors sa, sb of
// attribute source (kind, index).
FGL_D void emit_line(const DrawArgs &a, Emitter &em, uint32_t kind, uint64_t index, uint32_t sa, uint32_t sb,
                     V3 s0, V3 s1, double Wa, double Wb) {
    double half = a.st.line_width / 2;
    V3 n = vmuls(perpendicular(vsub(s1, s0)), a.st.line_width / 2);
    s0 = vadd(s0, vmuls(vnormalize(vsub(s0, s1)), half));
    s1 = vadd(s1, vmuls(vnormalize(vsub(s1, s0)), half));
    V3 s00 = vadd(s0, n);
    V3 s01 = vsub(s0, n);
    V3 s10 = vadd(s1, n);
    V3 s11 = vsub(s1, n);
    em.add(a, s11, s01, s00, Wb, Wa, Wa, make_attr(kind, sb, sa, sa, index));
    em.add(a, s10, s11, s00, Wb, Wb, Wa, make_attr(kind, sb, sb, sa, index));
}

// drawClippedTriangle (context.go:316-349) for three vertex outputs whose
// attributes are vertices sel[0..2] of source (kind, index).
FGL_D void emit_clipped_triangle(const DrawArgs &a, Emitter &em, uint32_t kind, uint64_t index, const uint32_t sel[3],
                                 V4 o0, V4 o1, V4 o2) {
    // normalized device coordinates (Output.DivScalar(W).Vector())
    V3 n0 = v3(o0.x / o0.w, o0.y / o0.w, o0.z / o0.w);
    V3 n1 = v3(o1.x / o1.w, o1.y / o1.w, o1.z / o1.w);
    V3 n2 = v3(o2.x / o2.w, o2.y / o2.w, o2.z / o2.w);
    // back face culling
    double ar = (n1.x - n0.x) * (n2.y - n0.y) - (n2.x - n0.x) * (n1.y - n0.y);
    uint32_t q0 = sel[0], q1 = sel[1], q2 = sel[2];
    double w0 = o0.w, w1 = o1.w, w2 = o2.w;
    if (ar < 0) {
        uint32_t tq = q0; q0 = q2; q2 = tq;
        double tw = w0; w0 = w2; w2 = tw;
        V3 tn = n0; n0 = n2; n2 = tn;
    }
    if (a.st.cull == 2 /* CullFront */) ar = -ar;
    if (a.st.front_face == 1 /* FaceCW */) ar = -ar;
    if (a.st.cull != 1 /* CullNone */ && ar <= 0) return;
    // screen coordinates
    V3 s0 = mul_position(a.S, n0);
    V3 s1 = mul_position(a.S, n1);
    V3 s2 = mul_position(a.S, n2);
    if (a.st.wireframe) {  // wireframe (context.go:296-301)
        emit_line(a, em, kind, index, q0, q1, s0, s1, w0, w1);
        emit_line(a, em, kind, index, q1, q2, s1, s2, w1, w2);
        emit_line(a, em, kind, index, q2, q0, s2, s0, w2, w0);
    } else {
        em.add(a, s0, s1, s2, w0, w1, w2, make_attr(kind, q0, q1, q2, index));
    }
}

// ---------------------------------------------------------------------------
// Clipping (clipping.go)

__constant__ double kClipP[6][4] = {{1, 0, 0, 1}, {-1, 0, 0, 1}, {0, 1, 0, 1},
                                    {0, -1, 0, 1}, {0, 0, 1, 1}, {0, 0, -1, 1}};
__constant__ double kClipN[6][4] = {{-1, 0, 0, 1}, {1, 0, 0, 1}, {0, -1, 0, 1},
                                    {0, 1, 0, 1}, {0, 0, -1, 1}, {0, 0, 1, 1}};

FGL_D V4 clipP(int p) { return v4(kClipP[p][0], kClipP[p][1], kClipP[p][2], kClipP[p][3]); }
FGL_D V4 clipN(int p) { return v4(kClipN[p][0], kClipN[p][1], kClipN[p][2], kClipN[p][3]); }

FGL_D bool point_in_front(int p, V4 v) { return wdot(wsub(v, clipP(p)), clipN(p)) > 0; }  // 16-18
FGL_D V4 intersect_segment(int p, V4 v0, V4 v1) {                                        // 20-26
    V4 u = wsub(v1, v0);
    V4 w = wsub(v0, clipP(p));
    V4 N = clipN(p);
    double d = wdot(N, u);
    double n = -wdot(N, w);
    return wadd(v0, wmuls(u, n / d));
}

// sutherlandHodgman (clipping.go:28-52); a triangle yields at most 9
// points.  buf[0][0..2] holds the triangle on entry; the polygon is left in
// buf[*which] and its point count returned.  buf is the caller's (rows of
// S >= 10 points; LDS in the clipping kernel).  sutherland_hodgman is the
// out-of-line form the generic vertex kernel calls.
template <int S>
FGL_D int sh_passes(V4 (*buf)[S], int *which) {
    int n = 3;
    int cur = 0;
    for (int p = 0; p < 6; p++) {
        if (n == 0) break;
        const V4 *input = buf[cur];
        V4 *output = buf[cur ^ 1];
        int m = 0;
        V4 s = input[n - 1];
        bool sf = point_in_front(p, s);
        for (int k = 0; k < n; k++) {
            V4 e = input[k];
            bool ef = point_in_front(p, e);
            if (ef) {
                if (!sf && m < 10) output[m++] = intersect_segment(p, s, e);
                if (m < 10) output[m++] = e;
            } else if (sf) {
                if (m < 10) output[m++] = intersect_segment(p, s, e);
            }
            s = e;
            sf = ef;
        }
        n = m;
        cur ^= 1;
    }
    *which = cur;
    return n;
}

__device__ __noinline__ int sutherland_hodgman(V4 (*buf)[10], int *which) { return sh_passes(buf, which); }

// sutherlandHodgman (clipping.go:28-52) for the clipping kernel: the same
// point tests, intersections and output order as sutherland_hodgman, with
// fewer dependent steps per point.  Each plane reads the polygon's points
// into registers at once and computes every intersection without branching
// (the ones used are those the reference computes); the outputs are written
// without branches, a write the reference does not make going to slot
// SH_MAX.  A plane that no lane of the wave clips is skipped (it would copy
// every polygon unchanged).  buf[0][0..2] holds the triangle on entry.
// Returns -1 for a lane whose polygon outgrows the SH_IN points a pass
// reads (a triangle has at most 9 after six planes; only rounding can add
// a tenth): the caller runs the plain passes (sh_passes) for it.
constexpr int SH_IN = 9;    // points a pass reads into registers
constexpr int SH_MAX = 10;  // points a pass keeps, as in sutherland_hodgman
constexpr int SH_SLOTS = SH_MAX + 1;
FGL_D int sutherland_hodgman_ilp(V4 (*buf)[SH_SLOTS], int *which) {
    int n = 3;
    int cur = 0;
    for (int p = 0; p < 6; p++) {
        if (n > SH_IN) return -1;
        const V4 *input = buf[cur];
        V4 *output = buf[cur ^ 1];
        V4 pt[SH_IN];
        bool fr[SH_IN];
        bool clips = false;
#pragma unroll
        for (int k = 0; k < SH_IN; k++) {
            pt[k] = input[k];
            fr[k] = point_in_front(p, pt[k]);
            clips |= k < n && !fr[k];
        }
        if (__ballot(clips) == 0) continue;
        const V4 s0 = input[n > 0 ? n - 1 : 0];
        V4 x[SH_IN];
#pragma unroll
        for (int k = 0; k < SH_IN; k++) x[k] = intersect_segment(p, k == 0 ? s0 : pt[k - 1], pt[k]);
        bool sf = point_in_front(p, s0);
        int m = 0;
#pragma unroll
        for (int k = 0; k < SH_IN; k++) {
            const bool act = k < n, ef = fr[k];
            const bool wx = act && ef != sf && m < SH_MAX;
            output[wx ? m : SH_MAX] = x[k];
            m += wx ? 1 : 0;
            const bool we = act && ef && m < SH_MAX;
            output[we ? m : SH_MAX] = pt[k];
            m += we ? 1 : 0;
            sf = act ? ef : sf;
        }
        n = m;
        cur ^= 1;
    }
    *which = cur;
    return n;
}

// Barycentric (vertex.go:81-95)
FGL_D V4 barycentric(V3 p1, V3 p2, V3 p3, V3 p) {
    V3 v0 = vsub(p2, p1);
    V3 v1 = vsub(p3, p1);
    V3 v2 = vsub(p, p1);
    double d00 = vdot(v0, v0);
    double d01 = vdot(v0, v1);
    double d11 = vdot(v1, v1);
    double d20 = vdot(v2, v0);
    double d21 = vdot(v2, v1);
    double d = d00 * d11 - d01 * d01;
    double v = (d11 * d20 - d01 * d21) / d;
    double w = (d00 * d21 - d01 * d20) / d;
    double u = 1 - v - w;
    return v4(u, v, w, 1);
}

// InterpolateVertexes (vertex.go:18-79) with b.W == 1, as ClipTriangle calls it.
FGL_D double ip(double a, double b, double c, V4 k) {
    double n = 0;
    n = n + a * k.x;
    n = n + b * k.y;
    n = n + c * k.z;
    return n * k.w;
}
FGL_D Vtx interpolate_clip(const Vtx &A, const Vtx &B, const Vtx &C, V4 b) {
    Vtx o;
    o.pos = v3(ip(A.pos.x, B.pos.x, C.pos.x, b), ip(A.pos.y, B.pos.y, C.pos.y, b), ip(A.pos.z, B.pos.z, C.pos.z, b));
    o.nrm = vnormalize(
        v3(ip(A.nrm.x, B.nrm.x, C.nrm.x, b), ip(A.nrm.y, B.nrm.y, C.nrm.y, b), ip(A.nrm.z, B.nrm.z, C.nrm.z, b)));
    o.u = ip(A.u, B.u, C.u, b);
    o.v = ip(A.v, B.v, C.v, b);
    o.col = c4(ip(A.col.r, B.col.r, C.col.r, b), ip(A.col.g, B.col.g, C.col.g, b), ip(A.col.b, B.col.b, C.col.b, b),
               ip(A.col.a, B.col.a, C.col.a, b));
    o.out = v4(ip(A.out.x, B.out.x, C.out.x, b), ip(A.out.y, B.out.y, C.out.y, b), ip(A.out.z, B.out.z, C.out.z, b),
               ip(A.out.w, B.out.w, C.out.w, b));
    return o;
}

// Triangle.FixNormals (triangle.go:33-37, 46-58)
FGL_D void fix_normals(Vtx &A, Vtx &B, Vtx &C) {
    V3 e1 = vsub(B.pos, A.pos);
    V3 e2 = vsub(C.pos, A.pos);
    V3 n = vnormalize(vcross(e1, e2));
    if (A.nrm.x == 0 && A.nrm.y == 0 && A.nrm.z == 0) A.nrm = n;
    if (B.nrm.x == 0 && B.nrm.y == 0 && B.nrm.z == 0) B.nrm = n;
    if (C.nrm.x == 0 && C.nrm.y == 0 && C.nrm.z == 0) C.nrm = n;
}

FGL_D void store_side_vertex(const DrawArgs &a, uint64_t v, const Vtx &F) {
    double *d = a.side_vtx + v * SIDE_VTX_DOUBLES;
    d[0] = F.pos.x; d[1] = F.pos.y; d[2] = F.pos.z;
    d[3] = F.nrm.x; d[4] = F.nrm.y; d[5] = F.nrm.z;
    d[6] = F.u; d[7] = F.v;
    d[8] = F.col.r; d[9] = F.col.g; d[10] = F.col.b; d[11] = F.col.a;
}

// ClipTriangle + per-fan drawClippedTriangle: the fan triangles are built
// (vertex attributes written at side_vtx[vbase..]) and emitted into em.
// Barycentric + InterpolateVertexes of a polygon point give the same
// vertex every time the reference recomputes them (clipping.go:64-70), so
// each point is interpolated once; NewTriangle's FixNormals is applied per
// fan to copies.
FGL_D void clip_fans(const DrawArgs &a, const Vtx &V0, const Vtx &V1, const Vtx &V2, const V4 *poly, int np,
                     Emitter &em, uint32_t vbase) {
    const V3 p1 = v3(V0.out.x, V0.out.y, V0.out.z);
    const V3 p2 = v3(V1.out.x, V1.out.y, V1.out.z);
    const V3 p3 = v3(V2.out.x, V2.out.y, V2.out.z);
    const Vtx F0 = interpolate_clip(V0, V1, V2, barycentric(p1, p2, p3, v3(poly[0].x, poly[0].y, poly[0].z)));
    Vtx Fp = interpolate_clip(V0, V1, V2, barycentric(p1, p2, p3, v3(poly[1].x, poly[1].y, poly[1].z)));
    for (int i = 2; i < np; i++) {
        const Vtx Fi = interpolate_clip(V0, V1, V2, barycentric(p1, p2, p3, v3(poly[i].x, poly[i].y, poly[i].z)));
        Vtx T0 = F0, T1 = Fp, T2 = Fi;
        fix_normals(T0, T1, T2);  // NewTriangle
        const uint64_t vb = (uint64_t)vbase + 3u * (uint32_t)(i - 2);
        store_side_vertex(a, vb, T0);
        store_side_vertex(a, vb + 1, T1);
        store_side_vertex(a, vb + 2, T2);
        const uint32_t sel[3] = {0, 1, 2};
        emit_clipped_triangle(a, em, ATTR_SIDE, vb, sel, T0.out, T1.out, T2.out);
        Fp = Fi;
    }
}

// ClipLine (clipping.go:76-98): clips the outputs only; false for nil.
FGL_D bool clip_line(V4 &w1, V4 &w2) {
    for (int p = 0; p < 6; p++) {
        bool f1 = point_in_front(p, w1);
        bool f2 = point_in_front(p, w2);
        if (f1 && f2) continue;
        else if (f1) w2 = intersect_segment(p, w1, w2);
        else if (f2) w1 = intersect_segment(p, w2, w1);
        else return false;
    }
    return true;
}

FGL_D bool alloc_side(const DrawArgs &a, uint32_t nrec, uint32_t nvtx, uint32_t &rbase, uint32_t &vbase) {
    rbase = nrec ? atomicAdd(&a.ctr->side_rec, nrec) : 0u;
    vbase = nvtx ? atomicAdd(&a.ctr->side_vtx, nvtx) : 0u;
    bool ok = true;
    if (nrec && (uint64_t)rbase + nrec > a.side_rec_cap) {
        atomicOr(&a.ctr->overflow, 1u);
        ok = false;
    }
    if (nvtx && (uint64_t)vbase + nvtx > a.side_vtx_cap) {
        atomicOr(&a.ctr->overflow, 2u);
        ok = false;
    }
    return ok;
}

// The clipped branch of DrawTriangle (context.go:376-384):
// ClipTriangle(NewTriangle(v1, v2, v3)) and each fan triangle in order.
// Side records are reserved for every fan triangle (six per fan in
// wireframe) and filled from the first; nr counts the ones written.
__device__ __noinline__ void vs_triangle_clip(const DrawArgs &a, int64_t i, V4 o0, V4 o1, V4 o2, uint32_t &nr,
                                              uint32_t &nt, uint32_t &ss, V4 (*buf)[10]) {
    buf[0][0] = o0;
    buf[0][1] = o1;
    buf[0][2] = o2;
    int which = 0;
    const int np = sutherland_hodgman(buf, &which);
    if (np < 3) return;
    const uint32_t nfan = (uint32_t)(np - 2);
    Vtx V0 = load_vertex(a.in, 3 * i), V1 = load_vertex(a.in, 3 * i + 1), V2 = load_vertex(a.in, 3 * i + 2);
    V0.out = o0;
    V1.out = o1;
    V2.out = o2;
    fix_normals(V0, V1, V2);  // NewTriangle (context.go:382)
    uint32_t rb, vb;
    if (!alloc_side(a, nfan * (a.st.wireframe ? 6u : 1u), 3 * nfan, rb, vb)) return;
    Emitter em{a.side_rec + rb, 0, 0};
    clip_fans(a, V0, V1, V2, buf[which], np, em, vb);
    if (em.n == 0) return;
    nr = em.n | REF_SIDE;
    nt = em.ntiles;
    ss = rb;
}

// One input triangle (DrawTriangle, context.go:370-389), any state.
FGL_D void vs_triangle(const DrawArgs &a, int64_t i, uint32_t &nr, uint32_t &nt, uint32_t &ss) {
    V3 P0 = load_pos(a.in, 3 * i), P1 = load_pos(a.in, 3 * i + 1), P2 = load_pos(a.in, 3 * i + 2);
    V4 o0 = mul_position_w(a.sh.M, P0);
    V4 o1 = mul_position_w(a.sh.M, P1);
    V4 o2 = mul_position_w(a.sh.M, P2);
    const uint32_t ident[3] = {0, 1, 2};
    if (!(woutside(o0) || woutside(o1) || woutside(o2))) {
        if (!a.st.wireframe) {
            Emitter em{a.rec_main + i, 0, 0, a.tbox + i};
            emit_clipped_triangle(a, em, ATTR_TRI, (uint64_t)i, ident, o0, o1, o2);
            nr = em.n;
            nt = em.ntiles;
            return;
        }
        Emitter cnt{nullptr, 0, 0};
        emit_clipped_triangle(a, cnt, ATTR_TRI, (uint64_t)i, ident, o0, o1, o2);
        if (cnt.n == 0) return;
        uint32_t rb, vb;
        if (!alloc_side(a, cnt.n, 0, rb, vb)) return;
        Emitter em{a.side_rec + rb, 0, 0};
        emit_clipped_triangle(a, em, ATTR_TRI, (uint64_t)i, ident, o0, o1, o2);
        nr = em.n | REF_SIDE;
        nt = em.ntiles;
        ss = rb;
        return;
    }
    V4 buf[2][10];
    vs_triangle_clip(a, i, o0, o1, o2, nr, nt, ss, buf);
}

// Filled triangles without clipping in one lean pass; triangles with a
// vertex outside the view volume are queued for k_vs_deferred, which fills
// in their counts afterwards.
template <class G>
FGL_D void vs_fast_body(const DrawArgs &a, const G &g) {
    __shared__ uint64_t wsum[VS_THREADS / 64];
    const int lane = threadIdx.x & 63;
    int64_t i = (int64_t)g.bid() * VS_THREADS + threadIdx.x;
    uint32_t nr = 0, nt = 0;
    bool defer = false;
    if (i < a.in.n) {
        V3 P0 = load_pos(a.in, 3 * i), P1 = load_pos(a.in, 3 * i + 1), P2 = load_pos(a.in, 3 * i + 2);
        V4 o0 = mul_position_w(a.sh.M, P0);
        V4 o1 = mul_position_w(a.sh.M, P1);
        V4 o2 = mul_position_w(a.sh.M, P2);
        if (woutside(o0) || woutside(o1) || woutside(o2)) {
            // sutherlandHodgman (clipping.go:28-52) leaves the triangle
            // untouched by every plane all three vertices are in front of;
            // if the first plane that is not so has no vertex in front, its
            // pass emits nothing and ClipTriangle returns no triangles.
            bool rejected = false;
            for (int p = 0; p < 6; p++) {
                bool f0 = point_in_front(p, o0), f1 = point_in_front(p, o1), f2 = point_in_front(p, o2);
                if (f0 && f1 && f2) continue;
                rejected = !f0 && !f1 && !f2;
                break;
            }
            defer = !rejected;
        } else {
            const uint32_t ident[3] = {0, 1, 2};
            Emitter em{a.rec_main + i, 0, 0, a.tbox + i};
            emit_clipped_triangle(a, em, ATTR_TRI, (uint64_t)i, ident, o0, o1, o2);
            nr = em.n;
            nt = em.ntiles;
        }
        // (sstart is read only for side records; k_vs_deferred writes it
        // for the triangles queued here)
        a.nrast[i] = nr;
        a.ntl[i] = nt;
    }
    // queue the deferred triangles: one counter update per wave
    const uint64_t m = __ballot(defer);
    if (m) {
        const int leader = __ffsll((unsigned long long)m) - 1;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&a.ctr->ndeferred, (uint32_t)__popcll(m));
        base = __shfl(base, leader);
        if (defer) a.deferred[base + __popcll(m & ((1ull << lane) - 1))] = (uint32_t)i;
    }
    uint64_t v = nt;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if (lane == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t s = 0;
        for (int w = 0; w < VS_THREADS / 64; w++) s += wsum[w];
        a.blocksum[g.bid()] = s;
    }
}
FGL_DRAW_KERNELS(k_vs_fast, vs_fast_body, __launch_bounds__(VS_THREADS))

// Exclusive positions of this lane's nrec side records and 3 nrec side
// vertices in blocks of the two counters reserved for the whole wave (one
// atomic on each, both in flight together).  All 64 lanes call it.
FGL_D void wave_reserve_fans(const DrawArgs &a, uint32_t nrec, uint32_t &rb, uint32_t &vb) {
    const int lane = threadIdx.x & 63;
    uint32_t incl = nrec;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off);
        if (lane >= off) incl += y;
    }
    const uint32_t total = __shfl(incl, 63);
    uint32_t r0 = 0, v0 = 0;
    if (lane == 63 && total) {
        r0 = atomicAdd(&a.ctr->side_rec, total);
        v0 = atomicAdd(&a.ctr->side_vtx, 3 * total);
    }
    r0 = __shfl(r0, 63);
    v0 = __shfl(v0, 63);
    rb = r0 + incl - nrec;
    vb = v0 + 3 * (incl - nrec);
}

// The triangles k_vs_fast queued, one per lane: clip (polygon in LDS),
// reserve the side records and vertices of the whole wave at once, then
// build and emit the fans.
template <class G>
FGL_D void vs_deferred_body(const DrawArgs &a, const G &g) {
    __shared__ V4 clipbuf[64][2][SH_SLOTS];
    const uint32_t n = a.ctr->ndeferred;
    V4(*buf)[SH_SLOTS] = clipbuf[threadIdx.x];
    for (uint32_t q0 = g.bid() * 64; q0 < n; q0 += g.nblk() * 64) {
        const uint32_t q = q0 + threadIdx.x;
        const bool act = q < n;
        const int64_t i = act ? (int64_t)a.deferred[q] : 0;
        int np = 0, which = 0;
        V4 o0, o1, o2;
        if (act) {
            V3 P0 = load_pos(a.in, 3 * i), P1 = load_pos(a.in, 3 * i + 1), P2 = load_pos(a.in, 3 * i + 2);
            o0 = mul_position_w(a.sh.M, P0);
            o1 = mul_position_w(a.sh.M, P1);
            o2 = mul_position_w(a.sh.M, P2);
            buf[0][0] = o0;
            buf[0][1] = o1;
            buf[0][2] = o2;
            np = sutherland_hodgman_ilp(buf, &which);
            if (np < 0) {
                buf[0][0] = o0;
                buf[0][1] = o1;
                buf[0][2] = o2;
                np = sh_passes(buf, &which);
            }
        }
        const uint32_t nfan = np >= 3 ? (uint32_t)(np - 2) : 0u;
        uint32_t rb, vb;
        wave_reserve_fans(a, nfan, rb, vb);
        uint32_t nr = 0, nt = 0, ss = 0;
        if (nfan) {
            bool ok = true;
            if ((uint64_t)rb + nfan > a.side_rec_cap) {
                atomicOr(&a.ctr->overflow, 1u);
                ok = false;
            }
            if ((uint64_t)vb + 3 * nfan > a.side_vtx_cap) {
                atomicOr(&a.ctr->overflow, 2u);
                ok = false;
            }
            if (ok) {
                Vtx V0 = load_vertex(a.in, 3 * i), V1 = load_vertex(a.in, 3 * i + 1),
                    V2 = load_vertex(a.in, 3 * i + 2);
                V0.out = o0;
                V1.out = o1;
                V2.out = o2;
                fix_normals(V0, V1, V2);  // NewTriangle (context.go:382)
                Emitter em{a.side_rec + rb, 0, 0};
                clip_fans(a, V0, V1, V2, buf[which], np, em, vb);
                if (em.n) {
                    nr = em.n | REF_SIDE;
                    nt = em.ntiles;
                    ss = rb;
                }
            }
        }
        if (act) {
            a.nrast[i] = nr;
            a.ntl[i] = nt;
            a.sstart[i] = ss;
            if (nt) atomicAdd((unsigned long long *)&a.blocksum[i / VS_THREADS], (unsigned long long)nt);
        }
    }
}
FGL_DRAW_KERNELS(k_vs_deferred, vs_deferred_body, __launch_bounds__(64))

// One input line (DrawLine, context.go:351-368).
FGL_D void vs_line(const DrawArgs &a, int64_t i, uint32_t &nr, uint32_t &nt, uint32_t &ss) {
    V3 P0 = load_pos(a.in, 2 * i), P1 = load_pos(a.in, 2 * i + 1);
    V4 w1 = mul_position_w(a.sh.M, P0);
    V4 w2 = mul_position_w(a.sh.M, P1);
    if (woutside(w1) || woutside(w2)) {
        if (!clip_line(w1, w2)) return;
    }
    // drawClippedLine (context.go:303-314)
    V3 n0 = v3(w1.x / w1.w, w1.y / w1.w, w1.z / w1.w);
    V3 n1 = v3(w2.x / w2.w, w2.y / w2.w, w2.z / w2.w);
    V3 s0 = mul_position(a.S, n0);
    V3 s1 = mul_position(a.S, n1);
    Emitter cnt{nullptr, 0, 0};
    emit_line(a, cnt, ATTR_LINE, (uint64_t)i, 0, 1, s0, s1, w1.w, w2.w);
    if (cnt.n == 0) return;
    uint32_t rb, vb;
    if (!alloc_side(a, cnt.n, 0, rb, vb)) return;
    Emitter em{a.side_rec + rb, 0, 0};
    emit_line(a, em, ATTR_LINE, (uint64_t)i, 0, 1, s0, s1, w1.w, w2.w);
    nr = em.n | REF_SIDE;
    nt = em.ntiles;
    ss = rb;
}

template <int KIND, class G>
FGL_D void vs_body(const DrawArgs &a, const G &g) {
    __shared__ uint64_t wsum[VS_THREADS / 64];
    int64_t i = (int64_t)g.bid() * VS_THREADS + threadIdx.x;
    uint32_t nr = 0, nt = 0, ss = 0;
    if (i < a.in.n) {
        if (KIND == PRIM_TRI) vs_triangle(a, i, nr, nt, ss);
        else vs_line(a, i, nr, nt, ss);
        a.nrast[i] = nr;
        a.ntl[i] = nt;
        a.sstart[i] = ss;
    }
    // block sum of pairs
    uint64_t v = nt;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t s = 0;
        for (int w = 0; w < VS_THREADS / 64; w++) s += wsum[w];
        a.blocksum[g.bid()] = s;
    }
}
template <int KIND>
__global__ __launch_bounds__(VS_THREADS) void k_vs(DrawArgs a) {
    vs_body<KIND>(a, OneDraw{});
}
template <int KIND>
__global__ __launch_bounds__(VS_THREADS) void k_vs_m(MultiArgs m) {
    const ManyDraws g = many_draws(m);
    vs_body<KIND>(draw_of(g), g);
}

hipError_t launch_vs(const DrawArgs &a, hipStream_t s) {
    if (a.kind == PRIM_TRI && !a.st.wireframe)
        hipLaunchKernelGGL(k_vs_fast, dim3(a.nvsblocks), dim3(VS_THREADS), 0, s, a);
    else if (a.kind == PRIM_TRI)
        hipLaunchKernelGGL(k_vs<PRIM_TRI>, dim3(a.nvsblocks), dim3(VS_THREADS), 0, s, a);
    else
        hipLaunchKernelGGL(k_vs<PRIM_LINE>, dim3(a.nvsblocks), dim3(VS_THREADS), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_vs_multi(const MultiArgs &m, uint32_t nb, const DrawArgs &rep, hipStream_t s) {
    if (rep.kind == PRIM_TRI && !rep.st.wireframe)
        hipLaunchKernelGGL(k_vs_fast_m, dim3(nb), dim3(VS_THREADS), 0, s, m);
    else if (rep.kind == PRIM_TRI)
        hipLaunchKernelGGL(k_vs_m<PRIM_TRI>, dim3(nb), dim3(VS_THREADS), 0, s, m);
    else
        hipLaunchKernelGGL(k_vs_m<PRIM_LINE>, dim3(nb), dim3(VS_THREADS), 0, s, m);
    return hipGetLastError();
}

// The queued triangles are taken 64 per block in a grid-stride loop.  A
// launch for many draws gives each fewer blocks: most draws queue few
// triangles, and every block holds 44 KB of LDS while it runs.
uint32_t blocks_vs_deferred(const DrawArgs &a, bool many) {
    int64_t blocks = (a.in.n + 63) / 64;
    const int64_t cap = many ? 16 : 4096;
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    return (uint32_t)blocks;
}
hipError_t launch_vs_deferred(const DrawArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_vs_deferred, dim3(blocks_vs_deferred(a, false)), dim3(64), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_vs_deferred_multi(const MultiArgs &m, uint32_t nb, const DrawArgs &, hipStream_t s) {
    hipLaunchKernelGGL(k_vs_deferred_m, dim3(nb), dim3(64), 0, s, m);
    return hipGetLastError();
}

}  // namespace FGL_TILE_NS
}  // namespace fgl
