/*
 * fauxgl_oracle.c -- CPU restatement of fauxgl's draw path (test
 * infrastructure only: the checker and the timed CPU baseline; the product
 * never links it).  See fauxgl_oracle.h for the semantics and what pins each
 * part.  Every function cites the reference file:line it restates.
 *
 * Build: gcc -O2 -ffp-contract=off -fPIC -shared (see oracle/Makefile).
 * Do not build with -ffast-math or -march flags that enable FMA contraction.
 */
#include "fauxgl_oracle.h"

#include <ctype.h>
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* Raster triangles whose integer bounding box leaves [-LIM, LIM] are
 * skipped.  In the reference such boxes come only from NaN/Inf screen
 * coordinates (int() of NaN/Inf is INT64_MIN on amd64), where its loops
 * either never end or wrap int64; no finite, clipped input reaches it. */
#define ORC_BBOX_LIM ((int64_t)1 << 26)

/* Planted mistakes for the mutation tests (tests/test_oracle_mutants.py
 * builds the oracle with -DORC_MUTANT=<n> and expects the pinning tests to
 * fail).  The default build defines none of them. */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

/* ------------------------------------------------------------------ */
/* L0 math types (vector.go:8-10, 204-206; color.go:17-19; matrix.go:5-10) */

typedef struct { double X, Y, Z; } Vec;
typedef struct { double X, Y, Z, W; } VecW;
typedef struct { double R, G, B, A; } Col;
typedef struct { double m[16]; } Mat; /* X00 X01 X02 X03 X10 ... X33 */

/* vertex.go:3-12 */
typedef struct {
    Vec Position, Normal, Texture;
    Col Color;
    VecW Output;
} Vertex;
typedef struct { Vertex V1, V2, V3; } Triangle; /* triangle.go:3-5 */
typedef struct { Vertex V1, V2; } Line;         /* line.go:3-5 */

/* ------------------------------------------------------------------ */
/* Go semantics helpers */

/* math.Min: -Inf wins over NaN, NaN propagates, Min(-0,+0) = -0. */
static double go_min(double x, double y) {
    if ((isinf(x) && x < 0) || (isinf(y) && y < 0)) return -INFINITY;
    if (isnan(x) || isnan(y)) return NAN;
    if (x == 0 && x == y) return signbit(x) ? x : y;
    return x < y ? x : y;
}

/* math.Max: +Inf wins over NaN, NaN propagates, Max(+0,-0) = +0. */
static double go_max(double x, double y) {
    if ((isinf(x) && x > 0) || (isinf(y) && y > 0)) return INFINITY;
    if (isnan(x) || isnan(y)) return NAN;
    if (x == 0 && x == y) return signbit(x) ? y : x;
    return x > y ? x : y;
}

/* int(x) on amd64 (CVTTSD2SQ): truncation; NaN / out of range -> INT64_MIN. */
static int64_t go_f2i(double x) {
    if (!(x >= -9223372036854775808.0 && x < 9223372036854775808.0)) return INT64_MIN;
    return (int64_t)x;
}

/* uint8(x) / uint16(x) on amd64 convert through int32 (CVTTSD2SL), then
 * truncate; NaN / out of range -> INT32_MIN, whose low bits are 0. */
static int32_t go_f2i32(double x) {
    if (!(x > -2147483649.0 && x < 2147483648.0)) return INT32_MIN;
    return (int32_t)x;
}
static uint8_t go_u8(double x) { return (uint8_t)(uint32_t)go_f2i32(x); }
static uint16_t go_u16(double x) { return (uint16_t)(uint32_t)go_f2i32(x); }

/* util.go:74-82 (NaN passes through). */
static double clampd(double x, double lo, double hi) {
    if (x < lo) return lo;
    if (x > hi) return hi;
    return x;
}

static int go_is_odd_int(double x) {
    if (fabs(x) >= 9007199254740992.0) return 0;
    double xi;
    double xf = modf(x, &xi);
    return xf == 0 && (((int64_t)xi) & 1) == 1;
}

/* Go math.Pow: special cases, then x**yi by repeated squaring of the
 * Frexp mantissa with the binary exponent tracked separately, times
 * x**yf = exp(yf*log(x)) with the C library's exp and log.  For non-integer
 * y the value is pinned against exact powers within a stated ulp bound;
 * its last bits against Go's own Exp and Log are not. */
static double go_pow(double x, double y) {
    if (y == 0 || x == 1) return 1;
    if (y == 1) return x;
    if (isnan(x) || isnan(y)) return NAN;
    if (x == 0) {
        if (y < 0) {
            if (signbit(x) && go_is_odd_int(y)) return -INFINITY;
            return INFINITY;
        }
        if (signbit(x) && go_is_odd_int(y)) return x;
        return 0;
    }
    if (isinf(y)) {
        if (x == -1) return 1;
        if ((fabs(x) < 1) == (y > 0)) return 0;
        return INFINITY;
    }
    if (isinf(x)) {
        if (x < 0) return go_pow(1 / x, -y);
        if (y < 0) return 0;
        return INFINITY;
    }
    if (y == 0.5) return sqrt(x);
    if (y == -0.5) return 1 / sqrt(x);

    double yi;
    double yf = modf(fabs(y), &yi);
    if (yf != 0 && x < 0) return NAN;
    if (yi >= 9223372036854775808.0) {
        if (x == -1) return 1;
        if ((fabs(x) < 1) == (y > 0)) return 0;
        return INFINITY;
    }
    double a1 = 1.0;
    int64_t ae = 0;
    if (yf != 0) {
        if (yf > 0.5) {
            yf--;
            if (ORC_MUTANT != 16) yi++;
        }
        a1 = ORC_MUTANT == 15 ? exp(yf * log10(x)) : exp(yf * log(x));
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
            if (ORC_MUTANT != 1) xe--;
        }
    }
    if (y < 0) {
        a1 = 1 / a1;
        ae = -ae;
    }
    if (ae > 100000) ae = 100000;
    if (ae < -100000) ae = -100000;
    return ldexp(a1, (int)ae);
}

/* ------------------------------------------------------------------ */
/* vector.go */

static Vec vadd(Vec a, Vec b) { return (Vec){a.X + b.X, a.Y + b.Y, a.Z + b.Z}; }  /* 96-98 */
static Vec vsub(Vec a, Vec b) { return (Vec){a.X - b.X, a.Y - b.Y, a.Z - b.Z}; }  /* 100-102 */
static Vec vmuls(Vec a, double b) { return (Vec){a.X * b, a.Y * b, a.Z * b}; }     /* 128-130 */
static Vec vdivs(Vec a, double b) { return (Vec){a.X / b, a.Y / b, a.Z / b}; }     /* 132-134 */
static double vdot(Vec a, Vec b) { return a.X * b.X + a.Y * b.Y + a.Z * b.Z; }    /* 72-74 */
static Vec vneg(Vec a) { return (Vec){-a.X, -a.Y, -a.Z}; }                        /* 88-90 */
static Vec vcross(Vec a, Vec b) {                                                  /* 76-81 */
    double x = a.Y * b.Z - a.Z * b.Y;
    double y = a.Z * b.X - a.X * b.Z;
    double z = a.X * b.Y - a.Y * b.X;
    return (Vec){x, y, z};
}
static Vec vnormalize(Vec a) { /* 83-86 */
    double r = 1 / sqrt(a.X * a.X + a.Y * a.Y + a.Z * a.Z);
    return (Vec){a.X * r, a.Y * r, a.Z * r};
}
static Vec vmin(Vec a, Vec b) { return (Vec){go_min(a.X, b.X), go_min(a.Y, b.Y), go_min(a.Z, b.Z)}; } /* 140-142 */
static Vec vmax(Vec a, Vec b) { return (Vec){go_max(a.X, b.X), go_max(a.Y, b.Y), go_max(a.Z, b.Z)}; } /* 144-146 */
static Vec vfloor(Vec a) { return (Vec){floor(a.X), floor(a.Y), floor(a.Z)}; }    /* 148-150 */
static Vec vceil(Vec a) { return (Vec){ceil(a.X), ceil(a.Y), ceil(a.Z)}; }        /* 152-154 */
static Vec vreflect(Vec i, Vec n) { /* 175-177 */
    if (ORC_MUTANT == 2) return vadd(i, vmuls(n, 2 * vdot(n, i)));
    return vsub(i, vmuls(n, 2 * vdot(n, i)));
}
static Vec vperpendicular(Vec a) {                                                 /* 179-187 */
    if (a.X == 0 && a.Y == 0) {
        if (a.Z == 0) return (Vec){0, 0, 0};
        return (Vec){0, 1, 0};
    }
    return vnormalize((Vec){-a.Y, a.X, 0});
}
static int vequal(Vec a, Vec b) { return a.X == b.X && a.Y == b.Y && a.Z == b.Z; }

static Vec wvec(VecW a) { return (Vec){a.X, a.Y, a.Z}; }                           /* 208-210 */
static int woutside(VecW a) {                                                      /* 212-215 */
    double x = a.X, y = a.Y, z = a.Z, w = a.W;
    return x < -w || x > w || y < -w || y > w || z < -w || z > w;
}
static double wdot(VecW a, VecW b) { return a.X * b.X + a.Y * b.Y + a.Z * b.Z + a.W * b.W; } /* 217-219 */
static VecW wadd(VecW a, VecW b) { return (VecW){a.X + b.X, a.Y + b.Y, a.Z + b.Z, a.W + b.W}; }
static VecW wsub(VecW a, VecW b) { return (VecW){a.X - b.X, a.Y - b.Y, a.Z - b.Z, a.W - b.W}; }
static VecW wmuls(VecW a, double b) { return (VecW){a.X * b, a.Y * b, a.Z * b, a.W * b}; }
static VecW wdivs(VecW a, double b) { return (VecW){a.X / b, a.Y / b, a.Z / b, a.W / b}; }

/* ------------------------------------------------------------------ */
/* color.go */

static Col cadd(Col a, Col b) { return (Col){a.R + b.R, a.G + b.G, a.B + b.B, a.A + b.A}; }  /* 80-82 */
static Col cmul(Col a, Col b) { return (Col){a.R * b.R, a.G * b.G, a.B * b.B, a.A * b.A}; }  /* 88-90 */
static Col cmuls(Col a, double b) { return (Col){a.R * b, a.G * b, a.B * b, a.A * b}; }      /* 104-106 */
static Col cmin(Col a, Col b) {                                                               /* 116-118 */
    return (Col){go_min(a.R, b.R), go_min(a.G, b.G), go_min(a.B, b.B), go_min(a.A, b.A)};
}
static int cdiscard(Col c) { return c.R == 0 && c.G == 0 && c.B == 0 && c.A == 0; } /* context.go:241 */

typedef struct { uint8_t R, G, B, A; } NRGBA8;

/* color.go:56-63 */
static NRGBA8 cnrgba(Col c) {
    const double d = 0xff;
    double r = clampd(c.R, 0, 1), g = clampd(c.G, 0, 1), b = clampd(c.B, 0, 1), a = clampd(c.A, 0, 1);
    if (ORC_MUTANT == 3) { /* rounds instead of truncating */
        r = r + 0.5 / d;
        g = g + 0.5 / d;
        b = b + 0.5 / d;
        a = a + 0.5 / d;
    }
    NRGBA8 o = {go_u8(r * d), go_u8(g * d), go_u8(b * d), go_u8(a * d)};
    return o;
}

/* image/color NRGBA.RGBA(): 16-bit alpha-premultiplied components. */
static void nrgba_rgba(NRGBA8 c, uint32_t *r, uint32_t *g, uint32_t *b, uint32_t *a) {
    uint32_t x;
    x = c.R; x |= x << 8; x *= c.A; x /= 0xff; *r = x;
    x = c.G; x |= x << 8; x *= c.A; x /= 0xff; *g = x;
    x = c.B; x |= x << 8; x *= c.A; x /= 0xff; *b = x;
    x = c.A; x |= x << 8; *a = x;
}

/* ------------------------------------------------------------------ */
/* matrix.go */

#define M(mat, r, c) ((mat).m[(r) * 4 + (c)])

static Mat mat_mul(Mat a, Mat b) { /* 188-207 */
    Mat o;
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++)
            M(o, r, c) = M(a, r, 0) * M(b, 0, c) + M(a, r, 1) * M(b, 1, c) + M(a, r, 2) * M(b, 2, c) +
                         M(a, r, 3) * M(b, 3, c);
    return o;
}
static Vec mat_mul_position(Mat a, Vec b) { /* 209-214 */
    double x = M(a, 0, 0) * b.X + M(a, 0, 1) * b.Y + M(a, 0, 2) * b.Z + M(a, 0, 3);
    double y = M(a, 1, 0) * b.X + M(a, 1, 1) * b.Y + M(a, 1, 2) * b.Z + M(a, 1, 3);
    double z = M(a, 2, 0) * b.X + M(a, 2, 1) * b.Y + M(a, 2, 2) * b.Z + M(a, 2, 3);
    return (Vec){x, y, z};
}
static VecW mat_mul_position_w(Mat a, Vec b) { /* 216-222 */
    double x = M(a, 0, 0) * b.X + M(a, 0, 1) * b.Y + M(a, 0, 2) * b.Z + M(a, 0, 3);
    double y = M(a, 1, 0) * b.X + M(a, 1, 1) * b.Y + M(a, 1, 2) * b.Z + M(a, 1, 3);
    double z = M(a, 2, 0) * b.X + M(a, 2, 1) * b.Y + M(a, 2, 2) * b.Z + M(a, 2, 3);
    double w = M(a, 3, 0) * b.X + M(a, 3, 1) * b.Y + M(a, 3, 2) * b.Z + M(a, 3, 3);
    return (VecW){x, y, z, w};
}
static Mat mat_screen(int w, int h) { /* 119-128 */
    double w2 = (double)w / 2;
    double h2 = (double)h / 2;
    Mat o = {{w2, 0, 0, w2, 0, -h2, 0, h2, 0, 0, 0.5, 0.5, 0, 0, 0, 1}};
    return o;
}
static Mat mat_frustum(double l, double r, double b, double t, double n, double f) { /* 69-79 */
    double t1 = 2 * n, t2 = r - l, t3 = t - b, t4 = f - n;
    Mat o = {{t1 / t2, 0, (r + l) / t2, 0, 0, t1 / t3, (t + b) / t3, 0, 0, 0, (-f - n) / t4, (-t1 * f) / t4, 0, 0, -1, 0}};
    return o;
}
static Mat mat_perspective(double fovy, double aspect, double near, double far) { /* 89-93 */
    double ymax = near * tan(fovy * M_PI / 360);
    double xmax = ymax * aspect;
    return mat_frustum(-xmax, xmax, -ymax, ymax, near, far);
}
static Mat mat_look_at(Vec eye, Vec center, Vec up) { /* 95-105 */
    Vec z = vnormalize(vsub(eye, center));
    Vec x = vnormalize(vcross(up, z));
    Vec y = ORC_MUTANT == 4 ? vcross(x, z) : vcross(z, x);
    Mat o = {{x.X, x.Y, x.Z, -vdot(x, eye), y.X, y.Y, y.Z, -vdot(y, eye), z.X, z.Y, z.Z, -vdot(z, eye), 0, 0, 0, 1}};
    return o;
}

/* ------------------------------------------------------------------ */
/* vertex.go / triangle.go */

static Vec interp_vec(Vec v1, Vec v2, Vec v3, VecW b) { /* vertex.go:59-65 */
    Vec n = {0, 0, 0};
    n = vadd(n, vmuls(v1, b.X));
    n = vadd(n, vmuls(v2, b.Y));
    n = vadd(n, vmuls(v3, b.Z));
    return vmuls(n, b.W);
}
static Col interp_col(Col v1, Col v2, Col v3, VecW b) { /* vertex.go:51-57 */
    Col n = {0, 0, 0, 0};
    n = cadd(n, cmuls(v1, b.X));
    n = cadd(n, cmuls(v2, b.Y));
    n = cadd(n, cmuls(v3, b.Z));
    return cmuls(n, b.W);
}
static VecW interp_vecw(VecW v1, VecW v2, VecW v3, VecW b) { /* vertex.go:67-73 */
    VecW n = {0, 0, 0, 0};
    n = wadd(n, wmuls(v1, b.X));
    n = wadd(n, wmuls(v2, b.Y));
    n = wadd(n, wmuls(v3, b.Z));
    return wmuls(n, b.W);
}
/* vertex.go:18-47 (all fields, including those no built-in shader reads) */
static Vertex interpolate_vertexes(const Vertex *v1, const Vertex *v2, const Vertex *v3, VecW b) {
    Vertex v;
    v.Position = interp_vec(v1->Position, v2->Position, v3->Position, b);
    v.Normal = vnormalize(interp_vec(v1->Normal, v2->Normal, v3->Normal, b));
    v.Texture = interp_vec(v1->Texture, v2->Texture, v3->Texture, b);
    v.Color = interp_col(v1->Color, v2->Color, v3->Color, b);
    v.Output = interp_vecw(v1->Output, v2->Output, v3->Output, b);
    return v;
}
static VecW barycentric(Vec p1, Vec p2, Vec p3, Vec p) { /* vertex.go:81-95 */
    Vec v0 = vsub(p2, p1);
    Vec v1 = vsub(p3, p1);
    Vec v2 = vsub(p, p1);
    double d00 = vdot(v0, v0);
    double d01 = vdot(v0, v1);
    double d11 = vdot(v1, v1);
    double d20 = vdot(v2, v0);
    double d21 = vdot(v2, v1);
    double d = d00 * d11 - d01 * d01;
    double v = (d11 * d20 - d01 * d21) / d;
    double w = (d00 * d21 - d01 * d20) / d;
    double u = ORC_MUTANT == 5 ? 1 - v + w : 1 - v - w;
    return (VecW){u, v, w, 1};
}
static void fix_normals(Triangle *t) { /* triangle.go:33-37, 46-58 */
    Vec e1 = vsub(t->V2.Position, t->V1.Position);
    Vec e2 = vsub(t->V3.Position, t->V1.Position);
    Vec n = vnormalize(vcross(e1, e2));
    Vec zero = {0, 0, 0};
    if (vequal(t->V1.Normal, zero)) t->V1.Normal = n;
    if (vequal(t->V2.Normal, zero)) t->V2.Normal = n;
    if (vequal(t->V3.Normal, zero)) t->V3.Normal = n;
}

/* ------------------------------------------------------------------ */
/* clipping.go */

static const VecW clip_P[6] = {{1, 0, 0, 1}, {-1, 0, 0, 1}, {0, 1, 0, 1}, {0, -1, 0, 1}, {0, 0, 1, 1}, {0, 0, -1, 1}};
static const VecW clip_N[6] = {{-1, 0, 0, 1}, {1, 0, 0, 1}, {0, -1, 0, 1}, {0, 1, 0, 1}, {0, 0, -1, 1}, {0, 0, 1, 1}};

static int point_in_front(int p, VecW v) { /* 16-18 */
    if (ORC_MUTANT == 11 && p == 4) return wdot(wsub(v, clip_P[p]), clip_P[p]) > 0;
    return wdot(wsub(v, clip_P[p]), clip_N[p]) > 0;
}
static VecW intersect_segment(int p, VecW v0, VecW v1) {                                   /* 20-26 */
    VecW u = wsub(v1, v0);
    VecW w = wsub(v0, clip_P[p]);
    double d = wdot(clip_N[p], u);
    double n = -wdot(clip_N[p], w);
    return wadd(v0, wmuls(u, n / d));
}
/* 28-52; at most 9 output points for a triangle. */
static int sutherland_hodgman(const VecW *points, int npoints, VecW *out) {
    VecW buf[2][16];
    int n = npoints;
    for (int i = 0; i < n; i++) buf[0][i] = points[i];
    int cur = 0;
    for (int p = 0; p < 6; p++) {
        VecW *input = buf[cur];
        VecW *output = buf[cur ^ 1];
        int m = 0;
        if (n == 0) return 0;
        VecW s = input[n - 1];
        for (int k = 0; k < n; k++) {
            VecW e = input[k];
            if (point_in_front(p, e)) {
                if (!point_in_front(p, s)) output[m++] = intersect_segment(p, s, e);
                output[m++] = e;
            } else if (point_in_front(p, s)) {
                output[m++] = intersect_segment(p, s, e);
            }
            s = e;
        }
        n = m;
        cur ^= 1;
    }
    for (int i = 0; i < n; i++) out[i] = buf[cur][i];
    return n;
}
/* 54-74: returns the fan triangles (NewTriangle -> FixNormals each). */
static int clip_triangle(const Triangle *t, Triangle *out) {
    VecW w1 = t->V1.Output, w2 = t->V2.Output, w3 = t->V3.Output;
    Vec p1 = wvec(w1), p2 = wvec(w2), p3 = wvec(w3);
    VecW points[3] = {w1, w2, w3};
    VecW np[16];
    int n = sutherland_hodgman(points, 3, np);
    int k = 0;
    for (int i = 2; i < n; i++) {
        VecW b1 = barycentric(p1, p2, p3, wvec(np[0]));
        VecW b2 = barycentric(p1, p2, p3, wvec(np[i - 1]));
        VecW b3 = barycentric(p1, p2, p3, wvec(np[i]));
        Triangle r;
        r.V1 = interpolate_vertexes(&t->V1, &t->V2, &t->V3, b1);
        r.V2 = interpolate_vertexes(&t->V1, &t->V2, &t->V3, b2);
        r.V3 = interpolate_vertexes(&t->V1, &t->V2, &t->V3, b3);
        fix_normals(&r);
        out[k++] = r;
    }
    return k;
}
/* 76-98: attributes are not re-interpolated; returns 0 for nil. */
static int clip_line(const Line *l, Line *out) {
    VecW w1 = l->V1.Output, w2 = l->V2.Output;
    for (int p = 0; p < 6; p++) {
        int f1 = point_in_front(p, w1);
        int f2 = point_in_front(p, w2);
        if (f1 && f2) continue;
        else if (f1) w2 = intersect_segment(p, w1, w2);
        else if (f2) w1 = intersect_segment(p, w2, w1);
        else return 0;
    }
    out->V1 = l->V1;
    out->V2 = l->V2;
    out->V1.Output = w1;
    out->V2.Output = w2;
    return 1;
}

/* ------------------------------------------------------------------ */
/* texture.go */

static Col tex_at(const orc_texture *t, int64_t x, int64_t y) {
    /* image.Image.At(x,y).RGBA() then MakeColor (color.go:25-29). */
    uint32_t r, g, b, a;
    if (x < 0 || y < 0 || x >= t->width || y >= t->height) {
        r = g = b = a = 0;
    } else {
        int64_t i = (y * t->width + x) * 4;
        if (t->kind == ORC_TEX_RGBA16) {
            const uint16_t *p = (const uint16_t *)t->pix;
            r = p[i]; g = p[i + 1]; b = p[i + 2]; a = p[i + 3];
        } else {
            const uint8_t *p = (const uint8_t *)t->pix;
            NRGBA8 c = {p[i], p[i + 1], p[i + 2], p[i + 3]};
            if (t->kind == ORC_TEX_NRGBA8) {
                nrgba_rgba(c, &r, &g, &b, &a);
            } else { /* image.RGBA: already premultiplied */
                r = c.R; r |= r << 8;
                g = c.G; g |= g << 8;
                b = c.B; b |= b << 8;
                a = c.A; a |= a << 8;
            }
        }
    }
    const double d = 0xffff;
    return (Col){(double)r / d, (double)g / d, (double)b / d, (double)a / d};
}
static Col bilinear_sample(const orc_texture *t, double u, double v) { /* texture.go:41-63 */
    v = 1 - v;
    u -= floor(u);
    v -= floor(v);
    double x = u * (double)(t->width - 1);
    double y = v * (double)(t->height - 1);
    int64_t x0 = go_f2i(x);
    int64_t y0 = go_f2i(y);
    int64_t x1 = x0 + 1;
    int64_t y1 = y0 + 1;
    x -= (double)x0;
    y -= (double)y0;
    Col c00 = tex_at(t, x0, y0);
    Col c01 = ORC_MUTANT == 6 ? tex_at(t, x1, y0) : tex_at(t, x0, y1);
    Col c10 = tex_at(t, x1, y0);
    Col c11 = tex_at(t, x1, y1);
    Col c = {0, 0, 0, 0};
    c = cadd(c, cmuls(c00, (1 - x) * (1 - y)));
    c = cadd(c, cmuls(c10, x * (1 - y)));
    c = cadd(c, cmuls(c01, (1 - x) * y));
    c = cadd(c, cmuls(c11, x * y));
    return c;
}

/* ------------------------------------------------------------------ */
/* shader.go */

typedef struct {
    int kind;
    Mat Matrix;
    Col Color;
    Vec LightDirection, CameraPosition;
    Col ObjectColor, AmbientColor, DiffuseColor, SpecularColor;
    double SpecularPower;
    orc_texture tex;
    int has_tex;
} Shader;

static Vertex shader_vertex(const Shader *s, Vertex v) { /* shader.go:20-23, 39-42, 70-73 */
    v.Output = mat_mul_position_w(s->Matrix, v.Position);
    return v;
}
static Col phong_fragment(const Shader *s, const Vertex *v) { /* shader.go:75-96 */
    Col light = s->AmbientColor;
    Col color = v->Color;
    if (!cdiscard(s->ObjectColor)) color = s->ObjectColor;
    if (s->has_tex) color = bilinear_sample(&s->tex, v->Texture.X, v->Texture.Y);
    double diffuse = go_max(vdot(v->Normal, s->LightDirection), 0);
    light = cadd(light, cmuls(s->DiffuseColor, diffuse));
    if (diffuse > 0 && s->SpecularPower > 0) {
        Vec camera = vnormalize(vsub(s->CameraPosition, v->Position));
        Vec reflected = vreflect(vneg(s->LightDirection), v->Normal);
        double specular = go_max(vdot(camera, reflected), 0);
        if (specular > 0) {
            specular = go_pow(specular, s->SpecularPower);
            light = cadd(light, cmuls(s->SpecularColor, specular));
        }
    }
    Col o = cmin(cmul(color, light), (Col){1, 1, 1, 1});
    o.A = color.A;
    return o;
}
/* Fragment dispatch through a function pointer, like the reference's
 * interface call (context.go:240), so the full interpolated Vertex is
 * materialised per fragment as in the reference. */
typedef Col (*frag_fn)(const Shader *, const Vertex *);
static Col solid_fragment(const Shader *s, const Vertex *v) { (void)v; return s->Color; }           /* 25-27 */
static Col texture_fragment(const Shader *s, const Vertex *v) {                                     /* 44-46 */
    return bilinear_sample(&s->tex, v->Texture.X, v->Texture.Y);
}

/* ------------------------------------------------------------------ */
/* context.go */

struct orc_ctx {
    int32_t Width, Height;
    uint8_t *Pix;   /* ColorBuffer.Pix, stride 4*Width */
    double *Depth;  /* DepthBuffer */
    Shader shader;
    frag_fn fragment;
    int ReadDepth, WriteDepth, WriteColor, AlphaBlend, Wireframe;
    int FrontFace, Cull;
    double LineWidth, DepthBias;
    Mat screenMatrix;
    pthread_mutex_t locks[256];
    int threaded;
    uint64_t counters[ORC_CNT_N];
};

typedef struct {
    uint64_t total, updated;
} Info;

typedef struct {
    orc_ctx *dc;
    uint64_t cnt[ORC_CNT_N];
} Worker;

static double load_depth(const orc_ctx *dc, int64_t i) {
    /* The reference reads the depth buffer without a lock (context.go:232);
     * relaxed atomic access keeps the threaded mirror free of UB. */
    uint64_t bits = __atomic_load_n((const uint64_t *)&dc->Depth[i], __ATOMIC_RELAXED);
    double d;
    memcpy(&d, &bits, 8);
    return d;
}
static void store_depth(orc_ctx *dc, int64_t i, double d) {
    uint64_t bits;
    memcpy(&bits, &d, 8);
    __atomic_store_n((uint64_t *)&dc->Depth[i], bits, __ATOMIC_RELAXED);
}

/* context.go:147-149 */
static double edge(Vec a, Vec b, Vec c) {
    if (ORC_MUTANT == 7) return (b.X - c.X) * (a.Y - c.Y) - (b.Y - c.Y) * (a.X - b.X);
    return (b.X - c.X) * (a.Y - c.Y) - (b.Y - c.Y) * (a.X - c.X);
}

/* context.go:151-281 */
static Info rasterize(Worker *wk, const Vertex *v0, const Vertex *v1, const Vertex *v2, Vec s0, Vec s1, Vec s2) {
    orc_ctx *dc = wk->dc;
    Info info = {0, 0};
    wk->cnt[ORC_CNT_RASTER]++;

    /* integer bounding box */
    Vec mn = vfloor(vmin(s0, vmin(s1, s2)));
    Vec mx = vceil(vmax(s0, vmax(s1, s2)));
    int64_t x0 = go_f2i(mn.X);
    int64_t x1 = go_f2i(mx.X);
    int64_t y0 = go_f2i(mn.Y);
    int64_t y1 = go_f2i(mx.Y);
    if (x0 < -ORC_BBOX_LIM || x0 > ORC_BBOX_LIM || x1 < -ORC_BBOX_LIM || x1 > ORC_BBOX_LIM ||
        y0 < -ORC_BBOX_LIM || y0 > ORC_BBOX_LIM || y1 < -ORC_BBOX_LIM || y1 > ORC_BBOX_LIM) {
        wk->cnt[ORC_CNT_SKIPPED]++;
        return info;
    }

    /* forward differencing variables */
    Vec p = {(double)x0 + 0.5, (double)y0 + 0.5, 0};
    double w00 = edge(s1, s2, p);
    double w01 = edge(s2, s0, p);
    double w02 = edge(s0, s1, p);
    double a01 = s1.Y - s0.Y;
    double b01 = s0.X - s1.X;
    double a12 = s2.Y - s1.Y;
    double b12 = s1.X - s2.X;
    double a20 = s0.Y - s2.Y;
    double b20 = s2.X - s0.X;

    /* reciprocals */
    double ra = 1 / edge(s0, s1, s2);
    double r0 = 1 / v0->Output.W;
    double r1 = 1 / v1->Output.W;
    double r2 = 1 / v2->Output.W;
    double ra12 = 1 / a12;
    double ra20 = 1 / a20;
    double ra01 = 1 / a01;

    const int64_t W = dc->Width;
    const int64_t npix = (int64_t)dc->Width * dc->Height;
    for (int64_t y = y0; y <= y1; y++) {
        wk->cnt[ORC_CNT_ROWS]++;
        double d = 0;
        double d0 = -w00 * ra12;
        double d1 = -w01 * ra20;
        double d2 = -w02 * ra01;
        if (w00 < 0 && d0 > d) d = d0;
        if (w01 < 0 && d1 > d) d = d1;
        if (w02 < 0 && d2 > d) d = d2;
        d = (double)go_f2i(d);
        if (d < 0) d = 0; /* occurs in pathological cases */
        double w0 = w00 + a12 * d;
        double w1 = w01 + a20 * d;
        double w2 = w02 + a01 * d;
        int wasInside = 0;
        int64_t id = go_f2i(d);
        /* x0 + int(d) > x1 (also guards the int64 wrap of the sum) */
        int64_t xs = (id > x1 - x0) ? x1 + 1 : x0 + id;
        for (int64_t x = xs; x <= x1; x++) {
            wk->cnt[ORC_CNT_SCANNED]++;
            double b0 = w0 * ra;
            double b1 = w1 * ra;
            double b2 = w2 * ra;
            w0 += a12;
            w1 += a20;
            w2 += a01;
            /* check if inside triangle */
            if (b0 < 0 || b1 < 0 || b2 < 0) {
                if (wasInside) break;
                continue;
            }
            wasInside = 1;
            /* check depth buffer for early abort */
            int64_t i = y * W + x;
            if (i < 0 || i >= npix) continue;
            info.total++;
            wk->cnt[ORC_CNT_INSIDE]++;
            double z = b0 * s0.Z + b1 * s1.Z + b2 * s2.Z;
            double bz = z + dc->DepthBias;
            if (dc->ReadDepth && bz > load_depth(dc, i)) continue;
            /* perspective-correct interpolation of vertex data */
            VecW b = {b0 * r0, b1 * r1, b2 * r2, 0};
            b.W = 1 / (b.X + b.Y + b.Z);
            Vertex v = interpolate_vertexes(v0, v1, v2, b);
            /* invoke fragment shader */
            wk->cnt[ORC_CNT_SHADED]++;
            Col color = dc->fragment(&dc->shader, &v);
            if (cdiscard(color)) continue;
            /* update buffers atomically */
            pthread_mutex_t *lock = &dc->locks[(x + y) & 255];
            if (dc->threaded) pthread_mutex_lock(lock);
            if ((ORC_MUTANT == 8 ? bz < load_depth(dc, i) : bz <= load_depth(dc, i)) || !dc->ReadDepth) {
                info.updated++;
                if (dc->WriteDepth) store_depth(dc, i, z);
                if (dc->WriteColor) {
                    if (dc->AlphaBlend && color.A < 1) {
                        uint32_t sr, sg, sb, sa;
                        nrgba_rgba(cnrgba(color), &sr, &sg, &sb, &sa);
                        uint32_t a = ORC_MUTANT == 9 ? (0xffff - sa) : (0xffff - sa) * 0x101;
                        int64_t j = (y * W + x) * 4; /* PixOffset(x, y) */
                        uint8_t *dr = &dc->Pix[j + 0];
                        uint8_t *dg = &dc->Pix[j + 1];
                        uint8_t *db = &dc->Pix[j + 2];
                        uint8_t *da = &dc->Pix[j + 3];
                        *dr = (uint8_t)(((uint32_t)*dr * a / 0xffff + sr) >> 8);
                        *dg = (uint8_t)(((uint32_t)*dg * a / 0xffff + sg) >> 8);
                        *db = (uint8_t)(((uint32_t)*db * a / 0xffff + sb) >> 8);
                        *da = (uint8_t)(((uint32_t)*da * a / 0xffff + sa) >> 8);
                    } else if (x >= 0 && y >= 0 && x < W && y < dc->Height) {
                        /* SetNRGBA bounds-checks x and y */
                        NRGBA8 c = cnrgba(color);
                        int64_t j = (y * W + x) * 4;
                        dc->Pix[j + 0] = c.R;
                        dc->Pix[j + 1] = c.G;
                        dc->Pix[j + 2] = c.B;
                        dc->Pix[j + 3] = c.A;
                    }
                }
            }
            if (dc->threaded) pthread_mutex_unlock(lock);
        }
        w00 += b12;
        w01 += b20;
        w02 += b01;
    }
    return info;
}

static Info info_add(Info a, Info b) { return (Info){a.total + b.total, a.updated + b.updated}; }

/* context.go:283-294 */
static Info line(Worker *wk, const Vertex *v0, const Vertex *v1, Vec s0, Vec s1) {
    orc_ctx *dc = wk->dc;
    Vec n = vmuls(vperpendicular(vsub(s1, s0)), dc->LineWidth / 2);
    s0 = vadd(s0, vmuls(vnormalize(vsub(s0, s1)), dc->LineWidth / 2));
    s1 = vadd(s1, vmuls(vnormalize(vsub(s1, s0)), dc->LineWidth / 2));
    Vec s00 = vadd(s0, n);
    Vec s01 = vsub(s0, n);
    Vec s10 = vadd(s1, n);
    Vec s11 = vsub(s1, n);
    Info info1 = rasterize(wk, v1, v0, v0, s11, s01, s00);
    Info info2 = rasterize(wk, v1, v1, v0, s10, s11, s00);
    return info_add(info1, info2);
}

/* context.go:296-301 */
static Info wireframe(Worker *wk, const Vertex *v0, const Vertex *v1, const Vertex *v2, Vec s0, Vec s1, Vec s2) {
    Info info1 = line(wk, v0, v1, s0, s1);
    Info info2 = line(wk, v1, v2, s1, s2);
    Info info3 = line(wk, v2, v0, s2, s0);
    return info_add(info_add(info1, info2), info3);
}

/* context.go:303-314 */
static Info draw_clipped_line(Worker *wk, const Vertex *v0, const Vertex *v1) {
    orc_ctx *dc = wk->dc;
    Vec ndc0 = wvec(wdivs(v0->Output, v0->Output.W));
    Vec ndc1 = wvec(wdivs(v1->Output, v1->Output.W));
    Vec s0 = mat_mul_position(dc->screenMatrix, ndc0);
    Vec s1 = mat_mul_position(dc->screenMatrix, ndc1);
    return line(wk, v0, v1, s0, s1);
}

/* context.go:316-349 */
static Info draw_clipped_triangle(Worker *wk, const Vertex *v0in, const Vertex *v1in, const Vertex *v2in) {
    orc_ctx *dc = wk->dc;
    const Vertex *v0 = v0in, *v1 = v1in, *v2 = v2in;
    Vec ndc0 = wvec(wdivs(v0->Output, v0->Output.W));
    Vec ndc1 = wvec(wdivs(v1->Output, v1->Output.W));
    Vec ndc2 = wvec(wdivs(v2->Output, v2->Output.W));
    /* back face culling */
    double a = (ndc1.X - ndc0.X) * (ndc2.Y - ndc0.Y) - (ndc2.X - ndc0.X) * (ndc1.Y - ndc0.Y);
    if (a < 0 && ORC_MUTANT != 10) {
        const Vertex *tv = v0; v0 = v2; v2 = tv;
        Vec tn = ndc0; ndc0 = ndc2; ndc2 = tn;
    }
    if (dc->Cull == ORC_CULL_FRONT) a = -a;
    if (dc->FrontFace == ORC_FACE_CW) a = -a;
    if (dc->Cull != ORC_CULL_NONE && a <= 0) return (Info){0, 0};
    /* screen coordinates */
    Vec s0 = mat_mul_position(dc->screenMatrix, ndc0);
    Vec s1 = mat_mul_position(dc->screenMatrix, ndc1);
    Vec s2 = mat_mul_position(dc->screenMatrix, ndc2);
    if (dc->Wireframe) return wireframe(wk, v0, v1, v2, s0, s1, s2);
    return rasterize(wk, v0, v1, v2, s0, s1, s2);
}

/* context.go:351-368 */
static Info draw_line(Worker *wk, const Line *t) {
    orc_ctx *dc = wk->dc;
    wk->cnt[ORC_CNT_INPUT]++;
    Vertex v1 = shader_vertex(&dc->shader, t->V1);
    Vertex v2 = shader_vertex(&dc->shader, t->V2);
    if (woutside(v1.Output) || woutside(v2.Output)) {
        wk->cnt[ORC_CNT_CLIPPED]++;
        Line in = {v1, v2}, out;
        if (clip_line(&in, &out)) return draw_clipped_line(wk, &out.V1, &out.V2);
        return (Info){0, 0};
    }
    return draw_clipped_line(wk, &v1, &v2);
}

/* context.go:370-389 */
static Info draw_triangle(Worker *wk, const Triangle *t) {
    orc_ctx *dc = wk->dc;
    wk->cnt[ORC_CNT_INPUT]++;
    Vertex v1 = shader_vertex(&dc->shader, t->V1);
    Vertex v2 = shader_vertex(&dc->shader, t->V2);
    Vertex v3 = shader_vertex(&dc->shader, t->V3);
    if (woutside(v1.Output) || woutside(v2.Output) || woutside(v3.Output)) {
        wk->cnt[ORC_CNT_CLIPPED]++;
        Triangle nt = {v1, v2, v3};
        fix_normals(&nt); /* NewTriangle */
        Triangle fan[8];
        int n = clip_triangle(&nt, fan);
        Info result = {0, 0};
        for (int k = 0; k < n; k++) result = info_add(result, draw_clipped_triangle(wk, &fan[k].V1, &fan[k].V2, &fan[k].V3));
        return result;
    }
    return draw_clipped_triangle(wk, &v1, &v2, &v3);
}

/* ------------------------------------------------------------------ */
/* meshes and the DrawTriangles / DrawLines fan-out (context.go:391-439) */

struct orc_mesh {
    int64_t ntri, nline;
    Triangle **tris; /* []*Triangle */
    Line **lines;    /* []*Line */
};

static void load_vertex(Vertex *v, int64_t k, const double *pos, const double *nrm, const double *uv,
                        const double *col) {
    memset(v, 0, sizeof *v);
    if (pos) v->Position = (Vec){pos[3 * k], pos[3 * k + 1], pos[3 * k + 2]};
    if (nrm) v->Normal = (Vec){nrm[3 * k], nrm[3 * k + 1], nrm[3 * k + 2]};
    if (uv) v->Texture = (Vec){uv[2 * k], uv[2 * k + 1], 0};
    if (col) v->Color = (Col){col[4 * k], col[4 * k + 1], col[4 * k + 2], col[4 * k + 3]};
}

orc_mesh *orc_mesh_new(int64_t ntri, const double *tpos, const double *tnrm, const double *tuv,
                       const double *tcol, int64_t nline, const double *lpos, const double *lnrm,
                       const double *luv, const double *lcol) {
    orc_mesh *m = (orc_mesh *)calloc(1, sizeof *m);
    m->ntri = ntri;
    m->nline = nline;
    m->tris = (Triangle **)calloc(ntri > 0 ? ntri : 1, sizeof(Triangle *));
    m->lines = (Line **)calloc(nline > 0 ? nline : 1, sizeof(Line *));
    for (int64_t i = 0; i < ntri; i++) {
        Triangle *t = (Triangle *)malloc(sizeof *t);
        load_vertex(&t->V1, 3 * i + 0, tpos, tnrm, tuv, tcol);
        load_vertex(&t->V2, 3 * i + 1, tpos, tnrm, tuv, tcol);
        load_vertex(&t->V3, 3 * i + 2, tpos, tnrm, tuv, tcol);
        m->tris[i] = t;
    }
    for (int64_t i = 0; i < nline; i++) {
        Line *l = (Line *)malloc(sizeof *l);
        load_vertex(&l->V1, 2 * i + 0, lpos, lnrm, luv, lcol);
        load_vertex(&l->V2, 2 * i + 1, lpos, lnrm, luv, lcol);
        m->lines[i] = l;
    }
    return m;
}

void orc_mesh_free(orc_mesh *m) {
    if (!m) return;
    for (int64_t i = 0; i < m->ntri; i++) free(m->tris[i]);
    for (int64_t i = 0; i < m->nline; i++) free(m->lines[i]);
    free(m->tris);
    free(m->lines);
    free(m);
}

typedef struct {
    Worker wk;
    const orc_mesh *m;
    int wi, wn, lines;
    Info result;
} Job;

static void *job_run(void *arg) {
    Job *j = (Job *)arg;
    Info result = {0, 0};
    if (j->lines) {
        for (int64_t i = 0; i < j->m->nline; i++)
            if (i % j->wn == j->wi) result = info_add(result, draw_line(&j->wk, j->m->lines[i]));
    } else {
        for (int64_t i = 0; i < j->m->ntri; i++)
            if (i % j->wn == j->wi) result = info_add(result, draw_triangle(&j->wk, j->m->tris[i]));
    }
    j->result = result;
    return NULL;
}

static Info draw_prims(orc_ctx *c, const orc_mesh *m, int nthreads, int lines) {
    int wn = nthreads < 1 ? 1 : nthreads;
    Job *jobs = (Job *)calloc(wn, sizeof(Job));
    pthread_t *th = (pthread_t *)calloc(wn, sizeof(pthread_t));
    c->threaded = wn > 1;
    for (int wi = 0; wi < wn; wi++) {
        jobs[wi].wk.dc = c;
        jobs[wi].m = m;
        jobs[wi].wi = wi;
        jobs[wi].wn = wn;
        jobs[wi].lines = lines;
    }
    if (wn == 1) {
        job_run(&jobs[0]);
    } else {
        for (int wi = 0; wi < wn; wi++) pthread_create(&th[wi], NULL, job_run, &jobs[wi]);
        for (int wi = 0; wi < wn; wi++) pthread_join(th[wi], NULL);
    }
    Info result = {0, 0};
    for (int wi = 0; wi < wn; wi++) {
        result = info_add(result, jobs[wi].result);
        for (int k = 0; k < ORC_CNT_N; k++) c->counters[k] += jobs[wi].wk.cnt[k];
    }
    free(jobs);
    free(th);
    return result;
}

void orc_draw_triangles(orc_ctx *c, const orc_mesh *m, int32_t nthreads, uint64_t info[2]) {
    Info r = draw_prims(c, m, nthreads, 0);
    info[0] = r.total;
    info[1] = r.updated;
}
void orc_draw_lines(orc_ctx *c, const orc_mesh *m, int32_t nthreads, uint64_t info[2]) {
    Info r = draw_prims(c, m, nthreads, 1);
    info[0] = r.total;
    info[1] = r.updated;
}
/* context.go:435-439 */
void orc_draw_mesh(orc_ctx *c, const orc_mesh *m, int32_t nthreads, uint64_t info[2]) {
    Info a = draw_prims(c, m, nthreads, 0);
    Info b = draw_prims(c, m, nthreads, 1);
    info[0] = a.total + b.total;
    info[1] = a.updated + b.updated;
}

/* ------------------------------------------------------------------ */
/* context construction, clears, readback */

orc_ctx *orc_ctx_new(int32_t width, int32_t height) { /* context.go:60-81 */
    orc_ctx *c = (orc_ctx *)calloc(1, sizeof *c);
    c->Width = width;
    c->Height = height;
    size_t n = (size_t)width * (size_t)height;
    c->Pix = (uint8_t *)calloc(n * 4 + 1, 1);
    c->Depth = (double *)malloc((n + 1) * sizeof(double));
    c->shader.kind = ORC_SHADER_SOLID;
    for (int k = 0; k < 16; k++) c->shader.Matrix.m[k] = (k % 5 == 0) ? 1 : 0;
    c->shader.Color = (Col){1, 0, 1, 1};
    c->fragment = solid_fragment;
    c->ReadDepth = c->WriteDepth = c->WriteColor = c->AlphaBlend = 1;
    c->Wireframe = 0;
    c->FrontFace = ORC_FACE_CCW;
    c->Cull = ORC_CULL_BACK;
    c->LineWidth = 2;
    c->DepthBias = 0;
    c->screenMatrix = mat_screen(width, height);
    for (int k = 0; k < 256; k++) pthread_mutex_init(&c->locks[k], NULL);
    orc_clear_depth(c, DBL_MAX);
    return c;
}

void orc_ctx_free(orc_ctx *c) {
    if (!c) return;
    for (int k = 0; k < 256; k++) pthread_mutex_destroy(&c->locks[k]);
    free(c->Pix);
    free(c->Depth);
    free(c);
}

void orc_set_state(orc_ctx *c, const orc_state *s) {
    c->ReadDepth = s->read_depth != 0;
    c->WriteDepth = s->write_depth != 0;
    c->WriteColor = s->write_color != 0;
    c->AlphaBlend = s->alpha_blend != 0;
    c->Wireframe = s->wireframe != 0;
    c->FrontFace = s->front_face;
    c->Cull = s->cull;
    c->LineWidth = s->line_width;
    c->DepthBias = s->depth_bias;
}

void orc_set_shader(orc_ctx *c, const orc_shader *s) {
    Shader *d = &c->shader;
    memset(d, 0, sizeof *d);
    d->kind = s->kind;
    memcpy(d->Matrix.m, s->matrix, sizeof d->Matrix.m);
    d->Color = (Col){s->color[0], s->color[1], s->color[2], s->color[3]};
    d->LightDirection = (Vec){s->light_dir[0], s->light_dir[1], s->light_dir[2]};
    d->CameraPosition = (Vec){s->camera_pos[0], s->camera_pos[1], s->camera_pos[2]};
    d->ObjectColor = (Col){s->object_color[0], s->object_color[1], s->object_color[2], s->object_color[3]};
    d->AmbientColor = (Col){s->ambient[0], s->ambient[1], s->ambient[2], s->ambient[3]};
    d->DiffuseColor = (Col){s->diffuse[0], s->diffuse[1], s->diffuse[2], s->diffuse[3]};
    d->SpecularColor = (Col){s->specular[0], s->specular[1], s->specular[2], s->specular[3]};
    d->SpecularPower = s->specular_power;
    if (s->texture) {
        d->tex = *s->texture;
        d->has_tex = 1;
    }
    if (s->kind == ORC_SHADER_TEXTURE) c->fragment = texture_fragment;
    else if (s->kind == ORC_SHADER_PHONG) c->fragment = phong_fragment;
    else c->fragment = solid_fragment;
}

void orc_clear_color(orc_ctx *c, const double rgba[4]) { /* context.go:119-131 */
    NRGBA8 v = cnrgba((Col){rgba[0], rgba[1], rgba[2], rgba[3]});
    size_t n = (size_t)c->Width * (size_t)c->Height;
    for (size_t i = 0; i < n; i++) {
        c->Pix[4 * i + 0] = v.R;
        c->Pix[4 * i + 1] = v.G;
        c->Pix[4 * i + 2] = v.B;
        c->Pix[4 * i + 3] = v.A;
    }
}

void orc_clear_depth(orc_ctx *c, double v) { /* context.go:137-141 */
    size_t n = (size_t)c->Width * (size_t)c->Height;
    for (size_t i = 0; i < n; i++) c->Depth[i] = v;
}

uint8_t *orc_color(orc_ctx *c) { return c->Pix; }
double *orc_depth(orc_ctx *c) { return c->Depth; }

void orc_depth_image(const orc_ctx *c, uint16_t *out) { /* context.go:87-117 */
    double lo = DBL_MAX;
    double hi = -DBL_MAX;
    size_t n = (size_t)c->Width * (size_t)c->Height;
    for (size_t i = 0; i < n; i++) {
        double d = c->Depth[i];
        if (d == DBL_MAX) continue;
        if (d < lo) lo = d;
        if (d > hi) hi = d;
    }
    for (size_t i = 0; i < n; i++) {
        double d = c->Depth[i];
        double t = (d - lo) / (hi - lo);
        if (d == DBL_MAX) t = 1;
        out[i] = go_u16(t * 0xffff);
    }
}

void orc_counters(const orc_ctx *c, uint64_t out[ORC_CNT_N]) { memcpy(out, c->counters, sizeof c->counters); }
void orc_reset_counters(orc_ctx *c) { memset(c->counters, 0, sizeof c->counters); }

/* ------------------------------------------------------------------ */
/* exported pieces for the pinning tests */

double orc_go_min(double a, double b) { return go_min(a, b); }
double orc_go_max(double a, double b) { return go_max(a, b); }
int64_t orc_go_f2i(double x) { return go_f2i(x); }
double orc_go_pow(double x, double y) { return go_pow(x, y); }
uint8_t orc_go_u8(double x) { return go_u8(x); }

void orc_color_nrgba(const double c[4], uint8_t out[4]) {
    NRGBA8 v = cnrgba((Col){c[0], c[1], c[2], c[3]});
    out[0] = v.R; out[1] = v.G; out[2] = v.B; out[3] = v.A;
}
void orc_nrgba_rgba(const uint8_t c[4], uint32_t out[4]) {
    NRGBA8 v = {c[0], c[1], c[2], c[3]};
    nrgba_rgba(v, &out[0], &out[1], &out[2], &out[3]);
}
void orc_blend(uint8_t dst[4], const double color[4]) {
    uint32_t sr, sg, sb, sa;
    nrgba_rgba(cnrgba((Col){color[0], color[1], color[2], color[3]}), &sr, &sg, &sb, &sa);
    uint32_t a = ORC_MUTANT == 9 ? (0xffff - sa) : (0xffff - sa) * 0x101;
    dst[0] = (uint8_t)(((uint32_t)dst[0] * a / 0xffff + sr) >> 8);
    dst[1] = (uint8_t)(((uint32_t)dst[1] * a / 0xffff + sg) >> 8);
    dst[2] = (uint8_t)(((uint32_t)dst[2] * a / 0xffff + sb) >> 8);
    dst[3] = (uint8_t)(((uint32_t)dst[3] * a / 0xffff + sa) >> 8);
}
void orc_barycentric(const double p1[3], const double p2[3], const double p3[3], const double p[3], double out[4]) {
    VecW b = barycentric((Vec){p1[0], p1[1], p1[2]}, (Vec){p2[0], p2[1], p2[2]}, (Vec){p3[0], p3[1], p3[2]},
                         (Vec){p[0], p[1], p[2]});
    out[0] = b.X; out[1] = b.Y; out[2] = b.Z; out[3] = b.W;
}
int32_t orc_clip_polygon(const double in[3][4], double out[9][4]) {
    VecW pts[3], res[16];
    for (int i = 0; i < 3; i++) pts[i] = (VecW){in[i][0], in[i][1], in[i][2], in[i][3]};
    int n = sutherland_hodgman(pts, 3, res);
    for (int i = 0; i < n && i < 9; i++) {
        out[i][0] = res[i].X; out[i][1] = res[i].Y; out[i][2] = res[i].Z; out[i][3] = res[i].W;
    }
    return n;
}
int32_t orc_clip_line(const double w1[4], const double w2[4], double out[2][4]) {
    Line l, o;
    memset(&l, 0, sizeof l);
    l.V1.Output = (VecW){w1[0], w1[1], w1[2], w1[3]};
    l.V2.Output = (VecW){w2[0], w2[1], w2[2], w2[3]};
    if (!clip_line(&l, &o)) return 0;
    out[0][0] = o.V1.Output.X; out[0][1] = o.V1.Output.Y; out[0][2] = o.V1.Output.Z; out[0][3] = o.V1.Output.W;
    out[1][0] = o.V2.Output.X; out[1][1] = o.V2.Output.Y; out[1][2] = o.V2.Output.Z; out[1][3] = o.V2.Output.W;
    return 1;
}
void orc_matrix_mul(const double a[16], const double b[16], double out[16]) {
    Mat x, y;
    memcpy(x.m, a, sizeof x.m);
    memcpy(y.m, b, sizeof y.m);
    Mat o = mat_mul(x, y);
    memcpy(out, o.m, sizeof o.m);
}
void orc_mul_position_w(const double m[16], const double p[3], double out[4]) {
    Mat x;
    memcpy(x.m, m, sizeof x.m);
    VecW o = mat_mul_position_w(x, (Vec){p[0], p[1], p[2]});
    out[0] = o.X; out[1] = o.Y; out[2] = o.Z; out[3] = o.W;
}
void orc_look_at(const double eye[3], const double center[3], const double up[3], double out[16]) {
    Mat o = mat_look_at((Vec){eye[0], eye[1], eye[2]}, (Vec){center[0], center[1], center[2]},
                        (Vec){up[0], up[1], up[2]});
    memcpy(out, o.m, sizeof o.m);
}
void orc_frustum(double l, double r, double b, double t, double n, double f, double out[16]) {
    Mat o = mat_frustum(l, r, b, t, n, f);
    memcpy(out, o.m, sizeof o.m);
}
void orc_perspective(double fovy, double aspect, double near, double far, double out[16]) {
    Mat o = mat_perspective(fovy, aspect, near, far);
    memcpy(out, o.m, sizeof o.m);
}
void orc_screen(int32_t w, int32_t h, double out[16]) {
    Mat o = mat_screen(w, h);
    memcpy(out, o.m, sizeof o.m);
}
void orc_bilinear_sample(const orc_texture *t, double u, double v, double out[4]) {
    Col c = bilinear_sample(t, u, v);
    out[0] = c.R; out[1] = c.G; out[2] = c.B; out[3] = c.A;
}
void orc_phong_fragment(const orc_shader *s, const double pos[3], const double nrm[3], const double uv[2],
                        const double col[4], double out[4]) {
    orc_ctx tmp;
    memset(&tmp, 0, sizeof tmp);
    orc_set_shader(&tmp, s);
    Vertex v;
    memset(&v, 0, sizeof v);
    v.Position = (Vec){pos[0], pos[1], pos[2]};
    v.Normal = (Vec){nrm[0], nrm[1], nrm[2]};
    v.Texture = (Vec){uv[0], uv[1], 0};
    v.Color = (Col){col[0], col[1], col[2], col[3]};
    Col c = phong_fragment(&tmp.shader, &v);
    out[0] = c.R; out[1] = c.G; out[2] = c.B; out[3] = c.A;
}
void orc_rasterize_screen(orc_ctx *c, const double s0[3], const double s1[3], const double s2[3],
                          const double w[3], uint64_t info[2]) {
    Worker wk;
    memset(&wk, 0, sizeof wk);
    wk.dc = c;
    c->threaded = 0;
    Vertex v[3];
    memset(v, 0, sizeof v);
    for (int k = 0; k < 3; k++) v[k].Output.W = w[k];
    Info r = rasterize(&wk, &v[0], &v[1], &v[2], (Vec){s0[0], s0[1], s0[2]}, (Vec){s1[0], s1[1], s1[2]},
                       (Vec){s2[0], s2[1], s2[2]});
    for (int k = 0; k < ORC_CNT_N; k++) c->counters[k] += wk.cnt[k];
    info[0] = r.total;
    info[1] = r.updated;
}

/* ---------------------------------------------------------------------
 * SSAA resolve (filter defined by this build; see the header).
 */

static int32_t tent_table(int32_t src, int32_t dst, int16_t *w, int32_t *start) {
    const double s = (double)src / (double)dst;
    const double factor = 1.0 / s < 1.0 ? 1.0 / s : 1.0;
    const int32_t half = (int32_t)ceil(s) < 1 ? 1 : (int32_t)ceil(s);
    const int32_t taps = 2 * half;
    if (!w) return taps;
    for (int32_t o = 0; o < dst; o++) {
        const double centre = s * ((double)o + 0.5) - 0.5;
        const int32_t st = (int32_t)centre - half + (ORC_MUTANT == 12 ? 0 : 1);
        start[o] = st;
        for (int32_t t = 0; t < taps; t++) {
            double x = (centre - (double)st - (double)t) * factor;
            if (x < 0) x = -x;
            const double k = x <= 1 ? 1 - x : 0;
            w[(size_t)o * taps + t] = (int16_t)(k * 256);
        }
    }
    return taps;
}

int32_t orc_tent_weights(int32_t src, int32_t dst, int16_t *w, int32_t *start) {
    return tent_table(src, dst, w, start);
}

static int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

void orc_resolve(const uint8_t *src, int32_t sw, int32_t sh, uint8_t *dst, int32_t dw, int32_t dh, int32_t filter) {
    if (filter == 2) {
        const int32_t f = sw / dw, n = f * f;
        for (int32_t y = 0; y < dh; y++)
            for (int32_t x = 0; x < dw; x++)
                for (int ch = 0; ch < 4; ch++) {
                    uint32_t acc = 0;
                    for (int32_t j = 0; j < f; j++)
                        for (int32_t i = 0; i < f; i++) acc += src[((size_t)(y * f + j) * sw + (size_t)(x * f + i)) * 4 + ch];
                    dst[((size_t)y * dw + x) * 4 + ch] = (uint8_t)((acc + (ORC_MUTANT == 14 ? 0u : (uint32_t)n / 2)) / (uint32_t)n);
                }
        return;
    }
    const int32_t tx = tent_table(sw, dw, NULL, NULL), ty = tent_table(sh, dh, NULL, NULL);
    int16_t *wx = malloc(sizeof(int16_t) * (size_t)dw * tx), *wy = malloc(sizeof(int16_t) * (size_t)dh * ty);
    int32_t *ox = malloc(sizeof(int32_t) * (size_t)dw), *oy = malloc(sizeof(int32_t) * (size_t)dh);
    uint8_t *mid = malloc((size_t)dw * sh * 4);  /* horizontal pass, premultiplied */
    tent_table(sw, dw, wx, ox);
    tent_table(sh, dh, wy, oy);
    for (int32_t y = 0; y < sh; y++)
        for (int32_t x = 0; x < dw; x++) {
            int32_t acc[4] = {0, 0, 0, 0}, wsum = 0;
            for (int32_t t = 0; t < tx; t++) {
                const int32_t c = wx[(size_t)x * tx + t];
                const uint8_t *q = src + ((size_t)y * sw + clampi(ox[x] + t, 0, sw - 1)) * 4;
                const int32_t a = q[3];
                for (int ch = 0; ch < 3; ch++) acc[ch] += c * (ORC_MUTANT == 13 ? (int32_t)q[ch] : (int32_t)q[ch] * a / 255);
                acc[3] += c * a;
                wsum += c;
            }
            for (int ch = 0; ch < 4; ch++)
                mid[((size_t)y * dw + x) * 4 + ch] = (uint8_t)clampi(wsum ? acc[ch] / wsum : 0, 0, 255);
        }
    for (int32_t y = 0; y < dh; y++)
        for (int32_t x = 0; x < dw; x++) {
            int32_t acc[4] = {0, 0, 0, 0}, wsum = 0;
            for (int32_t t = 0; t < ty; t++) {
                const int32_t c = wy[(size_t)y * ty + t];
                const uint8_t *q = mid + ((size_t)clampi(oy[y] + t, 0, sh - 1) * dw + x) * 4;
                for (int ch = 0; ch < 4; ch++) acc[ch] += c * (int32_t)q[ch];
                wsum += c;
            }
            for (int ch = 0; ch < 4; ch++)
                dst[((size_t)y * dw + x) * 4 + ch] = (uint8_t)clampi(wsum ? acc[ch] / wsum : 0, 0, 255);
        }
    free(wx);
    free(wy);
    free(ox);
    free(oy);
    free(mid);
}

/* ========================================================================
 * Mesh preparation (mesh.go)
 * ======================================================================== */

/* Box.Extend (box.go:48-53) with the triangle / line box (triangle.go:60-64,
 * line.go:17-21), folded in primitive order (mesh.go:153-165) */
static void box_extend_prims(const double *pos, int64_t nprim, int nv, double b[6]) {
    for (int64_t p = 0; p < nprim; p++) {
        const double *q = pos + 3 * nv * p;
        double mn[3] = {q[0], q[1], q[2]}, mx[3] = {q[0], q[1], q[2]};
        for (int k = 1; k < nv; k++)
            for (int c = 0; c < 3; c++) {
                mn[c] = go_min(mn[c], q[3 * k + c]);
                mx[c] = go_max(mx[c], q[3 * k + c]);
            }
        int empty = 1; /* a == EmptyBox: every field == 0 */
        for (int c = 0; c < 6; c++) empty &= b[c] == 0.0;
        if (empty) {
            for (int c = 0; c < 3; c++) {
                b[c] = mn[c];
                b[3 + c] = mx[c];
            }
        } else {
            for (int c = 0; c < 3; c++) {
                b[c] = go_min(b[c], mn[c]);
                b[3 + c] = go_max(b[3 + c], mx[c]);
            }
        }
    }
}

void orc_mesh_bounding_box(const double *tpos, int64_t ntri, const double *lpos, int64_t nline, double box[6]) {
    for (int c = 0; c < 6; c++) box[c] = 0.0; /* EmptyBox */
    if (tpos) box_extend_prims(tpos, ntri, 3, box);
    if (lpos) box_extend_prims(lpos, nline, 2, box);
}

/* MulPosition / MulDirection (matrix.go:209-229), Normalize (vector.go:83-86) */
void orc_mesh_transform(double *pos, double *nrm, int64_t nvert, const double a[16]) {
    for (int64_t v = 0; v < nvert; v++) {
        double *P = pos + 3 * v, *N = nrm + 3 * v;
        const double x = P[0], y = P[1], z = P[2];
        P[0] = a[0] * x + a[1] * y + a[2] * z + a[3];
        P[1] = a[4] * x + a[5] * y + a[6] * z + a[7];
        P[2] = a[8] * x + a[9] * y + a[10] * z + a[11];
        const double nx = N[0], ny = N[1], nz = N[2];
        const double dx = a[0] * nx + a[1] * ny + a[2] * nz;
        const double dy = a[4] * nx + a[5] * ny + a[6] * nz;
        const double dz = a[8] * nx + a[9] * ny + a[10] * nz;
        const double r = 1 / sqrt(dx * dx + dy * dy + dz * dz);
        N[0] = dx * r;
        N[1] = dy * r;
        N[2] = dz * r;
    }
}

static void mat_mul16(const double a[16], const double b[16], double m[16]) { /* matrix.go:188-207 */
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            m[4 * i + j] = a[4 * i] * b[j] + a[4 * i + 1] * b[4 + j] + a[4 * i + 2] * b[8 + j] + a[4 * i + 3] * b[12 + j];
}

void orc_mesh_fit_inside(double *tpos, double *tnrm, int64_t ntri, double *lpos, double *lnrm, int64_t nline,
                         const double box[6], const double anchor[3], double m_out[16]) {
    double bb[6];
    orc_mesh_bounding_box(tpos, ntri, lpos, nline, bb);
    Vec size = {bb[3] - bb[0], bb[4] - bb[1], bb[5] - bb[2]};          /* Box.Size, box.go:44-46 */
    Vec bsize = {box[3] - box[0], box[4] - box[1], box[5] - box[2]};
    Vec q = {bsize.X / size.X, bsize.Y / size.Y, bsize.Z / size.Z};  /* Vector.Div */
    const double scale = go_min(go_min(q.X, q.Y), q.Z);               /* MinComponent */
    Vec extra = {bsize.X - size.X * scale, bsize.Y - size.Y * scale, bsize.Z - size.Z * scale};
    const double I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    const double T1[16] = {1, 0, 0, -bb[0], 0, 1, 0, -bb[1], 0, 0, 1, -bb[2], 0, 0, 0, 1};
    const double S[16] = {scale, 0, 0, 0, 0, scale, 0, 0, 0, 0, scale, 0, 0, 0, 0, 1};
    const double T2[16] = {1, 0, 0, box[0] + extra.X * anchor[0], 0, 1, 0, box[1] + extra.Y * anchor[1],
                           0, 0, 1, box[2] + extra.Z * anchor[2], 0, 0, 0, 1};
    double m1[16], m2[16], m3[16];
    mat_mul16(T1, I, m1); /* Matrix.Translate(v) = Translate(v).Mul(m), matrix.go:143-145 */
    mat_mul16(S, m1, m2);
    mat_mul16(T2, m2, m3);
    if (tpos) orc_mesh_transform(tpos, tnrm, 3 * ntri, m3);
    if (lpos) orc_mesh_transform(lpos, lnrm, 2 * nline, m3);
    if (m_out) memcpy(m_out, m3, sizeof m3);
}

/* The map[Vector][]Vector of mesh.go:92-97 as groups of equal positions:
 * vertex indices sorted by position (-0 == +0; a position with a NaN is
 * its own key and is never found again), ties by index, so each group
 * lists its vertices in the order the reference appends them. */
static const double *g_sort_pos;
static int has_nan3(const double *p) { return p[0] != p[0] || p[1] != p[1] || p[2] != p[2]; }
static int cmp_vertex(const void *pa, const void *pb) {
    const int64_t a = *(const int64_t *)pa, b = *(const int64_t *)pb;
    const double *A = g_sort_pos + 3 * a, *B = g_sort_pos + 3 * b;
    const int na = has_nan3(A), nb = has_nan3(B);
    if (na != nb) return na - nb;
    if (!na)
        for (int c = 0; c < 3; c++) {
            if (A[c] < B[c]) return -1;
            if (A[c] > B[c]) return 1;
        }
    return a < b ? -1 : (a > b ? 1 : 0);
}
static int same_pos(const double *A, const double *B) { return A[0] == B[0] && A[1] == B[1] && A[2] == B[2]; }

static void smooth(const double *pos, double *nrm, int64_t ntri, int threshold_mode, double threshold) {
    const int64_t n = 3 * ntri;
    if (n == 0) return;
    int64_t *ord = malloc((size_t)n * sizeof *ord);
    double *out = malloc((size_t)n * 3 * sizeof *out);
    for (int64_t i = 0; i < n; i++) ord[i] = i;
    g_sort_pos = pos;
    qsort(ord, (size_t)n, sizeof *ord, cmp_vertex);
    int64_t g0 = 0;
    while (g0 < n) {
        int64_t g1 = g0 + 1;
        while (g1 < n && same_pos(pos + 3 * ord[g0], pos + 3 * ord[g1])) g1++;
        if (!same_pos(pos + 3 * ord[g0], pos + 3 * ord[g0])) {
            /* NaN key: its list is never found.  SmoothNormals reads the zero
             * Vector (mesh.go:121-123); the threshold version normalizes the
             * empty sum (mesh.go:81-88). */
            for (int64_t k = g0; k < g1; k++) {
                double *o = out + 3 * ord[k];
                if (threshold_mode) {
                    const double r = 1 / sqrt(0.0);
                    o[0] = 0.0 * r;
                    o[1] = 0.0 * r;
                    o[2] = 0.0 * r;
                } else {
                    o[0] = o[1] = o[2] = 0.0;
                }
            }
            g0 = g1;
            continue;
        }
        if (!threshold_mode) {
            /* lookup[p] = lookup[p].Add(n) in order, then Normalize (mesh.go:107-117) */
            Vec sum = {0, 0, 0};
            for (int64_t k = g0; k < g1; k++) {
                const double *x = nrm + 3 * ord[k];
                sum = (Vec){sum.X + x[0], sum.Y + x[1], sum.Z + x[2]};
            }
            const double r = 1 / sqrt(sum.X * sum.X + sum.Y * sum.Y + sum.Z * sum.Z);
            for (int64_t k = g0; k < g1; k++) {
                double *o = out + 3 * ord[k];
                o[0] = sum.X * r;
                o[1] = sum.Y * r;
                o[2] = sum.Z * r;
            }
        } else {
            /* smoothNormalsThreshold (mesh.go:80-89) per vertex of the group */
            for (int64_t k = g0; k < g1; k++) {
                const double *own = nrm + 3 * ord[k];
                Vec sum = {0, 0, 0};
                for (int64_t j = g0; j < g1; j++) {
                    const double *x = nrm + 3 * ord[j];
                    if (x[0] * own[0] + x[1] * own[1] + x[2] * own[2] >= threshold)
                        sum = (Vec){sum.X + x[0], sum.Y + x[1], sum.Z + x[2]};
                }
                const double r = 1 / sqrt(sum.X * sum.X + sum.Y * sum.Y + sum.Z * sum.Z);
                double *o = out + 3 * ord[k];
                o[0] = sum.X * r;
                o[1] = sum.Y * r;
                o[2] = sum.Z * r;
            }
        }
        g0 = g1;
    }
    memcpy(nrm, out, (size_t)n * 3 * sizeof *out);
    free(out);
    free(ord);
}

void orc_mesh_smooth_normals(const double *tpos, double *tnrm, int64_t ntri) { smooth(tpos, tnrm, ntri, 0, 0.0); }
void orc_mesh_smooth_normals_threshold(const double *tpos, double *tnrm, int64_t ntri, double threshold) {
    smooth(tpos, tnrm, ntri, 1, threshold);
}

/* ========================================================================
 * Mesh loaders (stl.go, obj.go)
 * ======================================================================== */

/* Triangle.Normal (triangle.go:33-37): Cross (vector.go:76-81), Normalize */
static Vec tri_normal(Vec a, Vec b, Vec c) {
    const Vec e1 = {b.X - a.X, b.Y - a.Y, b.Z - a.Z}, e2 = {c.X - a.X, c.Y - a.Y, c.Z - a.Z};
    const Vec n = {e1.Y * e2.Z - e1.Z * e2.Y, e1.Z * e2.X - e1.X * e2.Z, e1.X * e2.Y - e1.Y * e2.X};
    const double r = 1 / sqrt(n.X * n.X + n.Y * n.Y + n.Z * n.Z);
    return (Vec){n.X * r, n.Y * r, n.Z * r};
}

static double le_f32(const uint8_t *b) {
    uint32_t u = (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24);
    float f;
    memcpy(&f, &u, 4);
    return (double)f;
}

static int is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; }

/* strconv.ParseFloat(s, 64) with the error dropped (util.go:65-72) */
static double parse_float_z(char *buf, size_t n);
static double parse_float(const char *s, size_t n) {
    if (n == 0) return 0.0;
    char *buf = malloc(n + 1);
    memcpy(buf, s, n);
    buf[n] = 0;
    const double v = parse_float_z(buf, n);
    free(buf);
    return v;
}
static double parse_float_z(char *buf, size_t n) {
    (void)n;
    size_t i = 0;
    int neg = 0;
    if (buf[i] == '+' || buf[i] == '-') neg = buf[i++] == '-';
    if (!strcasecmp(buf + i, "inf") || !strcasecmp(buf + i, "infinity")) return neg ? -INFINITY : INFINITY;
    if (!strcasecmp(buf + i, "nan")) return NAN;
    if (buf[i] == '0' && (buf[i + 1] == 'x' || buf[i + 1] == 'X')) {
        /* hexadecimal mantissa with a mandatory binary exponent */
        size_t j = i + 2, nd = 0;
        while (isxdigit((unsigned char)buf[j])) j++, nd++;
        if (buf[j] == '.') {
            j++;
            while (isxdigit((unsigned char)buf[j])) j++, nd++;
        }
        if (nd == 0 || (buf[j] != 'p' && buf[j] != 'P')) return 0.0;
        size_t k = j + 1, ne = 0;
        if (buf[k] == '+' || buf[k] == '-') k++;
        while (buf[k] >= '0' && buf[k] <= '9') k++, ne++;
        if (ne == 0 || buf[k] != 0) return 0.0;
        return strtod(buf, NULL);
    }
    /* decimal: digits [. digits] or . digits, then [eE][+-]digits */
    size_t j = i, nd = 0;
    while (buf[j] >= '0' && buf[j] <= '9') j++, nd++;
    if (buf[j] == '.') {
        j++;
        while (buf[j] >= '0' && buf[j] <= '9') j++, nd++;
    }
    if (nd == 0) return 0.0;
    if (buf[j] == 'e' || buf[j] == 'E') {
        size_t k = j + 1, ne = 0;
        if (buf[k] == '+' || buf[k] == '-') k++;
        while (buf[k] >= '0' && buf[k] <= '9') k++, ne++;
        if (ne == 0) return 0.0;
        j = k;
    }
    if (buf[j] != 0) return 0.0;
    return strtod(buf, NULL);
}

/* strconv.ParseInt(s, 0, 0) with the error dropped (obj.go:10-17): sign,
 * 0x / 0o / 0b / leading-0 octal prefixes, underscores between digits */
static int64_t parse_int0(const char *s, size_t n) {
    size_t i = 0;
    int neg = 0;
    if (i < n && (s[i] == '+' || s[i] == '-')) neg = s[i++] == '-';
    if (i >= n) return 0;
    int base = 10;
    int prefixed = 0;
    if (s[i] == '0' && i + 1 < n) {
        const char c = s[i + 1];
        if (c == 'x' || c == 'X') base = 16, i += 2, prefixed = 1;
        else if (c == 'o' || c == 'O') base = 8, i += 2, prefixed = 1;
        else if (c == 'b' || c == 'B') base = 2, i += 2, prefixed = 1;
        else base = 8, i += 1, prefixed = 1;
    }
    /* underscores: each must sit between two digits, or right after a prefix */
    uint64_t v = 0;
    int nd = 0, prev_digit = prefixed;
    for (; i < n; i++) {
        const char c = s[i];
        if (c == '_') {
            if (!prev_digit) return 0;
            prev_digit = 0;
            continue;
        }
        int d;
        if (c >= '0' && c <= '9') d = c - '0';
        else if (c >= 'a' && c <= 'f') d = c - 'a' + 10;
        else if (c >= 'A' && c <= 'F') d = c - 'A' + 10;
        else return 0;
        if (d >= base) return 0;
        if (v > (UINT64_MAX - (uint64_t)d) / (uint64_t)base) v = UINT64_MAX; /* out of range */
        else v = v * (uint64_t)base + (uint64_t)d;
        nd++;
        prev_digit = 1;
    }
    if (!prev_digit) return 0;                /* trailing underscore or bare prefix */
    if (nd == 0 && !(base == 8 && prefixed)) return 0;
    if (!neg) return v > (uint64_t)INT64_MAX ? INT64_MAX : (int64_t)v;
    return v > (uint64_t)INT64_MAX + 1 ? INT64_MIN : (int64_t)(0 - v);
}

int orc_load_stl(const uint8_t *data, int64_t size, int64_t *ntri, double *pos, double *nrm) {
    if (size < 84) return -1; /* binary.Read of the header fails */
    const uint32_t count = (uint32_t)data[80] | ((uint32_t)data[81] << 8) | ((uint32_t)data[82] << 16) |
                           ((uint32_t)data[83] << 24);
    if ((int64_t)count * 50 + 84 == size) { /* loadSTLB (stl.go:86-154) */
        *ntri = count;
        if (!pos) return 0;
        for (int64_t i = 0; i < count; i++) {
            const uint8_t *r = data + 84 + 50 * i;
            Vec v[3];
            for (int k = 0; k < 3; k++) v[k] = (Vec){le_f32(r + 12 + 12 * k), le_f32(r + 16 + 12 * k), le_f32(r + 20 + 12 * k)};
            const Vec n = tri_normal(v[0], v[1], v[2]);
            for (int k = 0; k < 3; k++) {
                double *P = pos + 9 * i + 3 * k, *N = nrm + 9 * i + 3 * k;
                P[0] = v[k].X, P[1] = v[k].Y, P[2] = v[k].Z;
                N[0] = n.X, N[1] = n.Y, N[2] = n.Z;
            }
        }
        return 0;
    }
    /* loadSTLA (stl.go:59-84): lines with exactly 4 fields, "vertex x y z" */
    int64_t nv = 0;
    const char *t = (const char *)data;
    int64_t a = 0;
    while (a < size) {
        int64_t b = a;
        while (b < size && t[b] != '\n') b++;
        /* fields of line [a, b) */
        const char *f[5];
        size_t fl[5];
        int nf = 0;
        int64_t x = a;
        while (x < b) {
            while (x < b && is_space(t[x])) x++;
            if (x >= b) break;
            int64_t y = x;
            while (y < b && !is_space(t[y])) y++;
            if (nf < 5) f[nf] = t + x, fl[nf] = (size_t)(y - x);
            nf++;
            x = y;
        }
        if (nf == 4 && fl[0] == 6 && !memcmp(f[0], "vertex", 6)) {
            if (pos)
                for (int k = 0; k < 3; k++) pos[3 * nv + k] = parse_float(f[1 + k], fl[1 + k]);
            nv++;
        }
        a = b + 1;
    }
    if (nv % 3) return -1; /* the triangle loop indexes past the end */
    *ntri = nv / 3;
    if (!pos) return 0;
    for (int64_t i = 0; i < nv / 3; i++) {
        double *P = pos + 9 * i;
        const Vec n = tri_normal((Vec){P[0], P[1], P[2]}, (Vec){P[3], P[4], P[5]}, (Vec){P[6], P[7], P[8]});
        for (int k = 0; k < 3; k++) nrm[9 * i + 3 * k] = n.X, nrm[9 * i + 3 * k + 1] = n.Y, nrm[9 * i + 3 * k + 2] = n.Z;
    }
    return 0;
}

typedef struct {
    double *v;
    int64_t n, cap;
} VecList; /* n 3-vectors */

static void vl_push(VecList *l, double x, double y, double z) {
    if (l->n == l->cap) {
        l->cap = l->cap ? 2 * l->cap : 1024;
        l->v = realloc(l->v, (size_t)l->cap * 3 * sizeof(double));
    }
    l->v[3 * l->n] = x, l->v[3 * l->n + 1] = y, l->v[3 * l->n + 2] = z;
    l->n++;
}

int orc_load_obj(const char *t, int64_t size, int64_t *ntri, double *pos, double *nrm, double *uv) {
    VecList vs = {0}, vts = {0}, vns = {0};
    vl_push(&vs, 0, 0, 0); /* 1-based indexing (obj.go:25-27) */
    vl_push(&vts, 0, 0, 0);
    vl_push(&vns, 0, 0, 0);
    int64_t nt = 0;
    int rc = 0;
    int64_t a = 0;
    while (a < size && rc == 0) {
        int64_t b = a;
        while (b < size && t[b] != '\n') b++;
        /* fields: counted, then collected */
        int nf = 0;
        for (int64_t x = a; x < b;) {
            while (x < b && is_space(t[x])) x++;
            if (x >= b) break;
            while (x < b && !is_space(t[x])) x++;
            nf++;
        }
        const char **f = malloc((size_t)(nf + 1) * sizeof *f);
        size_t *fl = malloc((size_t)(nf + 1) * sizeof *fl);
        int64_t *iv = malloc((size_t)(nf + 1) * 3 * sizeof *iv), *it = iv + nf + 1, *in = it + nf + 1;
        nf = 0;
        for (int64_t x = a; x < b;) {
            while (x < b && is_space(t[x])) x++;
            if (x >= b) break;
            int64_t y = x;
            while (y < b && !is_space(t[y])) y++;
            f[nf] = t + x, fl[nf] = (size_t)(y - x);
            nf++;
            x = y;
        }
        a = b + 1;
        if (nf == 0) {
            free(f), free(fl), free(iv);
            continue;
        }
        const size_t kl = fl[0];
        if (kl == 1 && f[0][0] == 'v') {
            if (nf < 4) rc = -1;
            else
            vl_push(&vs, parse_float(f[1], fl[1]), parse_float(f[2], fl[2]), parse_float(f[3], fl[3]));
        } else if (kl == 2 && f[0][0] == 'v' && f[0][1] == 't') {
            if (nf < 3) rc = -1;
            else
            vl_push(&vts, parse_float(f[1], fl[1]), parse_float(f[2], fl[2]), 0.0);
        } else if (kl == 2 && f[0][0] == 'v' && f[0][1] == 'n') {
            if (nf < 4) rc = -1;
            else
            vl_push(&vns, parse_float(f[1], fl[1]), parse_float(f[2], fl[2]), parse_float(f[3], fl[3]));
        } else if (kl == 1 && f[0][0] == 'f') {
            const int na = nf - 1;
            for (int k = 0; k < na; k++) {
                /* strings.Split(arg + "//", "/"): the first three pieces */
                const char *s = f[1 + k];
                const size_t n = fl[1 + k];
                const char *piece[3] = {s, "", ""};
                size_t pl[3] = {n, 0, 0};
                size_t p0 = 0;
                while (p0 < n && s[p0] != '/') p0++;
                pl[0] = p0;
                if (p0 < n) {
                    size_t p1 = p0 + 1;
                    while (p1 < n && s[p1] != '/') p1++;
                    piece[1] = s + p0 + 1, pl[1] = p1 - p0 - 1;
                    if (p1 < n) {
                        size_t p2 = p1 + 1;
                        while (p2 < n && s[p2] != '/') p2++;
                        piece[2] = s + p1 + 1, pl[2] = p2 - p1 - 1;
                    }
                }
                const int64_t L[3] = {vs.n, vts.n, vns.n};
                int64_t *dst[3] = {&iv[k], &it[k], &in[k]};
                for (int q = 0; q < 3; q++) {
                    int64_t v = parse_int0(piece[q], pl[q]);
                    if (v < 0) v += L[q];
                    *dst[q] = v;
                }
            }
            for (int i = 1; i < na - 1; i++) {
                const int c[3] = {0, i, i + 1};
                for (int q = 0; q < 3; q++)
                    if (iv[c[q]] < 0 || iv[c[q]] >= vs.n || it[c[q]] < 0 || it[c[q]] >= vts.n || in[c[q]] < 0 ||
                        in[c[q]] >= vns.n)
                        rc = -1;
                if (rc) break;
                if (pos) {
                    Vec P[3], Nn[3];
                    for (int q = 0; q < 3; q++) {
                        const double *pv = vs.v + 3 * iv[c[q]], *nv = vns.v + 3 * in[c[q]], *tv = vts.v + 3 * it[c[q]];
                        P[q] = (Vec){pv[0], pv[1], pv[2]};
                        Nn[q] = (Vec){nv[0], nv[1], nv[2]};
                        uv[6 * nt + 2 * q] = tv[0];
                        uv[6 * nt + 2 * q + 1] = tv[1];
                    }
                    const Vec fn = tri_normal(P[0], P[1], P[2]); /* FixNormals (triangle.go:46-58) */
                    for (int q = 0; q < 3; q++) {
                        if (Nn[q].X == 0 && Nn[q].Y == 0 && Nn[q].Z == 0) Nn[q] = fn;
                        double *po = pos + 9 * nt + 3 * q, *no = nrm + 9 * nt + 3 * q;
                        po[0] = P[q].X, po[1] = P[q].Y, po[2] = P[q].Z;
                        no[0] = Nn[q].X, no[1] = Nn[q].Y, no[2] = Nn[q].Z;
                    }
                }
                nt++;
            }
        }
        free(f), free(fl), free(iv);
    }
    free(vs.v);
    free(vts.v);
    free(vns.v);
    if (rc) return rc;
    *ntri = nt;
    return 0;
}

/* ========================================================================
 * PNG row filters (Go image/png writer: filter, paeth, abs8)
 * ======================================================================== */

static int png_abs8(uint8_t d) { return d < 128 ? d : 256 - d; }
static uint8_t png_paeth(uint8_t a, uint8_t b, uint8_t c) {
    int pc = c, pa = (int)b - pc, pb = (int)a - pc;
    pc = abs(pa + pb);
    pa = abs(pa);
    pb = abs(pb);
    if (pa <= pb && pa <= pc) return a;
    if (pb <= pc) return b;
    return c;
}

int32_t orc_png_filter(const uint8_t *nrgba, int32_t w, int32_t h, uint8_t *out) {
    int opaque = 1;
    for (int64_t i = 0; i < (int64_t)w * h; i++) opaque &= nrgba[4 * i + 3] == 255;
    const int bpp = opaque ? 3 : 4;
    if (!out) return bpp;
    const int n = w * bpp;
    uint8_t *cr[5], *prev = calloc((size_t)n, 1), *cur = malloc((size_t)n);
    for (int f = 0; f < 5; f++) cr[f] = malloc((size_t)n);
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++)
            for (int k = 0; k < bpp; k++) cur[x * bpp + k] = nrgba[4 * ((int64_t)y * w + x) + k];
        /* filtered candidates: 0 None, 1 Sub, 2 Up, 3 Average, 4 Paeth */
        for (int i = 0; i < n; i++) {
            const uint8_t a = i >= bpp ? cur[i - bpp] : 0, b = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
            cr[0][i] = cur[i];
            cr[1][i] = (uint8_t)(cur[i] - a);
            cr[2][i] = (uint8_t)(cur[i] - b);
            cr[3][i] = (uint8_t)(cur[i] - (uint8_t)(((int)a + (int)b) / 2));
            cr[4][i] = (uint8_t)(cur[i] - png_paeth(a, b, c));
        }
        static const int order[5] = {2, 4, 0, 1, 3};
        int best = -1, pick = 2;
        for (int k = 0; k < 5; k++) {
            int sum = 0;
            for (int i = 0; i < n; i++) sum += png_abs8(cr[order[k]][i]);
            if (best < 0 || sum < best) best = sum, pick = order[k];
        }
        uint8_t *o = out + (int64_t)y * (n + 1);
        o[0] = (uint8_t)pick;
        memcpy(o + 1, cr[pick], (size_t)n);
        memcpy(prev, cur, (size_t)n);
    }
    for (int f = 0; f < 5; f++) free(cr[f]);
    free(prev);
    free(cur);
    return bpp;
}
