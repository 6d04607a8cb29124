// fgl_tiles.h -- which tiles a raster triangle's pixels can land in.
//
// rasterize() (context.go:151-281) does not clamp its box to the screen and
// bounds-checks only the linear index i = y*W + x (context.go:223-228), so a
// pixel with x outside [0, W) lands on row y + floor(x / W) at column
// x mod W.  Only fat lines reach that case; this mapping keeps it exact.
#pragma once

#include "fgl_device.h"

namespace fgl {

struct TileRect {
    int32_t tx0, ty0, tx1, ty1;  // inclusive
};

FGL_D int32_t floor_div(int32_t a, int32_t b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

// Tiles that pixels of the box [x0,x1] x [y0,y1] can map to; false if none.
FGL_D bool prim_tile_rect(int32_t x0, int32_t y0, int32_t x1, int32_t y1, int32_t W, int32_t H, TileRect &r) {
    if (x0 > x1 || y0 > y1) return false;
    int32_t kmin = floor_div(x0, W), kmax = floor_div(x1, W);
    int32_t cx0, cx1, ry0, ry1;
    if (kmin == kmax) {
        cx0 = x0 - kmin * W;
        cx1 = x1 - kmin * W;
        ry0 = y0 + kmin;
        ry1 = y1 + kmin;
    } else {
        cx0 = 0;
        cx1 = W - 1;
        ry0 = y0 + kmin;
        ry1 = y1 + kmax;
    }
    if (ry0 < 0) ry0 = 0;
    if (ry1 > H - 1) ry1 = H - 1;
    if (ry0 > ry1) return false;
    r.tx0 = cx0 / TILE_W;
    r.tx1 = cx1 / TILE_W;
    r.ty0 = ry0 / TILE_H;
    r.ty1 = ry1 / TILE_H;
    return true;
}

// Work items of one raster triangle inside one tile: (row y, wrap k) pairs,
// y-major, over y in [ya, ya+nY) and k in [kmin, kmin+nk).  Pixel (x, y)
// with x in [k*W + cx0, k*W + cx1] lands on tile pixel (x - k*W, y + k).
// wb bounds the fragments one item can produce.
struct ItemGeo {
    int32_t ya, nY, kmin, nk, wb;
};

FGL_D ItemGeo item_geo(int32_t x0, int32_t y0, int32_t x1, int32_t y1, int32_t W, int32_t cx0, int32_t cx1,
                       int32_t cy0, int32_t cy1) {
    ItemGeo g;
    int32_t kmin = floor_div(x0, W), kmax = floor_div(x1, W);
    int32_t ya = y0 > cy0 - kmax ? y0 : cy0 - kmax;
    int32_t yb = y1 < cy1 - kmin ? y1 : cy1 - kmin;
    g.ya = ya;
    g.nY = yb >= ya ? yb - ya + 1 : 0;
    g.kmin = kmin;
    g.nk = kmax - kmin + 1;
    int32_t bw = x1 - x0 + 1, tw = cx1 - cx0 + 1;
    g.wb = bw < tw ? bw : tw;
    return g;
}

// A conservative form of the reference's inside test (context.go:184-228),
// used to drop work that provably yields no pixel.  For a triangle whose box
// is at most `lim` + 1 pixels square (lim <= 23), a pixel center p is surely
// outside when one edge function at p has the sign opposite to
// edge(s0, s1, s2) (the sign of ra) by more than tol = (M + 16) * 1e-8, M
// bounding the coordinates: every way rasterize() evaluates that edge
// function there (the corner value, the row and pixel steps, the skip) stays
// within 1e-10 + 1e-13 * M of the exact value for a box this small, as does
// the direct evaluation here, so the reference's b for that edge is negative
// at p.  ok is false for a degenerate or non-finite triangle, which is left
// to the walk.
struct EdgeMargin {
    bool ok;
    double sg, tol, area;
};

FGL_D EdgeMargin edge_margin(double s0x, double s0y, double s1x, double s1y, double s2x, double s2y, int64_t x0,
                             int64_t y0, int64_t x1, int64_t y1, int64_t lim) {
    EdgeMargin m;
    m.ok = false;
    m.sg = 1;
    m.tol = 0;
    m.area = 0;
    if (x1 - x0 > lim || y1 - y0 > lim) return m;
    const double area = (s1x - s2x) * (s0y - s2y) - (s1y - s2y) * (s0x - s2x);
    if (!(__builtin_fabs(area) > 0) || !__builtin_isfinite(area)) return m;
    double M = __builtin_fmax(__builtin_fabs((double)x0), __builtin_fabs((double)x1));
    M = __builtin_fmax(M, __builtin_fmax(__builtin_fabs((double)y0), __builtin_fabs((double)y1)));
    M = __builtin_fmax(M, __builtin_fmax(__builtin_fabs(s0x), __builtin_fabs(s0y)));
    M = __builtin_fmax(M, __builtin_fmax(__builtin_fabs(s1x), __builtin_fabs(s1y)));
    M = __builtin_fmax(M, __builtin_fmax(__builtin_fabs(s2x), __builtin_fabs(s2y)));
    m.ok = true;
    m.sg = area > 0 ? 1.0 : -1.0;
    m.tol = (M + 16.0) * 1e-8;
    m.area = __builtin_fabs(area);
    return m;
}

FGL_D bool surely_outside(const EdgeMargin &m, double s0x, double s0y, double s1x, double s1y, double s2x,
                          double s2y, int64_t x, int64_t y) {
    const double px = (double)x + 0.5, py = (double)y + 0.5;
    const double e0 = m.sg * ((s2x - px) * (s1y - py) - (s2y - py) * (s1x - px));  // edge(s1, s2, p)
    const double e1 = m.sg * ((s0x - px) * (s2y - py) - (s0y - py) * (s2x - px));  // edge(s2, s0, p)
    const double e2 = m.sg * ((s1x - px) * (s0y - py) - (s1y - py) * (s0x - px));  // edge(s0, s1, p)
    return e0 < -m.tol || e1 < -m.tol || e2 < -m.tol;
}

// Whether every pixel center of row y is surely outside because the row's
// center line y + 0.5 misses the triangle's vertical extent [lo, hi] by
// delta > 0.  With P such a center, P = sum b_k s_k and sum b_k = 1 give
// sum over negative b_k of |b_k| >= delta / (delta + h), h = hi - lo, so one
// edge function has magnitude >= area * delta / (2 (delta + h)) on the
// outside; asking 4 (delta + h) tol of it leaves the rounding of delta, h
// and the area far behind.
FGL_D bool row_surely_outside(const EdgeMargin &m, double lo, double hi, int64_t y) {
    const double py = (double)y + 0.5;
    const double delta = py > hi ? py - hi : lo - py;
    return delta > 0 && m.area * delta > 4.0 * (delta + (hi - lo)) * m.tol;
}

}  // namespace fgl
