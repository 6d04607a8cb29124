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

}  // namespace fgl
