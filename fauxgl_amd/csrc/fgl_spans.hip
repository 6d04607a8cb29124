// fgl_spans.hip -- coverage spans.  After binning, every (tile, raster
// triangle) pair is expanded into one span per row the triangle's box
// reaches inside the tile (for a small box, per row whose pixel centers are
// not all surely outside the triangle): the reference's row walk (context.go:184-221) is
// run for that row, with the row recurrence carried from row to row and the
// per-triangle setup (edge values, steps, reciprocals, context.go:162-181)
// done once per pair.  A span records the first covered pixel inside the
// tile, the edge values the reference held there, and how many consecutive
// pixels are covered.  The tile kernel replays the per-pixel additions from
// those edge values, so every barycentric it uses is bit-identical to the
// reference's serial loop.
#include "fgl_internal.h"
#include "fgl_scan.h"
#include "fgl_tiles.h"
#include "fgl_draws.h"

namespace fgl {
namespace FGL_TILE_NS {

FGL_D const PrimRec *prim_rec(const DrawArgs &a, uint32_t ref) {
    return (ref & REF_SIDE) ? a.side_rec + (ref & ~REF_SIDE) : a.rec_main + ref;
}

FGL_D void tile_rect(const DrawArgs &a, uint32_t t, int32_t &cx0, int32_t &cx1, int32_t &cy0, int32_t &cy1) {
    int32_t tx = (int32_t)(t % (uint32_t)a.ntx), ty = (int32_t)(t / (uint32_t)a.ntx);
    cx0 = tx * TILE_W;
    cy0 = ty * TILE_H;
    cx1 = min(cx0 + TILE_W, a.W) - 1;
    cy1 = min(cy0 + TILE_H, a.H) - 1;
}

// The rows of a pair that get spans.  A triangle whose box is at most 24
// pixels square and whose rows in the tile do not wrap (nk == 1, at most 15
// rows) skips every row whose pixel centers inside the tile are all surely
// outside it (edge_margin): a row whose center line misses the triangle's
// vertical extent, and, in a box at most 8 pixels wide, a row whose centers
// each fail the edge test.  Such a row's span would cover nothing.  Bit r
// stands for row ya + r; ROWS_ALL keeps all nY * nk rows.
constexpr uint16_t ROWS_ALL = 0x8000;

FGL_D uint16_t pair_rows(const PrimRec *pr, const ItemGeo &g, int32_t W, int32_t cx0, int32_t cx1) {
    if (g.nk != 1 || g.nY > 15) return ROWS_ALL;
    const double s0x = pr->sx[0], s0y = pr->sy[0], s1x = pr->sx[1], s1y = pr->sy[1], s2x = pr->sx[2],
                 s2y = pr->sy[2];
    const EdgeMargin m = edge_margin(s0x, s0y, s1x, s1y, s2x, s2y, pr->x0, pr->y0, pr->x1, pr->y1, 23);
    if (!m.ok) return ROWS_ALL;
    const double lo = __builtin_fmin(s0y, __builtin_fmin(s1y, s2y));
    const double hi = __builtin_fmax(s0y, __builtin_fmax(s1y, s2y));
    const int32_t kw = g.kmin * W;
    const int32_t xa = pr->x0 > kw + cx0 ? pr->x0 : kw + cx0;
    const int32_t xb = pr->x1 < kw + cx1 ? pr->x1 : kw + cx1;
    const bool narrow = xb - xa <= 7;
    // the three edge functions (signed by sg) at the center of (xa, ya), and
    // their steps per pixel in x and in y; stepping adds at most 40 roundings
    // of values below 700, far inside tol
    const double px = (double)xa + 0.5, py = (double)g.ya + 0.5;
    double e0 = m.sg * ((s2x - px) * (s1y - py) - (s2y - py) * (s1x - px));  // edge(s1, s2, p)
    double e1 = m.sg * ((s0x - px) * (s2y - py) - (s0y - py) * (s2x - px));  // edge(s2, s0, p)
    double e2 = m.sg * ((s1x - px) * (s0y - py) - (s1y - py) * (s0x - px));  // edge(s0, s1, p)
    const double dx0 = m.sg * (s2y - s1y), dx1 = m.sg * (s0y - s2y), dx2 = m.sg * (s1y - s0y);
    const double dy0 = m.sg * (s1x - s2x), dy1 = m.sg * (s2x - s0x), dy2 = m.sg * (s0x - s1x);
    uint16_t rows = 0;
    for (int32_t r = 0; r < g.nY; r++, e0 += dy0, e1 += dy1, e2 += dy2) {
        if (row_surely_outside(m, lo, hi, g.ya + r)) continue;
        bool any = !narrow;
        double f0 = e0, f1 = e1, f2 = e2;
        for (int32_t x = xa; x <= xb && !any; x++, f0 += dx0, f1 += dx1, f2 += dx2)
            any = !(f0 < -m.tol || f1 < -m.tol || f2 < -m.tol);
        if (any) rows |= (uint16_t)(1u << r);
    }
    return rows;
}

// Spans of pair p: the rows pair_rows keeps (times the wraps).
FGL_D uint32_t pair_spans(const DrawArgs &a, const uint32_t *keys, const uint32_t *vals, uint32_t p, uint16_t &rows) {
    int32_t cx0, cx1, cy0, cy1;
    tile_rect(a, keys[p], cx0, cx1, cy0, cy1);
    const PrimRec *pr = prim_rec(a, vals[p]);
    ItemGeo g = item_geo(pr->x0, pr->y0, pr->x1, pr->y1, a.W, cx0, cx1, cy0, cy1);
    rows = pair_rows(pr, g, a.W, cx0, cx1);
    if (rows != ROWS_ALL) return (uint32_t)__builtin_popcount(rows);
    uint64_t n = (uint64_t)g.nY * (uint64_t)g.nk;
    return n > 0xffffffull ? 0xffffffu : (uint32_t)n;
}

// Spans per pair, summed per block of SPAN_THREADS pairs; also each tile's
// range in the sorted pairs, and the zeroed fragment counts per span block
// (the whole capacity, grid-wide) that k_spans accumulates into.
template <class G>
FGL_D void span_count_body(const DrawArgs &a, const G &g) {
    __shared__ uint64_t wsum[SPAN_THREADS / 64];
    if (a.ctr->overflow) return;
    const uint64_t nfb = ((uint64_t)a.span_cap + FRAG_BLOCK - 1) / FRAG_BLOCK;
    for (uint64_t k = (uint64_t)g.bid() * SPAN_THREADS + threadIdx.x; k < nfb; k += (uint64_t)g.nblk() * SPAN_THREADS)
        a.frag_blocksum[k] = 0;
    const uint32_t E = a.ctr->npairs;
    const uint32_t *keys = a.keys[a.sorted_buf];
    const uint32_t *vals = a.vals[a.sorted_buf];
    uint32_t p = g.bid() * SPAN_THREADS + threadIdx.x;
    uint16_t rows = ROWS_ALL;
    uint32_t n = p < E ? pair_spans(a, keys, vals, p, rows) : 0u;
    if (p < E) {
        a.pspans[p] = n;
        a.prows[p] = rows;
        // the tile's [start, end) in the sorted pairs
        const uint32_t t = keys[p];
        if (p == 0 || keys[p - 1] != t) a.tstart[t] = p;
        if (p + 1 == E || keys[p + 1] != t) a.tend[t] = p + 1;
    }
    uint64_t v = n;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t s = 0;
        for (int w = 0; w < SPAN_THREADS / 64; w++) s += wsum[w];
        a.span_blocksum[g.bid()] = s;
    }
}
FGL_DRAW_KERNELS(k_span_count, span_count_body, __launch_bounds__(SPAN_THREADS))

// Exclusive scan of the per-block span sums (one workgroup), total and
// capacity check.
template <class G>
FGL_D void span_scan_body(const DrawArgs &a, const G &) {
    __shared__ uint64_t lds[SCAN_THREADS / 64];
    if (a.ctr->overflow) return;
    const uint32_t E = a.ctr->npairs;
    const int nb = (int)((E + SPAN_THREADS - 1) / SPAN_THREADS);
    const uint64_t total = scan_block_sums(a.span_blocksum, a.span_blockoff, nb, lds);
    if (threadIdx.x == SCAN_THREADS - 1) {
        if (total > a.span_cap) atomicOr(&a.ctr->overflow, 16u);
        a.ctr->nspans = total > a.span_cap ? a.span_cap : (uint32_t)total;
        a.ctr->spans_lo = (uint32_t)total;
        a.ctr->spans_hi = (uint32_t)(total >> 32);
        a.poff[E] = total > a.span_cap ? a.span_cap : (uint32_t)total;
    }
}
FGL_DRAW_KERNELS(k_span_scan, span_scan_body, __launch_bounds__(SCAN_THREADS))

// edge(a, b, c) (context.go:147-149)
FGL_D double edge(double ax, double ay, double bx, double by, double cx, double cy) {
    return (bx - cx) * (ay - cy) - (by - cy) * (ax - cx);
}

// Fragment counts per FRAG_BLOCK-span block: collected in LDS for the
// FB_SLOTS blocks from the workgroup's first span on, in global memory
// beyond that.
constexpr uint32_t FB_SLOTS = 128;

FGL_D void add_frags(const DrawArgs &a, uint32_t *lfb, uint32_t blk0, uint32_t blk, uint64_t cnt) {
    if (!cnt) return;
    if (blk - blk0 < FB_SLOTS) atomicAdd(&lfb[blk - blk0], (uint32_t)cnt);
    else atomicAdd((unsigned long long *)&a.frag_blocksum[blk], (unsigned long long)cnt);
}

// Per-pair setup a wave shares through LDS: the edge values at the first
// row the tile needs (the row recurrence of context.go:275-277 run up to
// it), the steps, the reciprocals of the skip (context.go:185-206) and the
// box.
struct PairSetup {
    double w0, w1, w2;        // edge values at (x0 + 0.5, ya + 0.5)
    double a12, a20, a01;     // per-pixel steps
    double b12, b20, b01;     // per-row steps
    double ra12, ra20, ra01;  // 1 / a12, 1 / a20, 1 / a01
    double ra;                // 1 / edge(s0, s1, s2)
    int32_t x0, x1, ya, kmin, nk;
    int32_t cx0, cx1, cy0, cy1;
    uint32_t ref;
    uint32_t rows;  // pair_rows of the pair
};

FGL_D void pair_setup(const DrawArgs &a, const uint32_t *keys, const uint32_t *vals, uint32_t p, PairSetup &q) {
    const uint32_t ref = vals[p];
    int32_t cx0, cx1, cy0, cy1;
    tile_rect(a, keys[p], cx0, cx1, cy0, cy1);
    const PrimRec *pr = prim_rec(a, ref);
    const int32_t x0 = pr->x0, y0 = pr->y0, x1 = pr->x1, y1 = pr->y1;
    const ItemGeo g = item_geo(x0, y0, x1, y1, a.W, cx0, cx1, cy0, cy1);
    const double s0x = pr->sx[0], s0y = pr->sy[0], s1x = pr->sx[1], s1y = pr->sy[1], s2x = pr->sx[2],
                 s2y = pr->sy[2];
    // forward differencing variables and reciprocals (context.go:162-181)
    const double px = (double)x0 + 0.5, py = (double)y0 + 0.5;
    double w00 = edge(s1x, s1y, s2x, s2y, px, py);
    double w01 = edge(s2x, s2y, s0x, s0y, px, py);
    double w02 = edge(s0x, s0y, s1x, s1y, px, py);
    q.a01 = s1y - s0y;
    q.b01 = s0x - s1x;
    q.a12 = s2y - s1y;
    q.b12 = s1x - s2x;
    q.a20 = s0y - s2y;
    q.b20 = s2x - s0x;
    q.ra = pr->ra;  // 1 / edge(s0, s1, s2), computed with the record
    q.ra12 = 1.0 / q.a12;
    q.ra20 = 1.0 / q.a20;
    q.ra01 = 1.0 / q.a01;
    // rows y0 .. ya-1 of the row recurrence (context.go:275-277)
    for (int32_t y = y0; y < g.ya; y++) {
        w00 += q.b12;
        w01 += q.b20;
        w02 += q.b01;
    }
    q.w0 = w00;
    q.w1 = w01;
    q.w2 = w02;
    q.x0 = x0;
    q.x1 = x1;
    q.ya = g.ya;
    q.kmin = g.kmin;
    q.nk = g.nk;
    q.cx0 = cx0;
    q.cx1 = cx1;
    q.cy0 = cy0;
    q.cy1 = cy1;
    q.ref = ref;
    q.rows = a.prows[p];
}

// Span m of a pair (row yy = m / nk, wrapped column kk = m % nk, or the m-th
// row pair_rows kept): the row's
// edge values (yy row steps from ya), the skip to the first pixel that can
// be inside, then the pixel walk of the reference's row loop
// (context.go:185-221) through the tile's columns.  Returns the span and
// its covered-pixel count.
FGL_D Span pair_span(const PairSetup &q, uint32_t m, int32_t W, uint32_t &cnt_out) {
    int32_t yy = (int32_t)(m / (uint32_t)q.nk), kk = (int32_t)(m % (uint32_t)q.nk);
    if (q.rows != ROWS_ALL) {  // the m-th kept row (nk == 1)
        uint32_t r = q.rows;
        for (uint32_t k = 0; k < m; k++) r &= r - 1;
        yy = __builtin_ctz(r);
        kk = 0;
    }
    double w00 = q.w0, w01 = q.w1, w02 = q.w2;
    for (int32_t r = 0; r < yy; r++) {
        w00 += q.b12;
        w01 += q.b20;
        w02 += q.b01;
    }
    const int32_t y = q.ya + yy;
    // skip to the first pixel that can be inside (context.go:185-206)
    double d = 0;
    const double d0 = -w00 * q.ra12;
    const double d1 = -w01 * q.ra20;
    const double d2 = -w02 * q.ra01;
    if (w00 < 0 && d0 > d) d = d0;
    if (w01 < 0 && d1 > d) d = d1;
    if (w02 < 0 && d2 > d) d = d2;
    d = (double)go_f2i(d);
    if (d < 0) d = 0;
    double w0 = w00 + q.a12 * d;
    double w1 = w01 + q.a20 * d;
    double w2 = w02 + q.a01 * d;
    const int64_t id = go_f2i(d);
    const int32_t xs = (id > (int64_t)q.x1 - q.x0) ? q.x1 + 1 : q.x0 + (int32_t)id;
    const int32_t k = q.kmin + kk;
    const int32_t kw = k * W;
    const int32_t pyk = y + k;
    const int32_t xa = q.x0 > kw + q.cx0 ? q.x0 : kw + q.cx0;
    const int32_t xb = q.x1 < kw + q.cx1 ? q.x1 : kw + q.cx1;
    Span sp;
    sp.w[0] = 0;
    sp.w[1] = 0;
    sp.w[2] = 0;
    sp.ref = q.ref;
    sp.pix = 0;
    sp.n = 0;
    uint32_t cnt = 0;
    if (pyk >= q.cy0 && pyk <= q.cy1 && xa <= xb) {
        bool was_inside = false;
        for (int32_t x = xs; x <= xb; x++) {
            const double b0 = w0 * q.ra, b1 = w1 * q.ra, b2 = w2 * q.ra;
            if (b0 < 0 || b1 < 0 || b2 < 0) {
                if (was_inside) break;
            } else {
                was_inside = true;
                if (x >= xa) {
                    if (cnt == 0) {
                        sp.w[0] = w0;
                        sp.w[1] = w1;
                        sp.w[2] = w2;
                        sp.pix = (uint16_t)((pyk - q.cy0) * TILE_W + (x - kw - q.cx0));
                    }
                    cnt++;
                }
            }
            w0 += q.a12;
            w1 += q.a20;
            w2 += q.a01;
        }
        sp.n = (uint16_t)(cnt | (k == 0 ? SPAN_INB : 0u));
    }
    cnt_out = cnt;
    return sp;
}

// One thread per pair sets up; then each wave writes its pairs' spans 64
// at a time, one span per lane, so lanes share the walks evenly however
// many rows each pair has.  A lane finds its span's pair as the shading
// kernel finds a fragment's span (each pair marks its first span's window
// position, a max-scan carries the mark) and reads the pair's setup from
// LDS.  Also adds the covered pixels into the fragment count of each span
// block it writes.
template <class G>
FGL_D void spans_body(const DrawArgs &a, const G &g) {
    __shared__ uint64_t wsum[SPAN_THREADS / 64];
    __shared__ uint32_t lfb[FB_SLOTS];
    __shared__ uint32_t ostart[SPAN_THREADS / 64][64];
    __shared__ PairSetup setup[SPAN_THREADS / 64][64];
    if (a.ctr->overflow) return;
    const uint32_t E = a.ctr->npairs;
    const uint32_t *keys = a.keys[a.sorted_buf];
    const uint32_t *vals = a.vals[a.sorted_buf];
    const uint32_t p = g.bid() * SPAN_THREADS + threadIdx.x;
    const uint32_t n = p < E ? a.pspans[p] : 0u;
    for (uint32_t k = threadIdx.x; k < FB_SLOTS; k += SPAN_THREADS) lfb[k] = 0;
    // block-exclusive scan of n
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t v = n;
    for (int off = 1; off < 64; off <<= 1) {
        uint64_t y = __shfl_up(v, off);
        if (lane >= off) v += y;
    }
    if (lane == 63) wsum[wave] = v;
    __syncthreads();
    const uint64_t bbase = a.span_blockoff[g.bid()];
    const uint32_t blk0 = (uint32_t)(bbase / FRAG_BLOCK);
    uint64_t wbase = bbase;
    for (int w = 0; w < wave; w++) wbase += wsum[w];
    // this lane's pair's spans: [wbase + excl, wbase + excl + n)
    const uint32_t excl = (uint32_t)(v - n);
    const uint32_t T = (uint32_t)__shfl(v, 63);  // the wave's spans
    if (p < E) a.poff[p] = (uint32_t)(wbase + excl);
    if (n) pair_setup(a, keys, vals, p, setup[wave][lane]);
    __builtin_amdgcn_wave_barrier();
    Span *out = a.spans + wbase;
    uint32_t carry = 0;  // pair lane + 1 of the last span of the previous window
    for (uint32_t base = 0; base < T; base += 64) {
        const uint32_t q = base + (uint32_t)lane;
        ostart[wave][lane] = 0;
        if (n && excl >= base && excl - base < 64) ostart[wave][excl - base] = (uint32_t)lane + 1;
        __builtin_amdgcn_wave_barrier();
        uint32_t om = wave_incl_max(ostart[wave][lane]);
        om = om ? om : carry;
        carry = __builtin_amdgcn_readlane(om, 63);
        __builtin_amdgcn_wave_barrier();
        const int ow = (int)om - 1 < 0 ? 0 : (int)om - 1;
        const uint32_t oexcl = __shfl(excl, ow);
        if (q < T) {
            uint32_t cnt;
            const Span sp = pair_span(setup[wave][ow], q - oexcl, a.W, cnt);
            out[q] = sp;
            if (cnt) {
                const uint32_t blk = (uint32_t)((wbase + q) / FRAG_BLOCK);
                add_frags(a, lfb, blk0, blk, cnt);
            }
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < FB_SLOTS; k += SPAN_THREADS)
        if (lfb[k]) atomicAdd((unsigned long long *)&a.frag_blocksum[blk0 + k], (unsigned long long)lfb[k]);
}
FGL_DRAW_KERNELS(k_spans, spans_body, __launch_bounds__(SPAN_THREADS))

uint32_t span_blocks(const DrawArgs &a) {
    return (uint32_t)(((uint64_t)a.pair_cap + SPAN_THREADS - 1) / SPAN_THREADS);
}

hipError_t launch_span_count(const DrawArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_span_count, dim3(span_blocks(a)), dim3(SPAN_THREADS), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_span_scan(const DrawArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_span_scan, dim3(1), dim3(SCAN_THREADS), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_spans(const DrawArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_spans, dim3(span_blocks(a)), dim3(SPAN_THREADS), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_span_count_multi(const MultiArgs &m, uint32_t nb, const DrawArgs &, hipStream_t s) {
    hipLaunchKernelGGL(k_span_count_m, dim3(nb), dim3(SPAN_THREADS), 0, s, m);
    return hipGetLastError();
}
hipError_t launch_span_scan_multi(const MultiArgs &m, uint32_t nb, const DrawArgs &, hipStream_t s) {
    hipLaunchKernelGGL(k_span_scan_m, dim3(nb), dim3(SCAN_THREADS), 0, s, m);
    return hipGetLastError();
}
hipError_t launch_spans_multi(const MultiArgs &m, uint32_t nb, const DrawArgs &, hipStream_t s) {
    hipLaunchKernelGGL(k_spans_m, dim3(nb), dim3(SPAN_THREADS), 0, s, m);
    return hipGetLastError();
}

}  // namespace FGL_TILE_NS
}  // namespace fgl
