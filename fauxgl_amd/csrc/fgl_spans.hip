// fgl_spans.hip -- coverage spans.  After binning, every (tile, raster
// triangle) pair is expanded into one span per row the triangle's box
// reaches inside the tile: the reference's row walk (context.go:184-221) is
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

namespace fgl {

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

FGL_D uint32_t pair_spans(const DrawArgs &a, const uint32_t *keys, const uint32_t *vals, uint32_t p) {
    int32_t cx0, cx1, cy0, cy1;
    tile_rect(a, keys[p], cx0, cx1, cy0, cy1);
    const PrimRec *pr = prim_rec(a, vals[p]);
    ItemGeo g = item_geo(pr->x0, pr->y0, pr->x1, pr->y1, a.W, cx0, cx1, cy0, cy1);
    uint64_t n = (uint64_t)g.nY * (uint64_t)g.nk;
    return n > 0xffffffull ? 0xffffffu : (uint32_t)n;
}

// Spans per pair, summed per block of SPAN_THREADS pairs.
__global__ __launch_bounds__(SPAN_THREADS) void k_span_count(DrawArgs a) {
    __shared__ uint64_t wsum[SPAN_THREADS / 64];
    if (a.ctr->overflow) return;
    const uint32_t E = a.ctr->npairs;
    const uint32_t *keys = a.keys[a.sorted_buf];
    const uint32_t *vals = a.vals[a.sorted_buf];
    uint32_t p = blockIdx.x * SPAN_THREADS + threadIdx.x;
    uint32_t n = p < E ? pair_spans(a, keys, vals, p) : 0u;
    if (p < E) a.pspans[p] = n;
    uint64_t v = n;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t s = 0;
        for (int w = 0; w < SPAN_THREADS / 64; w++) s += wsum[w];
        a.span_blocksum[blockIdx.x] = s;
    }
}

// Exclusive scan of the per-block span sums (one workgroup), total and
// capacity check.
__global__ __launch_bounds__(SCAN_THREADS) void k_span_scan(DrawArgs a) {
    __shared__ uint64_t lds[SCAN_THREADS / 64];
    if (a.ctr->overflow) return;
    const uint32_t E = a.ctr->npairs;
    const int nb = (int)((E + SPAN_THREADS - 1) / SPAN_THREADS);
    const uint64_t total = scan_block_sums(a.span_blocksum, a.span_blockoff, nb, lds);
    // k_spans adds each span block's fragment count into frag_blocksum
    const uint64_t ns = total > a.span_cap ? a.span_cap : total;
    const uint32_t nfb = (uint32_t)((ns + SPAN_THREADS - 1) / SPAN_THREADS);
    for (uint32_t k = threadIdx.x; k < nfb; k += SCAN_THREADS) a.frag_blocksum[k] = 0;
    if (threadIdx.x == SCAN_THREADS - 1) {
        if (total > a.span_cap) atomicOr(&a.ctr->overflow, 16u);
        a.ctr->nspans = total > a.span_cap ? a.span_cap : (uint32_t)total;
        a.ctr->spans_lo = (uint32_t)total;
        a.ctr->spans_hi = (uint32_t)(total >> 32);
        a.poff[E] = total > a.span_cap ? a.span_cap : (uint32_t)total;
    }
}

// edge(a, b, c) (context.go:147-149)
FGL_D double edge(double ax, double ay, double bx, double by, double cx, double cy) {
    return (bx - cx) * (ay - cy) - (by - cy) * (ax - cx);
}

// Fragment counts per SPAN_THREADS-span block: collected in LDS for the
// FB_SLOTS blocks from the workgroup's first span on, in global memory
// beyond that.
constexpr uint32_t FB_SLOTS = 64;

FGL_D void add_frags(const DrawArgs &a, uint32_t *lfb, uint32_t blk0, uint32_t blk, uint64_t cnt) {
    if (!cnt) return;
    if (blk - blk0 < FB_SLOTS) atomicAdd(&lfb[blk - blk0], (uint32_t)cnt);
    else atomicAdd((unsigned long long *)&a.frag_blocksum[blk], (unsigned long long)cnt);
}

// One pair: setup once, then the rows in order (spans written at base..).
FGL_D void walk_pair(const DrawArgs &a, uint32_t p, uint32_t n, uint64_t base, const uint32_t *keys,
                     const uint32_t *vals, uint32_t *lfb, uint32_t blk0) {
    const uint32_t ref = vals[p];
    int32_t cx0, cx1, cy0, cy1;
    tile_rect(a, keys[p], cx0, cx1, cy0, cy1);
    const PrimRec *pr = prim_rec(a, ref);
    const int32_t x0 = pr->x0, y0 = pr->y0, x1 = pr->x1, y1 = pr->y1;
    ItemGeo g = item_geo(x0, y0, x1, y1, a.W, cx0, cx1, cy0, cy1);
    const double s0x = pr->sx[0], s0y = pr->sy[0], s1x = pr->sx[1], s1y = pr->sy[1], s2x = pr->sx[2],
                 s2y = pr->sy[2];
    // forward differencing variables and reciprocals (context.go:162-181)
    const double px = (double)x0 + 0.5, py = (double)y0 + 0.5;
    double w00 = edge(s1x, s1y, s2x, s2y, px, py);
    double w01 = edge(s2x, s2y, s0x, s0y, px, py);
    double w02 = edge(s0x, s0y, s1x, s1y, px, py);
    const double a01 = s1y - s0y, b01 = s0x - s1x;
    const double a12 = s2y - s1y, b12 = s1x - s2x;
    const double a20 = s0y - s2y, b20 = s2x - s0x;
    const double ra = pr->ra;  // 1 / edge(s0, s1, s2), computed with the record
    const double ra12 = 1.0 / a12, ra20 = 1.0 / a20, ra01 = 1.0 / a01;
    // rows y0 .. ya-1 of the row recurrence (context.go:275-277)
    for (int32_t y = y0; y < g.ya; y++) {
        w00 += b12;
        w01 += b20;
        w02 += b01;
    }
    Span *out = a.spans + base;
    uint32_t m = 0;
    // fragments per SPAN_THREADS-span block, added once per block touched
    uint32_t fblk = (uint32_t)(base / SPAN_THREADS);
    uint64_t facc = 0;
    for (int32_t yy = 0; yy < g.nY && m < n; yy++) {
        const int32_t y = g.ya + yy;
        // skip to the first pixel that can be inside (context.go:185-206)
        double d = 0;
        double d0 = -w00 * ra12;
        double d1 = -w01 * ra20;
        double d2 = -w02 * ra01;
        if (w00 < 0 && d0 > d) d = d0;
        if (w01 < 0 && d1 > d) d = d1;
        if (w02 < 0 && d2 > d) d = d2;
        d = (double)go_f2i(d);
        if (d < 0) d = 0;
        const double w0s = w00 + a12 * d;
        const double w1s = w01 + a20 * d;
        const double w2s = w02 + a01 * d;
        const int64_t id = go_f2i(d);
        const int32_t xs = (id > (int64_t)x1 - x0) ? x1 + 1 : x0 + (int32_t)id;
        for (int32_t kk = 0; kk < g.nk && m < n; kk++, m++) {
            const int32_t k = g.kmin + kk;
            const int32_t kw = k * a.W;
            const int32_t pyk = y + k;
            int32_t xa = x0 > kw + cx0 ? x0 : kw + cx0;
            int32_t xb = x1 < kw + cx1 ? x1 : kw + cx1;
            Span sp;
            sp.w[0] = 0;
            sp.w[1] = 0;
            sp.w[2] = 0;
            sp.ref = ref;
            sp.pix = 0;
            sp.n = 0;
            if (pyk >= cy0 && pyk <= cy1 && xa <= xb) {
                double w0 = w0s, w1 = w1s, w2 = w2s;
                bool was_inside = false;
                uint32_t cnt = 0;
                for (int32_t x = xs; x <= xb; x++) {
                    double b0 = w0 * ra, b1 = w1 * ra, b2 = w2 * ra;
                    if (b0 < 0 || b1 < 0 || b2 < 0) {
                        if (was_inside) break;
                    } else {
                        was_inside = true;
                        if (x >= xa) {
                            if (cnt == 0) {
                                sp.w[0] = w0;
                                sp.w[1] = w1;
                                sp.w[2] = w2;
                                sp.pix = (uint16_t)((pyk - cy0) * TILE_W + (x - kw - cx0));
                            }
                            cnt++;
                        }
                    }
                    w0 += a12;
                    w1 += a20;
                    w2 += a01;
                }
                sp.n = (uint16_t)(cnt | (k == 0 ? SPAN_INB : 0u));
                const uint32_t blk = (uint32_t)((base + m) / SPAN_THREADS);
                if (blk != fblk) {
                    add_frags(a, lfb, blk0, fblk, facc);
                    fblk = blk;
                    facc = 0;
                }
                facc += cnt;
            }
            out[m] = sp;
        }
        w00 += b12;
        w01 += b20;
        w02 += b01;
    }
    add_frags(a, lfb, blk0, fblk, facc);
}

// One thread per pair.  Also adds the covered pixels into the fragment
// count of each span block it writes.
__global__ __launch_bounds__(SPAN_THREADS) void k_spans(DrawArgs a) {
    __shared__ uint64_t wsum[SPAN_THREADS / 64];
    __shared__ uint32_t lfb[FB_SLOTS];
    if (a.ctr->overflow) return;
    const uint32_t E = a.ctr->npairs;
    const uint32_t *keys = a.keys[a.sorted_buf];
    const uint32_t *vals = a.vals[a.sorted_buf];
    const uint32_t p = blockIdx.x * SPAN_THREADS + threadIdx.x;
    const uint32_t n = p < E ? a.pspans[p] : 0u;
    if (threadIdx.x < FB_SLOTS) lfb[threadIdx.x] = 0;
    // block-exclusive scan of n
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t v = n;
    for (int off = 1; off < 64; off <<= 1) {
        uint64_t y = __shfl_up(v, off);
        if (lane >= off) v += y;
    }
    if (lane == 63) wsum[wave] = v;
    __syncthreads();
    const uint64_t bbase = a.span_blockoff[blockIdx.x];
    const uint32_t blk0 = (uint32_t)(bbase / SPAN_THREADS);
    uint64_t base = bbase;
    for (int w = 0; w < wave; w++) base += wsum[w];
    base += v - n;
    if (p < E) {
        a.poff[p] = (uint32_t)base;
        if (n) walk_pair(a, p, n, base, keys, vals, lfb, blk0);
    }
    __syncthreads();
    if (threadIdx.x < FB_SLOTS && lfb[threadIdx.x])
        atomicAdd((unsigned long long *)&a.frag_blocksum[blk0 + threadIdx.x], (unsigned long long)lfb[threadIdx.x]);
}

static unsigned span_blocks(const DrawArgs &a) {
    return (unsigned)(((uint64_t)a.pair_cap + SPAN_THREADS - 1) / SPAN_THREADS);
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

}  // namespace fgl
