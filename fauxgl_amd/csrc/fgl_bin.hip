// fgl_bin.hip -- sort-middle binning.  Every raster triangle becomes one
// (tile, triangle) pair per tile its box can reach, written in submission
// order (input index, then fan / line-quad order).  A stable LSD radix sort
// on the tile id then leaves each tile's triangles contiguous and still in
// submission order, which is the order the reference draws them in
// (context.go:413-433 with one worker).
#include "fgl_internal.h"
#include "fgl_scan.h"
#include "fgl_tiles.h"
#include "fgl_draws.h"

namespace fgl {
namespace FGL_TILE_NS {

// ---------------------------------------------------------------------------
// per-draw reset

template <class G>
FGL_D void zero_body(const DrawArgs &a, const G &g) {
    int64_t i = (int64_t)g.bid() * blockDim.x + threadIdx.x;
    if (i == 0) {
        DrawCounters z = {};
        *a.ctr = z;
    }
    for (int64_t t = i; t < a.ntiles; t += (int64_t)g.nblk() * blockDim.x) {
        a.tstart[t] = 0;
        a.tend[t] = 0;
    }
}
FGL_DRAW_KERNELS(k_zero, zero_body, __launch_bounds__(256))

uint32_t blocks_zero(const DrawArgs &a) {
    int blocks = (a.ntiles + 255) / 256;
    if (blocks < 1) blocks = 1;
    if (blocks > 1024) blocks = 1024;
    return (uint32_t)blocks;
}
hipError_t launch_zero(const DrawArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_zero, dim3(blocks_zero(a)), dim3(256), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_zero_multi(const MultiArgs &m, uint32_t nb, const DrawArgs &, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_m, dim3(nb), dim3(256), 0, s, m);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// exclusive scan of the vertex-stage block sums (one workgroup)

template <class G>
FGL_D void scan_top_body(const DrawArgs &a, const G &) {
    __shared__ uint64_t lds[SCAN_THREADS / 64];
    const int tid = threadIdx.x;
    const uint64_t total = scan_block_sums(a.blocksum, a.blockoff, a.nvsblocks, lds);
    if (tid == SCAN_THREADS - 1) {
        uint32_t ovf = a.ctr->overflow;
        if (total > a.pair_cap) ovf |= 4u;
        a.ctr->npairs = total > a.pair_cap ? a.pair_cap : (uint32_t)total;
        a.ctr->pairs_lo = (uint32_t)total;
        a.ctr->pairs_hi = (uint32_t)(total >> 32);
        a.ctr->overflow = ovf;
    }
}
FGL_DRAW_KERNELS(k_scan_top, scan_top_body, __launch_bounds__(SCAN_THREADS))

hipError_t launch_scan_top(const DrawArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_THREADS), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_scan_top_multi(const MultiArgs &m, uint32_t nb, const DrawArgs &, hipStream_t s) {
    hipLaunchKernelGGL(k_scan_top_m, dim3(nb), dim3(SCAN_THREADS), 0, s, m);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// pair emission

template <class G>
FGL_D void emit_body(const DrawArgs &a, const G &g) {
    __shared__ uint64_t wsum[VS_THREADS / 64];
    if (a.ctr->overflow) return;
    int64_t i = (int64_t)g.bid() * VS_THREADS + threadIdx.x;
    uint32_t nt = i < a.in.n ? a.ntl[i] : 0u;
    // block-exclusive scan of nt
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t v = nt;
    for (int off = 1; off < 64; off <<= 1) {
        uint64_t y = __shfl_up(v, off);
        if (lane >= off) v += y;
    }
    if (lane == 63) wsum[wave] = v;
    __syncthreads();
    uint64_t wbase = 0;
    for (int w = 0; w < wave; w++) wbase += wsum[w];
    uint64_t pos = a.blockoff[g.bid()] + wbase + v - nt;
    if (i >= a.in.n || nt == 0) return;
    uint32_t nr = a.nrast[i];
    bool side = (nr & REF_SIDE) != 0;
    uint32_t cnt = nr & ~REF_SIDE;
    uint32_t *keys = a.keys[0];
    uint32_t *vals = a.vals[0];
    if (!side) {
        // the main record (a triangle drawn without clipping): its box from
        // the compact copy the vertex stage wrote, not from the record
        const int4 bx = a.tbox[i];
        TileRect r;
        if (!prim_tile_rect(bx.x, bx.y, bx.z, bx.w, a.W, a.H, r)) return;
        for (int32_t ty = r.ty0; ty <= r.ty1; ty++)
            for (int32_t tx = r.tx0; tx <= r.tx1; tx++) {
                keys[pos] = (uint32_t)(ty * a.ntx + tx);
                vals[pos] = (uint32_t)i;
                pos++;
            }
        return;
    }
    for (uint32_t j = 0; j < cnt; j++) {
        const PrimRec *rec;
        uint32_t ref;
        if (side) {
            ref = REF_SIDE | (a.sstart[i] + j);
            rec = a.side_rec + (a.sstart[i] + j);
        } else {
            ref = (uint32_t)i;
            rec = a.rec_main + i;
        }
        TileRect r;
        if (!prim_tile_rect(rec->x0, rec->y0, rec->x1, rec->y1, a.W, a.H, r)) continue;
        for (int32_t ty = r.ty0; ty <= r.ty1; ty++)
            for (int32_t tx = r.tx0; tx <= r.tx1; tx++) {
                keys[pos] = (uint32_t)(ty * a.ntx + tx);
                vals[pos] = ref;
                pos++;
            }
    }
}

FGL_DRAW_KERNELS(k_emit, emit_body, __launch_bounds__(VS_THREADS))

hipError_t launch_emit(const DrawArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_emit, dim3(a.nvsblocks), dim3(VS_THREADS), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_emit_multi(const MultiArgs &m, uint32_t nb, const DrawArgs &, hipStream_t s) {
    hipLaunchKernelGGL(k_emit_m, dim3(nb), dim3(VS_THREADS), 0, s, m);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// stable LSD radix sort of (tile id, ref) pairs, 8 bits per pass

template <class G>
FGL_D void radix_hist_body(const DrawArgs &a, const G &g, int pass) {
    __shared__ uint32_t h[256];
    if (a.ctr->overflow) return;
    const uint32_t E = a.ctr->npairs;
    const uint32_t *keys = a.keys[pass & 1];
    const int shift = 8 * pass;
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)g.bid() * RS_BLOCK;
    if (base >= E) {
        __syncthreads();
        a.hist[(uint64_t)threadIdx.x * a.nrsblocks + g.bid()] = 0;
        return;
    }
    // unconditional loads (clamped to the last pair) issue back to back
    uint32_t k[RS_ITEMS];
#pragma unroll
    for (int it = 0; it < RS_ITEMS; it++) {
        const uint64_t idx = base + (uint64_t)it * RS_THREADS + threadIdx.x;
        k[it] = keys[idx < E ? idx : E - 1];
    }
#pragma unroll
    for (int it = 0; it < RS_ITEMS; it++) {
        const uint64_t idx = base + (uint64_t)it * RS_THREADS + threadIdx.x;
        if (idx < E) atomicAdd(&h[(k[it] >> shift) & 255u], 1u);
    }
    __syncthreads();
    a.hist[(uint64_t)threadIdx.x * a.nrsblocks + g.bid()] = h[threadIdx.x];
}
FGL_DRAW_KERNELS_PASS(k_radix_hist, radix_hist_body, __launch_bounds__(RS_THREADS))

// Block-wide exclusive scan of one value per thread (256 threads); returns
// the exclusive prefix and writes the block total to *total.
__device__ uint32_t block_exclusive_scan_256(uint32_t v, uint32_t *wsum, uint32_t *total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    uint32_t base = 0, tot = 0;
    for (int w = 0; w < 4; w++) {
        if (w < wave) base += wsum[w];
        tot += wsum[w];
    }
    __syncthreads();
    *total = tot;
    return base + x - v;
}

// One workgroup per digit: exclusive scan of that digit's per-block counts
// (hist is digit-major) and the digit total.
template <class G>
FGL_D void radix_scan_body(const DrawArgs &a, const G &g) {
    __shared__ uint32_t wsum[4];
    if (a.ctr->overflow) return;
    const int d = (int)g.bid();
    uint32_t *row = a.hist + (uint64_t)d * a.nrsblocks;
    uint32_t carry = 0;
    for (int base = 0; base < a.nrsblocks; base += RS_THREADS) {
        int k = base + threadIdx.x;
        uint32_t v = k < a.nrsblocks ? row[k] : 0u;
        uint32_t tot;
        uint32_t ex = block_exclusive_scan_256(v, wsum, &tot);
        if (k < a.nrsblocks) row[k] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) a.dtot[d] = carry;
}
FGL_DRAW_KERNELS(k_radix_scan, radix_scan_body, __launch_bounds__(RS_THREADS))

template <class G>
FGL_D void radix_scatter_body(const DrawArgs &a, const G &g, int pass) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[RS_THREADS / 64][256];
    if (a.ctr->overflow) return;
    const uint32_t E = a.ctr->npairs;
    const uint32_t *kin = a.keys[pass & 1];
    const uint32_t *vin = a.vals[pass & 1];
    uint32_t *kout = a.keys[(pass + 1) & 1];
    uint32_t *vout = a.vals[(pass + 1) & 1];
    const int shift = 8 * pass;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t ltmask = (1ull << lane) - 1ull;
    {
        // digit d starts after all pairs with a smaller digit
        __shared__ uint32_t wsum[4];
        uint32_t tot;
        uint32_t dbase = block_exclusive_scan_256(a.dtot[tid], wsum, &tot);
        run[tid] = dbase + a.hist[(uint64_t)tid * a.nrsblocks + g.bid()];
    }
    const uint64_t base = (uint64_t)g.bid() * RS_BLOCK;
    if (base >= E) return;
    // all of this thread's pairs, loaded up front (clamped, unconditional)
    uint32_t kk[RS_ITEMS], vv[RS_ITEMS];
#pragma unroll
    for (int it = 0; it < RS_ITEMS; it++) {
        const uint64_t idx = base + (uint64_t)it * RS_THREADS + tid;
        const uint64_t ci = idx < E ? idx : E - 1;
        kk[it] = kin[ci];
        vv[it] = vin[ci];
    }
#pragma unroll
    for (int it = 0; it < RS_ITEMS; it++) {
        for (int w = 0; w < RS_THREADS / 64; w++) wcnt[w][tid] = 0;
        __syncthreads();
        uint64_t idx = base + (uint64_t)it * RS_THREADS + tid;
        bool valid = idx < E;
        uint32_t key = valid ? kk[it] : 0u;
        uint32_t val = vv[it];
        uint32_t digit = (key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            bool bit = (digit >> b) & 1u;
            uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        uint32_t rank = (uint32_t)__popcll(peers & ltmask);
        if (valid && rank == 0) wcnt[wave][digit] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            uint32_t r = run[tid];
            for (int w = 0; w < RS_THREADS / 64; w++) {
                uint32_t c = wcnt[w][tid];
                wcnt[w][tid] = r;
                r += c;
            }
            run[tid] = r;
        }
        __syncthreads();
        if (valid) {
            uint32_t p = wcnt[wave][digit] + rank;
            kout[p] = key;
            vout[p] = val;
        }
        __syncthreads();
    }
}

FGL_DRAW_KERNELS_PASS(k_radix_scatter, radix_scatter_body, __launch_bounds__(RS_THREADS))

hipError_t launch_radix_hist(const DrawArgs &a, int pass, hipStream_t s) {
    hipLaunchKernelGGL(k_radix_hist, dim3(a.nrsblocks), dim3(RS_THREADS), 0, s, a, pass);
    return hipGetLastError();
}
hipError_t launch_radix_scan(const DrawArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_radix_scan, dim3(256), dim3(RS_THREADS), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_radix_scatter(const DrawArgs &a, int pass, hipStream_t s) {
    hipLaunchKernelGGL(k_radix_scatter, dim3(a.nrsblocks), dim3(RS_THREADS), 0, s, a, pass);
    return hipGetLastError();
}
hipError_t launch_radix_hist_multi(const MultiArgs &m, uint32_t nb, const DrawArgs &, hipStream_t s) {
    hipLaunchKernelGGL(k_radix_hist_m, dim3(nb), dim3(RS_THREADS), 0, s, m);
    return hipGetLastError();
}
hipError_t launch_radix_scan_multi(const MultiArgs &m, uint32_t nb, const DrawArgs &, hipStream_t s) {
    hipLaunchKernelGGL(k_radix_scan_m, dim3(nb), dim3(RS_THREADS), 0, s, m);
    return hipGetLastError();
}
hipError_t launch_radix_scatter_multi(const MultiArgs &m, uint32_t nb, const DrawArgs &, hipStream_t s) {
    hipLaunchKernelGGL(k_radix_scatter_m, dim3(nb), dim3(RS_THREADS), 0, s, m);
    return hipGetLastError();
}

}  // namespace FGL_TILE_NS
}  // namespace fgl
