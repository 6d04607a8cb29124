// fgl_scan.h -- exclusive scan of per-block sums by a single workgroup, the
// middle step of the reduce / scan / apply passes that place variable-length
// outputs (pairs, spans, fragments) in submission order.
#pragma once

#include "fgl_device.h"

namespace fgl {

// Inclusive max-scan over the 64 lanes of a wave with DPP row shifts and
// row broadcasts (values >= 0; lanes with no source contribute 0).
FGL_D uint32_t wave_incl_max(uint32_t v) {
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false));  // row_shr:1
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false));  // row_shr:2
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false));  // row_shr:4
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false));  // row_shr:8
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false));  // row_bcast:15
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false));  // row_bcast:31
    return v;
}

constexpr int SCAN_THREADS = 1024;
constexpr int SCAN_ITEMS = 16;  // sums loaded per batch

// Writes off[k] = sum[0] + ... + sum[k-1] for k < nb and returns the total
// in every thread.  Called by all SCAN_THREADS threads of the workgroup;
// lds holds SCAN_THREADS / 64 entries.
FGL_D uint64_t scan_block_sums(const uint64_t *sum, uint64_t *off, int nb, uint64_t *lds) {
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int per = (nb + SCAN_THREADS - 1) / SCAN_THREADS;
    const int lo = tid * per;
    const int hi = lo + per < nb ? lo + per : nb;
    // each thread's run of sums, SCAN_ITEMS independent loads at a time
    uint64_t s = 0;
    for (int c = lo; c < hi; c += SCAN_ITEMS) {
        uint64_t v[SCAN_ITEMS];
        // clamped addresses keep the loads unconditional (issued together)
#pragma unroll
        for (int q = 0; q < SCAN_ITEMS; q++) v[q] = sum[c + q < hi ? c + q : hi - 1];
#pragma unroll
        for (int q = 0; q < SCAN_ITEMS; q++) s += c + q < hi ? v[q] : 0;
    }
    uint64_t incl = s;
    for (int o = 1; o < 64; o <<= 1) {
        uint64_t y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    if (lane == 63) lds[wave] = incl;
    __syncthreads();
    uint64_t base = 0, total = 0;
    for (int w = 0; w < SCAN_THREADS / 64; w++) {
        const uint64_t x = lds[w];
        if (w < wave) base += x;
        total += x;
    }
    uint64_t run = base + incl - s;
    for (int c = lo; c < hi; c += SCAN_ITEMS) {
        uint64_t v[SCAN_ITEMS];
#pragma unroll
        for (int q = 0; q < SCAN_ITEMS; q++) v[q] = sum[c + q < hi ? c + q : hi - 1];
#pragma unroll
        for (int q = 0; q < SCAN_ITEMS; q++) {
            if (c + q < hi) off[c + q] = run;
            run += v[q];
        }
    }
    return total;
}

}  // namespace fgl
