// fgl_misc.hip -- framebuffer fills (deferred clears made concrete), the
// SSAA resolve, and DepthImage.
#include <mutex>

#include "fgl_internal.h"

namespace fgl {

// ClearColorBufferWith / ClearDepthBufferWith (context.go:119-141)
__global__ void k_fill_color(uint32_t *color, int64_t n, uint32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        color[i] = v;
}
__global__ void k_fill_depth(double *depth, int64_t n, double v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        depth[i] = v;
}

static unsigned grid_for(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    return (unsigned)b;
}

// word copy in a kernel (graph-capturable, unlike a memcpy node between
// arbitrary device buffers on every runtime): the batch gathers each
// view's draw counters with it
__global__ void k_copy_words(const uint32_t *src, uint32_t *dst, int32_t n) {
    for (int32_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}
hipError_t launch_copy_words(const uint32_t *src, uint32_t *dst, int32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_copy_words, dim3(1), dim3(64), 0, s, src, dst, n);
    return hipGetLastError();
}

// one workgroup per record: record k's n words from src[k] to dst + slot[k] * n
__global__ void k_gather_words(const uint32_t *const *src, const uint32_t *slot, uint32_t *dst, int32_t n) {
    const uint32_t *s = src[blockIdx.x];
    uint32_t *d = dst + (uint64_t)slot[blockIdx.x] * (uint64_t)n;
    for (int32_t i = threadIdx.x; i < n; i += blockDim.x) d[i] = s[i];
}
hipError_t launch_gather_words(const uint32_t *const *src, const uint32_t *slot, uint32_t *dst, int32_t n,
                               int32_t count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_words, dim3((unsigned)count), dim3(64), 0, s, src, slot, dst, n);
    return hipGetLastError();
}

// one workgroup per block: its marked channels get the clear values
__global__ __launch_bounds__(256) void k_cblk_fill(uint32_t *color, double *depth, int32_t W, int32_t H,
                                                  uint8_t *cblk, int32_t cbw, uint32_t cc, double cd, uint8_t chans) {
    const int b = blockIdx.x;
    const uint8_t m = cblk[b] & chans;
    if (!m) return;
    const int x0 = (b % cbw) * CB_W, y0 = (b / cbw) * CB_H;
    for (int p = threadIdx.x; p < CB_W * CB_H; p += 256) {
        const int x = x0 + p % CB_W, y = y0 + p / CB_W;
        if (x >= W || y >= H) continue;
        const int64_t i = (int64_t)y * W + x;
        if (m & CB_COLOR) color[i] = cc;
        if (m & CB_DEPTH) depth[i] = cd;
    }
    __syncthreads();
    if (threadIdx.x == 0) cblk[b] &= (uint8_t)~m;
}
hipError_t launch_cblk_fill(uint32_t *color, double *depth, int32_t W, int32_t H, uint8_t *cblk, int32_t cbw,
                            int32_t cbh, uint32_t clear_color, double clear_depth, uint8_t chans, hipStream_t s) {
    hipLaunchKernelGGL(k_cblk_fill, dim3((unsigned)(cbw * cbh)), dim3(256), 0, s, color, depth, W, H, cblk, cbw,
                       clear_color, clear_depth, chans);
    return hipGetLastError();
}
__global__ void k_cblk_drop(uint8_t *cblk, int64_t n, uint8_t bits) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        cblk[i] &= (uint8_t)~bits;
}
hipError_t launch_cblk_drop(uint8_t *cblk, int64_t n, uint8_t bits, hipStream_t s) {
    hipLaunchKernelGGL(k_cblk_drop, dim3(grid_for(n)), dim3(256), 0, s, cblk, n, bits);
    return hipGetLastError();
}

hipError_t launch_fill_color(uint32_t *color, int64_t n, uint32_t v, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_color, dim3(grid_for(n)), dim3(256), 0, s, color, n, v);
    return hipGetLastError();
}
hipError_t launch_fill_depth(double *depth, int64_t n, double v, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_depth, dim3(grid_for(n)), dim3(256), 0, s, depth, n, v);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// SSAA resolve.  Tent: separable, one pass per axis with 8-bit fixed-point
// weights int16(k(t) * 256) of the linear kernel widened by the scale
// factor, clamped edge taps and truncating division by the weight sum; the
// horizontal pass premultiplies NRGBA by alpha.  Box: mean of each f x f
// block, rounded to nearest.

__global__ void k_resolve_h(const uint32_t *src, int32_t sw, int32_t sh, uint32_t *tmp, int32_t dw,
                            const int16_t *wx, const int32_t *ox, int32_t flx) {
    // tmp is dw x sh, premultiplied RGBA8
    int64_t n = (int64_t)dw * sh;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int32_t x = (int32_t)(i % dw), y = (int32_t)(i / dw);
        const uint32_t *row = src + (int64_t)y * sw;
        int32_t acc[4] = {0, 0, 0, 0};
        int32_t sum = 0;
        int32_t start = ox[x];
        for (int32_t k = 0; k < flx; k++) {
            int32_t c = wx[x * flx + k];
            if (c == 0) continue;
            int32_t xi = start + k;
            if (xi < 0) xi = 0;
            if (xi >= sw - 1) xi = sw - 1;
            uint32_t p = row[xi];
            int32_t a = (int32_t)(p >> 24);
            acc[0] += c * ((int32_t)(p & 0xff) * a / 0xff);
            acc[1] += c * ((int32_t)((p >> 8) & 0xff) * a / 0xff);
            acc[2] += c * ((int32_t)((p >> 16) & 0xff) * a / 0xff);
            acc[3] += c * a;
            sum += c;
        }
        uint32_t o = 0;
        for (int ch = 0; ch < 4; ch++) {
            int32_t v = sum ? acc[ch] / sum : 0;
            v = v < 0 ? 0 : (v > 255 ? 255 : v);
            o |= (uint32_t)v << (8 * ch);
        }
        tmp[i] = o;
    }
}

__global__ void k_resolve_v(const uint32_t *tmp, int32_t dw, int32_t sh, uint32_t *dst, int32_t dh,
                            const int16_t *wy, const int32_t *oy, int32_t fly) {
    int64_t n = (int64_t)dw * dh;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int32_t x = (int32_t)(i % dw), y = (int32_t)(i / dw);
        int32_t acc[4] = {0, 0, 0, 0};
        int32_t sum = 0;
        int32_t start = oy[y];
        for (int32_t k = 0; k < fly; k++) {
            int32_t c = wy[y * fly + k];
            if (c == 0) continue;
            int32_t yi = start + k;
            if (yi < 0) yi = 0;
            if (yi >= sh - 1) yi = sh - 1;
            uint32_t p = tmp[(int64_t)yi * dw + x];
            for (int ch = 0; ch < 4; ch++) acc[ch] += c * (int32_t)((p >> (8 * ch)) & 0xff);
            sum += c;
        }
        uint32_t o = 0;
        for (int ch = 0; ch < 4; ch++) {
            int32_t v = sum ? acc[ch] / sum : 0;
            v = v < 0 ? 0 : (v > 255 ? 255 : v);
            o |= (uint32_t)v << (8 * ch);
        }
        dst[i] = o;
    }
}

__global__ void k_resolve_box(const uint32_t *src, int32_t sw, uint32_t *dst, int32_t dw, int32_t dh, int32_t f) {
    int64_t n = (int64_t)dw * dh;
    const int32_t cnt = f * f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int32_t x = (int32_t)(i % dw), y = (int32_t)(i / dw);
        uint32_t acc[4] = {0, 0, 0, 0};
        for (int32_t dy = 0; dy < f; dy++) {
            const uint32_t *row = src + (int64_t)(y * f + dy) * sw + (int64_t)x * f;
            for (int32_t dx = 0; dx < f; dx++) {
                uint32_t p = row[dx];
                for (int ch = 0; ch < 4; ch++) acc[ch] += (p >> (8 * ch)) & 0xff;
            }
        }
        uint32_t o = 0;
        for (int ch = 0; ch < 4; ch++) o |= ((acc[ch] + (uint32_t)cnt / 2) / (uint32_t)cnt) << (8 * ch);
        dst[i] = o;
    }
}

// Both tent passes in one kernel: a workgroup owns RT_OW x RT_OH output
// pixels, runs the horizontal pass over exactly the input rows its vertical
// taps reach (keeping those rows in LDS), then the vertical pass.  Each row
// is computed with the same integer steps as k_resolve_h / k_resolve_v, so
// the result is identical to the two-pass form.
constexpr int RT_OW = 64, RT_OH = 4, RT_MAXROWS = 48;

__global__ __launch_bounds__(256) void k_resolve_tent(const uint32_t *src, int32_t sw, int32_t sh, uint32_t *dst,
                                                      int32_t dw, int32_t dh, const int16_t *wx, const int32_t *ox,
                                                      int32_t flx, const int16_t *wy, const int32_t *oy, int32_t fly) {
    __shared__ uint32_t hrow[RT_MAXROWS][RT_OW];
    const int x0 = blockIdx.x * RT_OW, y0 = blockIdx.y * RT_OH;
    const int nx = min(RT_OW, dw - x0), ny = min(RT_OH, dh - y0);
    const int r0 = oy[y0];
    const int nrows = oy[y0 + ny - 1] + fly - r0;
    for (int q = threadIdx.x; q < nrows * RT_OW; q += blockDim.x) {
        int j = q / RT_OW, x = q % RT_OW;
        if (x >= nx) continue;
        int ry = r0 + j;
        if (ry < 0) ry = 0;
        if (ry >= sh - 1) ry = sh - 1;
        const uint32_t *row = src + (int64_t)ry * sw;
        int32_t acc[4] = {0, 0, 0, 0};
        int32_t sum = 0;
        const int gx = x0 + x;
        const int32_t start = ox[gx];
        for (int32_t k = 0; k < flx; k++) {
            int32_t c = wx[gx * flx + k];
            if (c == 0) continue;
            int32_t xi = start + k;
            if (xi < 0) xi = 0;
            if (xi >= sw - 1) xi = sw - 1;
            uint32_t p = row[xi];
            int32_t a = (int32_t)(p >> 24);
            acc[0] += c * ((int32_t)(p & 0xff) * a / 0xff);
            acc[1] += c * ((int32_t)((p >> 8) & 0xff) * a / 0xff);
            acc[2] += c * ((int32_t)((p >> 16) & 0xff) * a / 0xff);
            acc[3] += c * a;
            sum += c;
        }
        uint32_t o = 0;
        for (int ch = 0; ch < 4; ch++) {
            int32_t v = sum ? acc[ch] / sum : 0;
            v = v < 0 ? 0 : (v > 255 ? 255 : v);
            o |= (uint32_t)v << (8 * ch);
        }
        hrow[j][x] = o;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < RT_OH * RT_OW; q += blockDim.x) {
        int yy = q / RT_OW, x = q % RT_OW;
        if (x >= nx || yy >= ny) continue;
        const int gy = y0 + yy;
        int32_t acc[4] = {0, 0, 0, 0};
        int32_t sum = 0;
        const int32_t start = oy[gy];
        for (int32_t k = 0; k < fly; k++) {
            int32_t c = wy[gy * fly + k];
            if (c == 0) continue;
            uint32_t p = hrow[start + k - r0][x];
            for (int ch = 0; ch < 4; ch++) acc[ch] += c * (int32_t)((p >> (8 * ch)) & 0xff);
            sum += c;
        }
        uint32_t o = 0;
        for (int ch = 0; ch < 4; ch++) {
            int32_t v = sum ? acc[ch] / sum : 0;
            v = v < 0 ? 0 : (v > 255 ? 255 : v);
            o |= (uint32_t)v << (8 * ch);
        }
        dst[(int64_t)gy * dw + x0 + x] = o;
    }
}

// Integer factors up to RS_F: the workgroup stages the whole source window
// its RS_OW x RS_OH outputs read (rows and columns clamped, premultiplied
// once per pixel) in LDS with coalesced loads, then runs both passes from
// LDS with the same integer steps as k_resolve_h / k_resolve_v.  Square
// 16x16 output blocks stage the fewest halo texels per output (68x68 for
// 256 outputs at 4x); at 1920x1080 from 4x they resolve in 0.065 ms against
// 0.076 ms for 64x4 blocks.
#ifndef FGL_RS_OH
#define FGL_RS_OH 16
#endif
#ifndef FGL_RS_OW
#define FGL_RS_OW 16
#endif
constexpr int RS_F = 4, RS_OW = FGL_RS_OW, RS_OH = FGL_RS_OH;
constexpr int RS_ROWS = RS_OH * RS_F + RS_F, RS_COLS = RS_OW * RS_F + RS_F;
constexpr int RS_LOADS = (RS_ROWS * RS_COLS + 255) / 256;
static_assert(RS_COLS % 4 == 0, "staged rows must be 16-byte multiples");
static_assert(RS_OW <= 256 && RS_OH * 2 * RS_F <= 256, "one pass of 256 threads fills the offset and weight tables");

FGL_D uint32_t premultiply(uint32_t p) {
    const uint32_t a = p >> 24;
    const uint32_t r = (p & 0xff) * a / 0xff, g = ((p >> 8) & 0xff) * a / 0xff, b = ((p >> 16) & 0xff) * a / 0xff;
    return r | (g << 8) | (b << 16) | (a << 24);
}

// Weighted sums are non-negative, so when the weight sum is a power of two
// (every integer factor 2^k) the truncating division is a shift.
FGL_D uint32_t tent_tap_sum(const int32_t acc[4], int32_t sum) {
    const bool pow2 = sum > 0 && (sum & (sum - 1)) == 0;
    const int sh = pow2 ? __builtin_ctz((uint32_t)sum) : 0;
    uint32_t o = 0;
    for (int ch = 0; ch < 4; ch++) {
        int32_t v = pow2 ? (acc[ch] >> sh) : (sum ? acc[ch] / sum : 0);
        v = v < 0 ? 0 : (v > 255 ? 255 : v);
        o |= (uint32_t)v << (8 * ch);
    }
    return o;
}

__global__ __launch_bounds__(256) void k_resolve_tent_lds(const uint32_t *src, int32_t sw, int32_t sh, uint32_t *dst,
                                                          int32_t dw, int32_t dh, const int16_t *wx, const int32_t *ox,
                                                          int32_t flx, const int16_t *wy, const int32_t *oy,
                                                          int32_t fly) {
    __shared__ __align__(16) uint32_t s[RS_ROWS][RS_COLS];
    __shared__ uint32_t hrow[RS_ROWS][RS_OW];
    __shared__ int16_t lwx[RS_OW][2 * RS_F], lwy[RS_OH][2 * RS_F];
    __shared__ int32_t lox[RS_OW], loy[RS_OH];
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * RS_OW, y0 = blockIdx.y * RS_OH;
    const int nx = min(RS_OW, dw - x0), ny = min(RS_OH, dh - y0);
    const int r0 = oy[y0], c0 = ox[x0];
    const int nrows = oy[y0 + ny - 1] + fly - r0;
    const int ncols = ox[x0 + nx - 1] + flx - c0;
    // stage: all loads first (independent), then the premultiplied stores
    uint32_t v[RS_LOADS];
#pragma unroll
    for (int k = 0; k < RS_LOADS; k++) {
        const int q = k * 256 + tid;
        const int j = q / RS_COLS, c = q % RS_COLS;
        const int ry = min(max(r0 + min(j, nrows - 1), 0), sh - 1);
        const int cx = min(max(c0 + min(c, ncols - 1), 0), sw - 1);
        v[k] = src[(int64_t)ry * sw + cx];
    }
    // this workgroup's rows of the weight tables
    for (int q = tid; q < RS_OW * 2 * RS_F; q += 256) {
        const int x = q / (2 * RS_F), k = q % (2 * RS_F);
        lwx[x][k] = (x < nx && k < flx) ? wx[(x0 + x) * flx + k] : (int16_t)0;
    }
    if (tid < RS_OH * 2 * RS_F) {
        const int y = tid / (2 * RS_F), k = tid % (2 * RS_F);
        lwy[y][k] = (y < ny && k < fly) ? wy[(y0 + y) * fly + k] : (int16_t)0;
    }
    if (tid < RS_OW) lox[tid] = tid < nx ? ox[x0 + tid] - c0 : 0;
    if (tid < RS_OH) loy[tid] = tid < ny ? oy[y0 + tid] - r0 : 0;
#pragma unroll
    for (int k = 0; k < RS_LOADS; k++) {
        const int q = k * 256 + tid;
        if (q < RS_ROWS * RS_COLS) s[q / RS_COLS][q % RS_COLS] = premultiply(v[k]);
    }
    __syncthreads();
    for (int q = tid; q < nrows * RS_OW; q += 256) {
        const int j = q / RS_OW, x = q % RS_OW;
        if (x >= nx) continue;
        const int base = lox[x];
        int32_t acc[4] = {0, 0, 0, 0};
        int32_t sum = 0;
        if (flx == 2 * RS_F && (base & 3) == 0) {
            // the 8 taps as two aligned 16-byte LDS reads (rows are 16-byte
            // multiples, base = RS_F * x)
            const uint4 q0 = *reinterpret_cast<const uint4 *>(&s[j][base]);
            const uint4 q1 = *reinterpret_cast<const uint4 *>(&s[j][base + 4]);
            const uint32_t t[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int32_t c = lwx[x][k];
                const uint32_t p = t[k];
                acc[0] += c * (int32_t)(p & 0xff);
                acc[1] += c * (int32_t)((p >> 8) & 0xff);
                acc[2] += c * (int32_t)((p >> 16) & 0xff);
                acc[3] += c * (int32_t)(p >> 24);
                sum += c;
            }
        } else {
            for (int32_t k = 0; k < flx; k++) {
                const int32_t c = lwx[x][k];
                const uint32_t p = s[j][base + k];
                acc[0] += c * (int32_t)(p & 0xff);
                acc[1] += c * (int32_t)((p >> 8) & 0xff);
                acc[2] += c * (int32_t)((p >> 16) & 0xff);
                acc[3] += c * (int32_t)(p >> 24);
                sum += c;
            }
        }
        hrow[j][x] = tent_tap_sum(acc, sum);
    }
    __syncthreads();
    for (int q = tid; q < RS_OH * RS_OW; q += 256) {
        const int yy = q / RS_OW, x = q % RS_OW;
        if (x >= nx || yy >= ny) continue;
        const int base = loy[yy];
        int32_t acc[4] = {0, 0, 0, 0};
        int32_t sum = 0;
        for (int32_t k = 0; k < fly; k++) {
            const int32_t c = lwy[yy][k];
            const uint32_t p = hrow[base + k][x];
            for (int ch = 0; ch < 4; ch++) acc[ch] += c * (int32_t)((p >> (8 * ch)) & 0xff);
            sum += c;
        }
        dst[(int64_t)(y0 + yy) * dw + x0 + x] = tent_tap_sum(acc, sum);
    }
}

// The 4x tent with uniform 8-bit weights (ResolveTent4): the same sums as
// k_resolve_tent_lds, computed four taps per v_dot4_u32_u8.  A workgroup
// makes T4_OW x T4_OH outputs.  It stages source rows 4*y0 - 2 .. +63 and
// columns 4*x0 - 4 .. +71 with 16-byte loads (clamped into the image),
// premultiplies them (the identity on opaque pixels), and stores them
// transposed into one byte plane per channel, four columns to a dword.
// Output column x then reads staged columns 4x+2 .. 4x+9: dwords x, x+1 and
// x+2 of each plane against the twelve packed horizontal weights.  The
// horizontal results go to planes of four rows per dword, so the vertical
// pass reads output row y's taps (rows 4y .. 4y+7) as two aligned dwords.
// Each sum is non-negative with a power-of-two weight sum, so its division
// is the shift tent_tap_sum applies.
constexpr int T4_OW = 16, T4_OH = 15;
constexpr int T4_ROWS = 4 * T4_OH + 4;         // 64 staged rows
constexpr int T4_QW = T4_OW + 2;               // 18 dwords (72 columns) per staged row and plane
constexpr int T4_QS = 20;                      // plane row stride in dwords (4 rows apart: half the banks)
constexpr int T4_QUADS = T4_ROWS * T4_QW;      // 16-byte loads per workgroup
constexpr int T4_LOADS = (T4_QUADS + 255) / 256;
static_assert(T4_QS >= T4_QW, "a plane row holds the staged columns");
static_assert(T4_OW * (T4_ROWS / 4) == 256 && T4_OW * T4_OH <= 256, "one horizontal and one vertical item a thread");

FGL_D uint32_t t4_tap(uint32_t acc, int sh) {
    const uint32_t v = acc >> sh;
    return v > 255u ? 255u : v;
}

__global__ __launch_bounds__(256) void k_resolve_tent4(const uint32_t *src, int32_t sw, int32_t sh, uint32_t *dst,
                                                       int32_t dw, int32_t dh, ResolveTent4 t, const uint8_t *cblk,
                                                       int32_t cbw, uint32_t clear_color) {
    __shared__ uint32_t pl[4][T4_ROWS][T4_QS];
    __shared__ uint32_t hp[4][T4_OW][T4_ROWS / 4 + 1];
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * T4_OW, y0 = blockIdx.y * T4_OH;
    const int ca = 4 * x0 - 4, ra = 4 * y0 - 2;
    uint4 v[T4_LOADS];
    // source blocks a draw left at the clear color (cblk: the marks of
    // deferred clears, null when there are none) are not read: a quad of an
    // interior load lies in one 32-pixel-wide block
    uint8_t mk[T4_LOADS];
#pragma unroll
    for (int k = 0; k < T4_LOADS; k++) {
        const int q = min(k * 256 + tid, T4_QUADS - 1);
        const int j = q / T4_QW, m = q % T4_QW;
        const int ry = min(max(ra + j, 0), sh - 1), c = ca + 4 * m;
        mk[k] = cblk && c >= 0 && c + 3 < sw ? cblk[(ry / CB_H) * cbw + c / CB_W] : (uint8_t)0;
    }
#pragma unroll
    for (int k = 0; k < T4_LOADS; k++) {
        const int q = min(k * 256 + tid, T4_QUADS - 1);
        const int j = q / T4_QW, m = q % T4_QW;
        const int ry = min(max(ra + j, 0), sh - 1);
        const uint32_t *row = src + (int64_t)ry * sw;
        const int c = ca + 4 * m;
        if (mk[k] & CB_COLOR) {
            v[k] = make_uint4(clear_color, clear_color, clear_color, clear_color);
        } else if (c >= 0 && c + 3 < sw) {
            v[k] = *reinterpret_cast<const uint4 *>(row + c);
        } else if (cblk) {
            // an edge quad (clamped columns): each pixel by its own block
            const int c0 = min(max(c, 0), sw - 1), c1 = min(max(c + 1, 0), sw - 1);
            const int c2 = min(max(c + 2, 0), sw - 1), c3 = min(max(c + 3, 0), sw - 1);
            const uint8_t *mrow = cblk + (ry / CB_H) * cbw;
            v[k].x = (mrow[c0 / CB_W] & CB_COLOR) ? clear_color : row[c0];
            v[k].y = (mrow[c1 / CB_W] & CB_COLOR) ? clear_color : row[c1];
            v[k].z = (mrow[c2 / CB_W] & CB_COLOR) ? clear_color : row[c2];
            v[k].w = (mrow[c3 / CB_W] & CB_COLOR) ? clear_color : row[c3];
        } else {
            v[k].x = row[min(max(c, 0), sw - 1)];
            v[k].y = row[min(max(c + 1, 0), sw - 1)];
            v[k].z = row[min(max(c + 2, 0), sw - 1)];
            v[k].w = row[min(max(c + 3, 0), sw - 1)];
        }
    }
#pragma unroll
    for (int k = 0; k < T4_LOADS; k++) {
        const int q = k * 256 + tid;
        if (q >= T4_QUADS) break;
        uint4 p = v[k];
        if (((p.x & p.y & p.z & p.w) >> 24) != 255u) {
            p.x = premultiply(p.x);
            p.y = premultiply(p.y);
            p.z = premultiply(p.z);
            p.w = premultiply(p.w);
        }
        // 4x4 byte transpose: plane ch gets byte ch of the four pixels
        const uint32_t lo01 = __builtin_amdgcn_perm(p.y, p.x, 0x05010400u);  // x0 y0 x1 y1
        const uint32_t hi01 = __builtin_amdgcn_perm(p.y, p.x, 0x07030602u);  // x2 y2 x3 y3
        const uint32_t lo23 = __builtin_amdgcn_perm(p.w, p.z, 0x05010400u);
        const uint32_t hi23 = __builtin_amdgcn_perm(p.w, p.z, 0x07030602u);
        const int j = q / T4_QW, m = q % T4_QW;
        pl[0][j][m] = __builtin_amdgcn_perm(lo23, lo01, 0x05040100u);
        pl[1][j][m] = __builtin_amdgcn_perm(lo23, lo01, 0x07060302u);
        pl[2][j][m] = __builtin_amdgcn_perm(hi23, hi01, 0x05040100u);
        pl[3][j][m] = __builtin_amdgcn_perm(hi23, hi01, 0x07060302u);
    }
    __syncthreads();
    {
        const int x = tid % T4_OW, jq = tid / T4_OW;
        uint32_t o[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int j = 4 * jq + r;
#pragma unroll
            for (int ch = 0; ch < 4; ch++) {
                uint32_t acc = __builtin_amdgcn_udot4(pl[ch][j][x], t.wx[0], 0u, false);
                acc = __builtin_amdgcn_udot4(pl[ch][j][x + 1], t.wx[1], acc, false);
                acc = __builtin_amdgcn_udot4(pl[ch][j][x + 2], t.wx[2], acc, false);
                o[ch] |= t4_tap(acc, t.shx) << (8 * r);
            }
        }
#pragma unroll
        for (int ch = 0; ch < 4; ch++) hp[ch][x][jq] = o[ch];
    }
    __syncthreads();
    {
        const int x = tid % T4_OW, yy = tid / T4_OW;
        if (yy < T4_OH && x0 + x < dw && y0 + yy < dh) {
            uint32_t o = 0;
#pragma unroll
            for (int ch = 0; ch < 4; ch++) {
                uint32_t acc = __builtin_amdgcn_udot4(hp[ch][x][yy], t.wy[0], 0u, false);
                acc = __builtin_amdgcn_udot4(hp[ch][x][yy + 1], t.wy[1], acc, false);
                o |= t4_tap(acc, t.shy) << (8 * ch);
            }
            dst[(int64_t)(y0 + yy) * dw + x0 + x] = o;
        }
    }
}

bool resolve_reads_marks(int32_t filter, int32_t sw, int32_t sh, int32_t dw, int32_t dh, const ResolveTent4 &t4) {
    return filter != 2 && t4.ok && sw == 4 * dw && sh == 4 * dh && sw % 4 == 0;
}

hipError_t launch_resolve(const uint32_t *src, int32_t sw, int32_t sh, uint32_t *tmp, uint32_t *dst, int32_t dw,
                          int32_t dh, int32_t filter, const int16_t *wx, const int32_t *ox, int32_t flx,
                          const int16_t *wy, const int32_t *oy, int32_t fly, const ResolveTent4 &t4,
                          const uint8_t *cblk, int32_t cbw, uint32_t clear_color, hipStream_t s) {
    if (filter == 2) {
        hipLaunchKernelGGL(k_resolve_box, dim3(grid_for((int64_t)dw * dh)), dim3(256), 0, s, src, sw, dst, dw, dh,
                           sw / dw);
        return hipGetLastError();
    }
    // the 4x tent with uniform weights (checked with the tables on the host);
    // rows of 16-byte multiples keep the staged loads aligned
    if (resolve_reads_marks(filter, sw, sh, dw, dh, t4)) {
        dim3 grid((dw + T4_OW - 1) / T4_OW, (dh + T4_OH - 1) / T4_OH);
        hipLaunchKernelGGL(k_resolve_tent4, grid, dim3(256), 0, s, src, sw, sh, dst, dw, dh, t4, cblk, cbw,
                           clear_color);
        return hipGetLastError();
    }
    // integer factor up to RS_F: the source window fits the LDS stage
    if (sw % dw == 0 && sh % dh == 0 && sw / dw == sh / dh && sw / dw <= RS_F && flx == 2 * (sw / dw) &&
        fly == 2 * (sh / dh)) {
        dim3 grid((dw + RS_OW - 1) / RS_OW, (dh + RS_OH - 1) / RS_OH);
        hipLaunchKernelGGL(k_resolve_tent_lds, grid, dim3(256), 0, s, src, sw, sh, dst, dw, dh, wx, ox, flx, wy, oy,
                           fly);
        return hipGetLastError();
    }
    // rows one output tile's vertical taps reach: (RT_OH - 1) * step + fly
    if ((RT_OH - 1) * (sh / dh) + fly <= RT_MAXROWS) {
        dim3 grid((dw + RT_OW - 1) / RT_OW, (dh + RT_OH - 1) / RT_OH);
        hipLaunchKernelGGL(k_resolve_tent, grid, dim3(256), 0, s, src, sw, sh, dst, dw, dh, wx, ox, flx, wy, oy, fly);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_resolve_h, dim3(grid_for((int64_t)dw * sh)), dim3(256), 0, s, src, sw, sh, tmp, dw, wx, ox,
                       flx);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_resolve_v, dim3(grid_for((int64_t)dw * dh)), dim3(256), 0, s, tmp, dw, sh, dst, dh, wy, oy,
                       fly);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// DepthImage (context.go:87-117): lo/hi over depths != MaxFloat64, then
// uint16(t * 0xffff) with t = (d - lo) / (hi - lo) (t = 1 for cleared).

constexpr double kMaxFloat64 = 1.79769313486231570814527423731704356798070e+308;

__global__ __launch_bounds__(256) void k_depth_minmax(const double *depth, int64_t n, double *partial) {
    __shared__ double slo[256], shi[256];
    double lo = kMaxFloat64, hi = -kMaxFloat64;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double d = depth[i];
        if (d == kMaxFloat64) continue;
        if (d < lo) lo = d;
        if (d > hi) hi = d;
    }
    slo[threadIdx.x] = lo;
    shi[threadIdx.x] = hi;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            if (slo[threadIdx.x + off] < slo[threadIdx.x]) slo[threadIdx.x] = slo[threadIdx.x + off];
            if (shi[threadIdx.x + off] > shi[threadIdx.x]) shi[threadIdx.x] = shi[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = slo[0];
        partial[2 * blockIdx.x + 1] = shi[0];
    }
}

__global__ void k_depth_image(const double *depth, int64_t n, const double *partial, int32_t nblocks, uint16_t *out) {
    double lo = kMaxFloat64, hi = -kMaxFloat64;
    for (int32_t b = 0; b < nblocks; b++) {
        double l = partial[2 * b], h = partial[2 * b + 1];
        if (l < lo) lo = l;
        if (h > hi) hi = h;
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double d = depth[i];
        double t = (d - lo) / (hi - lo);
        if (d == kMaxFloat64) t = 1;
        double v = t * 65535.0;
        uint32_t u = (v > -2147483649.0 && v < 2147483648.0) ? (uint32_t)(int32_t)v : 0u;
        out[i] = (uint16_t)(u & 0xffffu);
    }
}

hipError_t launch_depth_minmax(const double *depth, int64_t n, double *partial, int32_t nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_depth_minmax, dim3(nblocks), dim3(256), 0, s, depth, n, partial);
    return hipGetLastError();
}
hipError_t launch_depth_image(const double *depth, int64_t n, const double *partial, int32_t nblocks, uint16_t *out,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_depth_image, dim3(grid_for(n)), dim3(256), 0, s, depth, n, partial, nblocks, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free DepthImage: k_depth_minmax's per-block
// (min, max) partials, kDepthScratchBlocks blocks (8 KiB).  Every GPU holds
// its own copy from the library's load until the process ends, and every
// call on that GPU shares it: two such calls in flight at once on one GPU
// race on it.  Host code never passes the array's name to a kernel (on the
// host it is only a stand-in for the symbol); it passes the device address
// hipGetSymbolAddress returns, looked up once per device.

static __device__ __align__(256) double scratch[2 * kDepthScratchBlocks];

namespace {
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t depth_scratch(int32_t device, double **ptr) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        // the caller made `device` current: the address is that GPU's copy
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = static_cast<double *>(p);
    }
    *ptr = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace fgl
