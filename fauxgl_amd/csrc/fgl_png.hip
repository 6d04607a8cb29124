// fgl_png.hip -- the context's color buffer as a PNG file, the way
// SavePNG (util.go:56-63) writes an *image.NRGBA through Go's image/png:
// RGB when every alpha is 255 (NRGBA.Opaque), RGBA otherwise; 8 bits per
// sample, no interlace; each row filtered with the filter image/png picks
// (the smallest sum of |byte| taken as signed, candidates in the order Up,
// Paeth, None, Sub, Average, the first minimum kept); one zlib stream.
//
// The compressed bytes are this build's own: the filtered stream is cut
// into CHUNK-byte pieces that one thread each deflates (fixed Huffman
// codes, runs of a repeated byte as distance-1 matches) and ends with an
// empty stored block, so the pieces are byte-aligned and concatenate into
// one valid deflate stream.  The Adler-32 of each piece is combined on the
// host.  Decoding gives back the same filtered bytes and the same pixels
// as Go's file.
#include "fgl_internal.h"

namespace fgl {

// any alpha below 255: flag[0] = 1 (every writer stores the same value)
__global__ __launch_bounds__(256) void k_png_opaque(const uint32_t *color, int64_t n, uint32_t *flag) {
    bool seen = false;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        seen |= (color[i] >> 24) != 255u;
    if (__ballot(seen) != 0 && (threadIdx.x & 63) == 0) flag[0] = 1u;
}

FGL_D uint32_t png_byte(const uint32_t *row, int i, int bpp) {
    const int x = bpp == 4 ? (i >> 2) : i / 3, ch = i - x * bpp;
    return (row[x] >> (8 * ch)) & 255u;
}
FGL_D int abs8(uint32_t d) { return d < 128u ? (int)d : 256 - (int)d; }
FGL_D uint32_t paeth(uint32_t a, uint32_t b, uint32_t c) {
    int pc = (int)c;
    int pa = (int)b - pc, pb = (int)a - pc;
    pc = abs(pa + pb);
    pa = abs(pa);
    pb = abs(pb);
    if (pa <= pb && pa <= pc) return a;
    if (pb <= pc) return b;
    return c;
}

// filtered value of byte i of the current row under filter f (PNG types
// 0 None, 1 Sub, 2 Up, 3 Average, 4 Paeth); p is the previous row (NULL
// for the first row: zeros)
FGL_D uint32_t png_filtered(int f, const uint32_t *cur, const uint32_t *prev, int i, int bpp) {
    const uint32_t x = png_byte(cur, i, bpp);
    const uint32_t b = prev ? png_byte(prev, i, bpp) : 0u;
    const uint32_t a = i >= bpp ? png_byte(cur, i - bpp, bpp) : 0u;
    const uint32_t c = (prev && i >= bpp) ? png_byte(prev, i - bpp, bpp) : 0u;
    switch (f) {
    case 0: return x;
    case 1: return (x - a) & 255u;
    case 2: return (x - b) & 255u;
    case 3: return (x - ((a + b) >> 1)) & 255u;
    default: return (x - paeth(a, b, c)) & 255u;
    }
}

// one workgroup per row: the five sums, the choice, the filtered row
__global__ __launch_bounds__(256) void k_png_filter(const uint32_t *color, int32_t W, int32_t H, int32_t bpp,
                                                    uint8_t *out) {
    __shared__ int part[5][4];
    __shared__ int chosen;
    const int y = blockIdx.x;
    const uint32_t *cur = color + (int64_t)y * W;
    const uint32_t *prev = y > 0 ? cur - W : nullptr;
    const int n = W * bpp;
    int sum[5] = {0, 0, 0, 0, 0};
    for (int i = threadIdx.x; i < n; i += blockDim.x)
#pragma unroll
        for (int f = 0; f < 5; f++) sum[f] += abs8(png_filtered(f, cur, prev, i, bpp));
#pragma unroll
    for (int f = 0; f < 5; f++) {
        int v = sum[f];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if ((threadIdx.x & 63) == 0) part[f][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot[5];
        for (int f = 0; f < 5; f++) tot[f] = part[f][0] + part[f][1] + part[f][2] + part[f][3];
        static const int order[5] = {2, 4, 0, 1, 3};  // Up, Paeth, None, Sub, Average
        int best = tot[2], pick = 2;
        for (int k = 1; k < 5; k++)
            if (tot[order[k]] < best) {
                best = tot[order[k]];
                pick = order[k];
            }
        chosen = pick;
    }
    __syncthreads();
    const int f = chosen;
    uint8_t *o = out + (int64_t)y * (n + 1);
    if (threadIdx.x == 0) o[0] = (uint8_t)f;
    for (int i = threadIdx.x; i < n; i += blockDim.x) o[1 + i] = (uint8_t)png_filtered(f, cur, prev, i, bpp);
}

// ---------------------------------------------------------------------------
// deflate, fixed Huffman (RFC 1951 3.2.6), one thread per piece

struct BitOut {
    uint32_t *w;
    uint64_t acc = 0;
    int nbits = 0;
    FGL_D void put(uint32_t v, int n) {  // n <= 24, LSB first
        acc |= (uint64_t)v << nbits;
        nbits += n;
        if (nbits >= 32) {
            *w++ = (uint32_t)acc;
            acc >>= 32;
            nbits -= 32;
        }
    }
    FGL_D void put_rev(uint32_t code, int n) {  // a Huffman code, MSB first
        put(__brev(code) >> (32 - n), n);
    }
};

FGL_D void put_litlen(BitOut &bo, uint32_t v) {
    if (v < 144) bo.put_rev(0x30u + v, 8);
    else if (v < 256) bo.put_rev(0x190u + (v - 144), 9);
    else if (v < 280) bo.put_rev(v - 256, 7);
    else bo.put_rev(0xC0u + (v - 280), 8);
}

__constant__ uint16_t c_len_base[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                                        31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
__constant__ uint8_t c_len_extra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};

FGL_D void put_match(BitOut &bo, int len) {  // distance 1
    int k = 28;
    while (c_len_base[k] > len) k--;
    put_litlen(bo, 257u + (uint32_t)k);
    if (c_len_extra[k]) bo.put((uint32_t)(len - c_len_base[k]), c_len_extra[k]);
    bo.put_rev(0u, 5);  // distance code 0 = 1
}

// piece c covers bytes [c * chunk, min(n, (c + 1) * chunk)); out words at
// c * cap_words; sizes[c] = bytes written; adl[2c], adl[2c+1] = sum of the
// bytes and sum of (len - k) * byte k, both mod 65521
__global__ __launch_bounds__(64) void k_png_deflate(const uint8_t *s, int64_t n, int32_t chunk, int32_t cap_words,
                                                    int64_t npieces, uint32_t *out, uint32_t *sizes, uint32_t *adl) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= npieces) return;
    const int64_t b0 = c * chunk;
    const int len = (int)((n - b0) < chunk ? (n - b0) : chunk);
    const uint8_t *p = s + b0;
    BitOut bo;
    bo.w = out + c * cap_words;
    bo.put(0u, 1);  // BFINAL 0
    bo.put(1u, 2);  // BTYPE 01, fixed codes
    uint32_t sa = 0, sb = 0;
    int i = 0;
    while (i < len) {
        const uint32_t v = p[i];
        int r = 1;
        while (i + r < len && p[i + r] == v) r++;
        // bytes i .. i + r - 1 are v: a literal, then distance-1 matches
        for (int k = 0; k < r; k++) {
            sa += v;
            sb += (uint32_t)(len - (i + k)) * v;
            if (sb >= 0x80000000u) sb %= 65521u;
        }
        sa %= 65521u;
        put_litlen(bo, v);
        int rest = r - 1;
        while (rest >= 3) {
            int m = rest < 258 ? rest : 258;
            if (rest - m > 0 && rest - m < 3) m -= 3;
            put_match(bo, m);
            rest -= m;
        }
        for (; rest > 0; rest--) put_litlen(bo, v);
        i += r;
    }
    put_litlen(bo, 256u);  // end of block
    // an empty stored block brings the piece to a byte boundary
    bo.put(0u, 3);
    if (bo.nbits & 7) bo.put(0u, 8 - (bo.nbits & 7));
    bo.put(0x0000u, 16);
    bo.put(0xFFFFu, 16);
    const uint64_t words = bo.w - (out + c * cap_words);
    // nbits is 0 here: every field after the alignment is whole bytes and
    // the total is a multiple of 32 bits, or a partial word remains
    uint32_t bytes = (uint32_t)(words * 4);
    if (bo.nbits) {
        *bo.w = (uint32_t)bo.acc;
        bytes += (uint32_t)(bo.nbits / 8);
    }
    sizes[c] = bytes;
    adl[2 * c] = sa % 65521u;
    adl[2 * c + 1] = sb % 65521u;
}

// pieces packed back to back at their exclusive-prefix offsets
__global__ __launch_bounds__(256) void k_png_pack(const uint32_t *pieces, const uint32_t *sizes, const uint64_t *offs,
                                                  int32_t cap_words, int64_t npieces, uint8_t *dst) {
    const int64_t c = blockIdx.x;
    if (c >= npieces) return;
    const uint8_t *src = reinterpret_cast<const uint8_t *>(pieces + c * cap_words);
    uint8_t *d = dst + offs[c];
    for (uint32_t k = threadIdx.x; k < sizes[c]; k += blockDim.x) d[k] = src[k];
}

hipError_t launch_png_opaque(const uint32_t *color, int64_t n, uint32_t *flag, hipStream_t s) {
    int64_t b = (n + 255) / 256;
    hipLaunchKernelGGL(k_png_opaque, dim3((unsigned)(b < 1 ? 1 : (b > 4096 ? 4096 : b))), dim3(256), 0, s, color, n,
                       flag);
    return hipGetLastError();
}
hipError_t launch_png_filter(const uint32_t *color, int32_t W, int32_t H, int32_t bpp, uint8_t *out, hipStream_t s) {
    hipLaunchKernelGGL(k_png_filter, dim3(H), dim3(256), 0, s, color, W, H, bpp, out);
    return hipGetLastError();
}
hipError_t launch_png_deflate(const uint8_t *stream, int64_t n, int32_t chunk, int32_t cap_words, int64_t npieces,
                              uint32_t *out, uint32_t *sizes, uint32_t *adl, hipStream_t s) {
    hipLaunchKernelGGL(k_png_deflate, dim3((unsigned)((npieces + 63) / 64)), dim3(64), 0, s, stream, n, chunk,
                       cap_words, npieces, out, sizes, adl);
    return hipGetLastError();
}
hipError_t launch_png_pack(const uint32_t *pieces, const uint32_t *sizes, const uint64_t *offs, int32_t cap_words,
                           int64_t npieces, uint8_t *dst, hipStream_t s) {
    hipLaunchKernelGGL(k_png_pack, dim3((unsigned)npieces), dim3(256), 0, s, pieces, sizes, offs, cap_words, npieces,
                       dst);
    return hipGetLastError();
}

}  // namespace fgl
