// fgl_png.hip -- the context's color buffer as a PNG file, the way
// SavePNG (util.go:56-63) writes an *image.NRGBA through Go's image/png:
// RGB when every alpha is 255 (NRGBA.Opaque), RGBA otherwise; 8 bits per
// sample, no interlace; each row filtered with the filter image/png picks
// (the smallest sum of |byte| taken as signed, candidates in the order Up,
// Paeth, None, Sub, Average, the first minimum kept); one zlib stream.
//
// The compressed bytes are this build's own: the filtered stream is cut
// into CHUNK-byte pieces that one thread each deflates (runs of a repeated
// byte as distance-1 matches, the piece's own dynamic Huffman code or the
// fixed one, whichever is shorter) and ends with an empty stored block, so
// the pieces are byte-aligned and concatenate into one valid deflate
// stream.  The Adler-32 of each piece is combined on the
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
// deflate (RFC 1951), one thread per piece: fixed codes (3.2.6)

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

// ---------------------------------------------------------------------------
// dynamic Huffman codes (RFC 1951 3.2.7), one table per piece

constexpr int PNG_THREADS = 16;  // pieces per workgroup (the scratch below is per thread)
constexpr int NLIT = 286;        // literal/length symbols
constexpr int NCL = 19;          // code-length symbols

// order in which the code-length code lengths are sent (RFC 1951 3.2.7)
__constant__ uint8_t c_cl_order[NCL] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};

struct HuffScratch {
    uint32_t freq[NLIT];
    uint32_t w[NLIT];
    uint16_t sym[NLIT];
    uint16_t code[NLIT];
    uint16_t rle[NLIT + 2];  // code-length symbol | extra << 5
    uint8_t len[NLIT + 2];
    uint32_t clfreq[NCL];
    uint16_t clcode[NCL];
    uint8_t cllen[NCL];
};

// Code lengths of a Huffman code for freq[0..n) (symbols with freq 0 get
// length 0), no longer than limit: minimum-redundancy lengths by Moffat and
// Katajainen's in-place method ("In-place calculation of minimum-redundancy
// codes", 1995) on the used symbols sorted by frequency; if the longest
// exceeds limit, the frequencies are halved (rounding up) and the code is
// built again, which keeps it a complete prefix code.  At least two symbols
// always get a code, so a decoder sees a complete code.
FGL_D void huff_lengths(const uint32_t *freq, int n, int limit, uint8_t *len, uint32_t *w, uint16_t *sym) {
    int m = 0;
    for (int i = 0; i < n; i++) {
        len[i] = 0;
        if (freq[i]) sym[m++] = (uint16_t)i;
    }
    for (int i = 0; m < 2 && i < n; i++)  // pad to two symbols
        if (!freq[i] && (m == 0 || sym[0] != i)) sym[m++] = (uint16_t)i;
    // insertion sort by frequency, ascending (stable)
    for (int i = 1; i < m; i++) {
        const uint16_t v = sym[i];
        int j = i - 1;
        while (j >= 0 && freq[sym[j]] > freq[v]) {
            sym[j + 1] = sym[j];
            j--;
        }
        sym[j + 1] = v;
    }
    for (int shift = 0;; shift++) {
        for (int i = 0; i < m; i++) {
            const uint32_t f = freq[sym[i]];
            w[i] = f ? ((f - 1) >> shift) + 1 : 0;
        }
        // phase 1: internal node weights, parents in place
        w[0] += w[1];
        int root = 0, leaf = 2;
        for (int next = 1; next < m - 1; next++) {
            if (leaf >= m || w[root] < w[leaf]) {
                w[next] = w[root];
                w[root++] = (uint32_t)next;
            } else {
                w[next] = w[leaf++];
            }
            if (leaf >= m || (root < next && w[root] < w[leaf])) {
                w[next] += w[root];
                w[root++] = (uint32_t)next;
            } else {
                w[next] += w[leaf++];
            }
        }
        // phase 2: internal node depths
        w[m - 2] = 0;
        for (int next = m - 3; next >= 0; next--) w[next] = w[w[next]] + 1;
        // phase 3: leaf depths
        int avail = 1, used = 0, depth = 0, next = m - 1;
        root = m - 2;
        while (avail > 0) {
            while (root >= 0 && (int)w[root] == depth) {
                used++;
                root--;
            }
            while (avail > used) {
                w[next--] = (uint32_t)depth;
                avail--;
            }
            avail = 2 * used;
            depth++;
            used = 0;
        }
        if ((int)w[0] <= limit) break;  // w[0]: the rarest symbol, the longest code
    }
    for (int i = 0; i < m; i++) len[sym[i]] = (uint8_t)w[i];
}

// canonical codes (RFC 1951 3.2.2), bit-reversed for LSB-first output
FGL_D void huff_codes(const uint8_t *len, int n, uint16_t *code) {
    uint16_t count[16] = {0}, next[16];
    for (int i = 0; i < n; i++) count[len[i]]++;
    count[0] = 0;
    uint16_t c = 0;
    for (int b = 1; b < 16; b++) {
        c = (uint16_t)((c + count[b - 1]) << 1);
        next[b] = c;
    }
    for (int i = 0; i < n; i++)
        code[i] = len[i] ? (uint16_t)(__brev((uint32_t)next[len[i]]++) >> (32 - len[i])) : (uint16_t)0;
}

FGL_D int len_sym(int len) {  // match length 3..258 -> index into the length tables
    int k = 28;
    while (c_len_base[k] > len) k--;
    return k;
}

// The tokens of a piece: each run of one byte value is a literal and then
// distance-1 matches of up to 258 (none leaving a remainder of 1 or 2,
// which go out as literals).  f(lit) / g(match length) per token.
template <typename Lit, typename Match>
FGL_D void tokens(const uint8_t *p, int len, Lit f, Match g) {
    int i = 0;
    while (i < len) {
        const uint32_t v = p[i];
        int r = 1;
        while (i + r < len && p[i + r] == v) r++;
        f(v);
        int rest = r - 1;
        while (rest >= 3) {
            int m = rest < 258 ? rest : 258;
            if (rest - m > 0 && rest - m < 3) m -= 3;
            g(m);
            rest -= m;
        }
        for (; rest > 0; rest--) f(v);
        i += r;
    }
}

FGL_D int fixed_len(int s) { return s < 144 ? 8 : s < 256 ? 9 : s < 280 ? 7 : 8; }

// piece c covers bytes [c * chunk, min(n, (c + 1) * chunk)); out words at
// c * cap_words; sizes[c] = bytes written; adl[2c], adl[2c+1] = sum of the
// bytes and sum of (len - k) * byte k, both mod 65521.  The block uses
// fixed or dynamic codes, whichever is shorter, so a piece never exceeds the
// fixed-code bound cap_words is sized for.
__global__ __launch_bounds__(PNG_THREADS) void k_png_deflate(const uint8_t *s, int64_t n, int32_t chunk,
                                                             int32_t cap_words, int64_t npieces, uint32_t *out,
                                                             uint32_t *sizes, uint32_t *adl) {
    __shared__ HuffScratch scratch[PNG_THREADS];
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= npieces) return;
    HuffScratch &h = scratch[threadIdx.x];
    const int64_t b0 = c * chunk;
    const int len = (int)((n - b0) < chunk ? (n - b0) : chunk);
    const uint8_t *p = s + b0;

    // Adler-32 parts of the piece
    uint32_t sa = 0, sb = 0;
    for (int i = 0; i < len; i++) {
        const uint32_t v = p[i];
        sa += v;
        sb += (uint32_t)(len - i) * v;
        if (sb >= 0x80000000u) sb %= 65521u;
        if ((i & 4095) == 4095) sa %= 65521u;
    }
    sa %= 65521u;

    // symbol frequencies and the fixed-code cost
    for (int i = 0; i < NLIT; i++) h.freq[i] = 0;
    uint64_t fixed_bits = 3 + 7;  // block header, end of block
    uint64_t extra_bits = 0;      // length extra bits + distance codes (same in both forms but the distance code)
    uint32_t nmatch = 0;
    tokens(
        p, len, [&](uint32_t v) { h.freq[v]++; },
        [&](int m) {
            const int k = len_sym(m);
            h.freq[257 + k]++;
            extra_bits += c_len_extra[k];
            nmatch++;
        });
    h.freq[256] = 1;
    for (int i = 0; i < NLIT; i++) fixed_bits += (uint64_t)h.freq[i] * (uint64_t)fixed_len(i);
    fixed_bits -= 7;  // EOB counted twice
    fixed_bits += extra_bits + 5ull * nmatch;

    // dynamic code: literal/length lengths, the run-length coded length
    // sequence, the code-length code
    huff_lengths(h.freq, NLIT, 15, h.len, h.w, h.sym);
    int hlit = NLIT;
    while (hlit > 257 && h.len[hlit - 1] == 0) hlit--;
    h.len[hlit] = 1;  // the one distance code (distance 1), one bit
    const int nlen = hlit + 1;
    int nrle = 0;
    for (int i = 0; i < NCL; i++) h.clfreq[i] = 0;
    for (int i = 0; i < nlen;) {
        const int v = h.len[i];
        int r = 1;
        while (i + r < nlen && h.len[i + r] == v) r++;
        int left = r;
        if (v == 0) {
            while (left >= 11) {
                const int k = left < 138 ? left : 138;
                h.rle[nrle++] = (uint16_t)(18 | ((k - 11) << 5));
                h.clfreq[18]++;
                left -= k;
            }
            if (left >= 3) {
                h.rle[nrle++] = (uint16_t)(17 | ((left - 3) << 5));
                h.clfreq[17]++;
                left = 0;
            }
        } else {
            h.rle[nrle++] = (uint16_t)v;
            h.clfreq[v]++;
            left--;
            while (left >= 3) {
                const int k = left < 6 ? left : 6;
                h.rle[nrle++] = (uint16_t)(16 | ((k - 3) << 5));
                h.clfreq[16]++;
                left -= k;
            }
        }
        for (; left > 0; left--) {
            h.rle[nrle++] = (uint16_t)v;
            h.clfreq[v]++;
        }
        i += r;
    }
    huff_lengths(h.clfreq, NCL, 7, h.cllen, h.w, h.sym);
    int hclen = NCL;
    while (hclen > 4 && h.cllen[c_cl_order[hclen - 1]] == 0) hclen--;
    uint64_t dyn_bits = 3 + 5 + 5 + 4 + 3ull * hclen;
    for (int i = 0; i < nrle; i++) {
        const int cs = h.rle[i] & 31;
        dyn_bits += h.cllen[cs] + (cs == 16 ? 2 : cs == 17 ? 3 : cs == 18 ? 7 : 0);
    }
    for (int i = 0; i < hlit; i++) dyn_bits += (uint64_t)h.freq[i] * h.len[i];
    dyn_bits += extra_bits + 1ull * nmatch;

    BitOut bo;
    bo.w = out + c * cap_words;
    bo.put(0u, 1);  // BFINAL 0
    if (dyn_bits < fixed_bits) {
        bo.put(2u, 2);  // BTYPE 10, dynamic codes
        huff_codes(h.len, hlit, h.code);
        huff_codes(h.cllen, NCL, h.clcode);
        bo.put((uint32_t)(hlit - 257), 5);
        bo.put(0u, 5);  // HDIST - 1: one distance code
        bo.put((uint32_t)(hclen - 4), 4);
        for (int i = 0; i < hclen; i++) bo.put(h.cllen[c_cl_order[i]], 3);
        for (int i = 0; i < nrle; i++) {
            const int cs = h.rle[i] & 31, ex = h.rle[i] >> 5;
            bo.put(h.clcode[cs], h.cllen[cs]);
            if (cs == 16) bo.put((uint32_t)ex, 2);
            else if (cs == 17) bo.put((uint32_t)ex, 3);
            else if (cs == 18) bo.put((uint32_t)ex, 7);
        }
        tokens(
            p, len, [&](uint32_t v) { bo.put(h.code[v], h.len[v]); },
            [&](int m) {
                const int k = len_sym(m);
                bo.put(h.code[257 + k], h.len[257 + k]);
                if (c_len_extra[k]) bo.put((uint32_t)(m - c_len_base[k]), c_len_extra[k]);
                bo.put(0u, 1);  // distance code 0 (distance 1): the one-bit code 0
            });
        bo.put(h.code[256], h.len[256]);  // end of block
    } else {
        bo.put(1u, 2);  // BTYPE 01, fixed codes
        tokens(
            p, len, [&](uint32_t v) { put_litlen(bo, v); }, [&](int m) { put_match(bo, m); });
        put_litlen(bo, 256u);  // end of block
    }
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
    adl[2 * c] = sa;
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
    hipLaunchKernelGGL(k_png_deflate, dim3((unsigned)((npieces + PNG_THREADS - 1) / PNG_THREADS)), dim3(PNG_THREADS),
                       0, s, stream, n, chunk, cap_words, npieces, out, sizes, adl);
    return hipGetLastError();
}
hipError_t launch_png_pack(const uint32_t *pieces, const uint32_t *sizes, const uint64_t *offs, int32_t cap_words,
                           int64_t npieces, uint8_t *dst, hipStream_t s) {
    hipLaunchKernelGGL(k_png_pack, dim3((unsigned)npieces), dim3(256), 0, s, pieces, sizes, offs, cap_words, npieces,
                       dst);
    return hipGetLastError();
}

}  // namespace fgl
