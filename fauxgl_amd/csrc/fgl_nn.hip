// fgl_nn.hip -- learned anti-aliasing: a small convolutional network maps a
// 1x render (color, depth and depth-derived normals) to an estimate of the
// 4x4-supersampled image.  Six 3x3 convolutions (features -> 32 -> 32 -> 32
// -> 32 -> 32 -> 3, ReLU between), the output added to the input color.
// No reference counterpart: this replaces 16x SSAA + resolve after the draw.
//
// Tensors are pixel-major with channels contiguous (H x W x C), BF16 for
// convolution operands, FP32 accumulation on the matrix cores
// (v_mfma_f32_16x16x32_bf16 for the convolutions, 16x16x16 for the weight
// gradient).  Inference runs the layers fused in pairs (k_nn_conv2), the
// layer between kept in LDS; training runs them one by one (k_nn_conv),
// keeping every activation for the backward pass, which uses the same
// convolution kernel for the input gradient (flipped, transposed weights,
// ReLU mask in the epilogue), a weight-gradient kernel (MFMA over pixels)
// with per-wave partial sums reduced without atomics, and Adam on FP32
// master weights.
#include "fgl_internal.h"

namespace fgl {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round-to-nearest-even FP32 -> BF16 (v_cvt_pk_bf16_f32)
FGL_D uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
FGL_D float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

constexpr double kMaxF64 = 1.79769313486231570814527423731704356798070e+308;

// ---------------------------------------------------------------------------
// input features: R, G, B (NRGBA / 255), depth t in [0, 1] (DepthImage's
// normalisation; 1 where cleared), the normal of the depth surface from
// central differences (nx, ny, nz), coverage mask; channels 8..15 zero.

// one wave: min / max over the depth kernel's per-block partials
__global__ __launch_bounds__(64) void k_nn_range(const double *partial, int32_t nblocks, double *range) {
    double lo = kMaxF64, hi = -kMaxF64;
    for (int32_t b = threadIdx.x; b < nblocks; b += 64) {
        const double l = partial[2 * b], h = partial[2 * b + 1];
        lo = l < lo ? l : lo;
        hi = h > hi ? h : hi;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double l = __shfl_xor(lo, off), h = __shfl_xor(hi, off);
        lo = l < lo ? l : lo;
        hi = h > hi ? h : hi;
    }
    if (threadIdx.x == 0) {
        range[0] = lo;
        range[1] = hi;
    }
}

FGL_D float depth_tv(double d, double lo, double hi) {
    if (d == kMaxF64 || !(hi > lo)) return 1.0f;
    return (float)((d - lo) / (hi - lo));
}

// the depth and color samples a pixel's features read
struct FeatIn {
    uint32_t c;
    double d, dl, dr, du, dd;  // centre, left, right, up, down (clamped at the border)
};

FGL_D FeatIn feat_load(const uint32_t *color, const double *depth, int32_t H, int32_t W, int32_t x, int32_t y) {
    const int64_t i = (int64_t)y * W + x;
    const int32_t xl = x > 0 ? x - 1 : x, xr = x + 1 < W ? x + 1 : x;
    const int32_t yu = y > 0 ? y - 1 : y, yd = y + 1 < H ? y + 1 : y;
    FeatIn v;
    v.c = color[i];
    v.d = depth[i];
    v.dl = depth[(int64_t)y * W + xl];
    v.dr = depth[(int64_t)y * W + xr];
    v.du = depth[(int64_t)yu * W + x];
    v.dd = depth[(int64_t)yd * W + x];
    return v;
}

// the 8 feature channels as bf16 pairs
FGL_D uint4 feat_eval(const FeatIn &v, double lo, double hi) {
    const float t = depth_tv(v.d, lo, hi);
    const float gx = depth_tv(v.dr, lo, hi) - depth_tv(v.dl, lo, hi);
    const float gy = depth_tv(v.dd, lo, hi) - depth_tv(v.du, lo, hi);
    const float sx = -64.0f * gx, sy = -64.0f * gy;
    const float rn = 1.0f / sqrtf(sx * sx + sy * sy + 1.0f);
    const uint32_t c = v.c;
    uint4 f;
    f.x = f2bf((float)(c & 0xff) * (1.0f / 255.0f)) | ((uint32_t)f2bf((float)((c >> 8) & 0xff) * (1.0f / 255.0f)) << 16);
    f.y = f2bf((float)((c >> 16) & 0xff) * (1.0f / 255.0f)) | ((uint32_t)f2bf(t) << 16);
    f.z = f2bf(sx * rn) | ((uint32_t)f2bf(sy * rn) << 16);
    f.w = f2bf(rn) | ((uint32_t)f2bf(v.d == kMaxF64 ? 0.0f : 1.0f) << 16);
    return f;
}

FGL_D uint4 pixel_features(const uint32_t *color, const double *depth, double lo, double hi, int32_t H, int32_t W,
                           int32_t x, int32_t y) {
    return feat_eval(feat_load(color, depth, H, W, x, y), lo, hi);
}

__global__ __launch_bounds__(256) void k_nn_features(const uint32_t *color, const double *depth, const double *range,
                                                     int32_t H, int32_t W, uint16_t *feat) {
    const int64_t n = (int64_t)W * H;
    const double lo = range[0], hi = range[1];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t x = (int32_t)(i % W), y = (int32_t)(i / W);
        uint4 *o = reinterpret_cast<uint4 *>(feat + i * NN_FEAT_C);
        o[0] = pixel_features(color, depth, lo, hi, H, W, x, y);
        o[1] = make_uint4(0, 0, 0, 0);
    }
}

// ---------------------------------------------------------------------------
// 3x3 convolution, zero padding.  in: H x W x CIN (bf16), w: 9 x COUT x CIN
// (bf16, tap = 3 * (dy + 1) + (dx + 1)).  EPI 0: + bias, ReLU -> bf16;
// 1: + bias -> f32; 2: times (mask > 0) -> bf16 (the ReLU derivative of the
// layer below); 3: + bias, added to the input color -> NRGBA8 (inference
// output; COUT 16, channels 0..2 used).
//
// A workgroup (4 waves) computes CT_W x CT_H pixel tiles.  The input tile
// with its one-pixel halo is staged in LDS (CT_S bytes per pixel: 64 of
// channels, padded), 32 channels wide (a 16-channel input is
// zero-extended).  Each wave keeps every weight fragment of the layer in
// registers and runs v_mfma_f32_16x16x32_bf16 with A = weights (16 output
// channels x 32 input channels of one tap) and B = input (32 channels x 16
// pixels), so that the accumulator holds 4 consecutive output channels of
// one pixel per lane.  For COUT 32 the rows of tile t are channel
// 8 * (m / 4) + 4 * t + m % 4, which gives every lane 8 consecutive
// channels: one 16-byte store.
//
// The workgroups are persistent (two per CU): the grid is 8 * g
// workgroups, workgroup b runs on XCD b % 8 (round-robin dispatch) and
// walks the tiles of band b % 8, a contiguous eighth of the tile rows, so
// that the halo rows two tiles share go through one L2.  The next tile's
// input is loaded into registers before the current tile is computed and
// written to LDS after it, so that the loads' latency hides behind the
// matrix work.  The weights and bias reach their registers through LDS, so
// that no register the tile loop reads waits on a global load.

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int CT_W = 64, CT_H = 8, CT_LW = CT_W + 2, CT_LH = CT_H + 2, CT_S = 80;
constexpr int CT_PIX = CT_LW * CT_LH;
constexpr int CT_WG_PER_XCD = 64;  // 2 resident workgroups (register-bound) x 32 CUs

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
FGL_D uint32_t pack_bf2(float a, float b) {
    bf16x2 v;
    v[0] = (__bf16)a;
    v[1] = (__bf16)b;
    return __builtin_bit_cast(uint32_t, v);
}
FGL_D float relu(float v) { return v > 0.0f ? v : 0.0f; }
// pack_bf2(relu(a), relu(b)) for every a, b but NaN, in fewer instructions:
// rounding to BF16 keeps the sign, and a BF16 is negative (or -0) exactly
// when its bits are negative as a 16-bit integer
typedef short i16x2 __attribute__((ext_vector_type(2)));
FGL_D uint32_t pack_relu_bf2(float a, float b) {
    uint32_t u = pack_bf2(a, b);
    asm("" : "+v"(u));  // keeps the two conversions one v_cvt_pk_bf16_f32
    const i16x2 v = __builtin_bit_cast(i16x2, u);
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(v, i16x2{0, 0}));
}
FGL_D float keep_if(uint32_t m16, float v) { return ((m16 & 0x8000u) == 0 && (m16 & 0x7fffu) != 0) ? v : 0.0f; }

template <int CIN, int COUT, int EPI, bool FEAT>
__global__ __launch_bounds__(256) void k_nn_conv(const uint16_t *in, const uint16_t *w, const float *bias, int32_t H,
                                                 int32_t W, void *out, const uint16_t *mask, const uint32_t *color,
                                                 const double *depth, const double *range) {
    static_assert(CIN == 16 || CIN == 32, "CIN");
    static_assert(COUT == 16 || COUT == 32, "COUT");
    static_assert(!FEAT || CIN == 16, "features are 16 channels");
    static_assert(EPI != 3 || COUT == 16, "EPI 3 needs COUT 16");
    constexpr int NT = COUT / 16, CH = CIN / 8;
    constexpr int ITEMS = CT_PIX * CH, ITER = (ITEMS + 255) / 256;
    __shared__ __attribute__((aligned(16))) uint8_t lds[CT_PIX * CT_S];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int pm = lane & 15, kg = lane >> 4;

    const int ntx = (W + CT_W - 1) / CT_W, nty = (H + CT_H - 1) / CT_H;
    const int band = blockIdx.x & 7, g = gridDim.x >> 3;
    const int by0 = (band * nty) >> 3, by1 = ((band + 1) * nty) >> 3;
    const int nband = (by1 - by0) * ntx;
    int j = blockIdx.x >> 3;
    if (j >= nband) return;

    // weights and bias through LDS into registers; lane (pm, kg) holds
    // w[tap][row pm of tile t][8 kg .. 8 kg + 7]
    constexpr int WBYTES = 9 * COUT * CIN * 2;
    static_assert(WBYTES + COUT * 4 <= CT_PIX * CT_S, "weights fit the tile area");
    for (int q = tid; q < WBYTES / 16; q += 256)
        *reinterpret_cast<uint4 *>(lds + 16 * q) = reinterpret_cast<const uint4 *>(w)[q];
    if (EPI != 2)
        for (int q = tid; q < COUT; q += 256) *reinterpret_cast<float *>(lds + WBYTES + 4 * q) = bias[q];
    __syncthreads();
    bf16x8 A[9][NT];
#pragma unroll
    for (int tap = 0; tap < 9; tap++)
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int co = NT == 2 ? 8 * (pm >> 2) + 4 * t + (pm & 3) : pm;
            if (CIN == 32 || kg < 2) {
                A[tap][t] = *reinterpret_cast<const bf16x8 *>(lds + 2 * ((tap * COUT + co) * CIN + 8 * kg));
            } else {
                A[tap][t] = bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
            }
        }
    const int ch0 = NT == 2 ? 8 * kg : 4 * kg;  // first output channel this lane holds
    float bb[4 * NT];
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int q = 0; q < 4; q++)
            bb[4 * t + q] = (EPI == 2) ? 0.0f : *reinterpret_cast<const float *>(lds + WBYTES + 4 * (ch0 + 4 * t + q));
    __syncthreads();
    // channels the input does not have stay zero in LDS for every tile
    for (int p = tid; p < CT_PIX; p += 256) {
        if (FEAT) *reinterpret_cast<uint4 *>(lds + p * CT_S + 16) = make_uint4(0, 0, 0, 0);
        if (CIN == 16) {
            *reinterpret_cast<uint4 *>(lds + p * CT_S + 32) = make_uint4(0, 0, 0, 0);
            *reinterpret_cast<uint4 *>(lds + p * CT_S + 48) = make_uint4(0, 0, 0, 0);
        }
    }

    // the input of tile jj: clamped unconditional 16-byte loads, zero
    // outside the image
    auto fetch = [&](int jj, uint4 *v) {
        const int x0 = (jj % ntx) * CT_W, y0 = (by0 + jj / ntx) * CT_H;
#pragma unroll
        for (int it = 0; it < ITER; it++) {
            const int idx = tid + it * 256;
            const int p = (idx < ITEMS ? idx : ITEMS - 1) / CH, c = idx % CH;
            const int ry = p / CT_LW, rx = p - ry * CT_LW;
            const int gy = y0 - 1 + ry, gx = x0 - 1 + rx;
            const int cy = gy < 0 ? 0 : (gy >= H ? H - 1 : gy), cx = gx < 0 ? 0 : (gx >= W ? W - 1 : gx);
            v[it] = *reinterpret_cast<const uint4 *>(in + ((int64_t)cy * W + cx) * CIN + c * 8);
            if (gy != cy || gx != cx) v[it] = make_uint4(0, 0, 0, 0);
        }
    };
    auto put = [&](const uint4 *v) {
#pragma unroll
        for (int it = 0; it < ITER; it++) {
            const int idx = tid + it * 256;
            if (idx < ITEMS) *reinterpret_cast<uint4 *>(lds + (idx / CH) * CT_S + (idx % CH) * 16) = v[it];
        }
    };
    // layer 0 of inference: the features of each tile pixel, computed here
    // from color and depth (k_nn_features' values, bit for bit)
    auto put_features = [&](int jj) {
        const int x0 = (jj % ntx) * CT_W, y0 = (by0 + jj / ntx) * CT_H;
        const double lo = range[0], hi = range[1];
        for (int p = tid; p < CT_PIX; p += 256) {
            const int ry = p / CT_LW, rx = p - ry * CT_LW;
            const int gy = y0 - 1 + ry, gx = x0 - 1 + rx;
            uint4 f = make_uint4(0, 0, 0, 0);
            if (gy >= 0 && gy < H && gx >= 0 && gx < W) f = pixel_features(color, depth, lo, hi, H, W, gx, gy);
            *reinterpret_cast<uint4 *>(lds + p * CT_S) = f;
        }
    };

    if (FEAT) {
        put_features(j);
    } else {
        uint4 v[ITER];
        fetch(j, v);
        put(v);
    }
    __syncthreads();
    for (;;) {
        const int x0 = (j % ntx) * CT_W, y0 = (by0 + j / ntx) * CT_H;
        const int jn = j + g;
        uint4 v[FEAT ? 1 : ITER];
        if (!FEAT && jn < nband) fetch(jn, v);

        for (int rr = 0; rr < 2; rr++) {
            const int r = wave * 2 + rr, y = y0 + r;
            if (y >= H) break;
            for (int xb = 0; xb < CT_W; xb += 16) {
                if (x0 + xb >= W) break;
                f32x4 acc[NT];
#pragma unroll
                for (int t = 0; t < NT; t++) acc[t] = f32x4{bb[4 * t], bb[4 * t + 1], bb[4 * t + 2], bb[4 * t + 3]};
#pragma unroll
                for (int tap = 0; tap < 9; tap++) {
                    const bf16x8 b = *reinterpret_cast<const bf16x8 *>(
                        lds + ((r + tap / 3) * CT_LW + xb + pm + tap % 3) * CT_S + 16 * kg);
#pragma unroll
                    for (int t = 0; t < NT; t++)
                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[tap][t], b, acc[t], 0, 0, 0);
                }
                const int x = x0 + xb + pm;
                if (x >= W) continue;
                const int64_t px = (int64_t)y * W + x;
                if (EPI == 0 || EPI == 2) {
                    float o[4 * NT];
#pragma unroll
                    for (int t = 0; t < NT; t++)
#pragma unroll
                        for (int q = 0; q < 4; q++) o[4 * t + q] = acc[t][q];
                    uint16_t *dst = static_cast<uint16_t *>(out) + px * COUT + ch0;
                    if (EPI == 0) {
#pragma unroll
                        for (int k = 0; k < 4 * NT; k++) o[k] = relu(o[k]);
                    } else if (NT == 2) {
                        const uint4 m = *reinterpret_cast<const uint4 *>(mask + px * COUT + ch0);
                        const uint32_t mw[4] = {m.x, m.y, m.z, m.w};
#pragma unroll
                        for (int k = 0; k < 8; k++) o[k] = keep_if((mw[k >> 1] >> (16 * (k & 1))) & 0xffffu, o[k]);
                    } else {
                        const uint2 m = *reinterpret_cast<const uint2 *>(mask + px * COUT + ch0);
                        const uint32_t mw[2] = {m.x, m.y};
#pragma unroll
                        for (int k = 0; k < 4; k++) o[k] = keep_if((mw[k >> 1] >> (16 * (k & 1))) & 0xffffu, o[k]);
                    }
                    if (NT == 2) {
                        *reinterpret_cast<uint4 *>(dst) = make_uint4(pack_bf2(o[0], o[1]), pack_bf2(o[2], o[3]),
                                                                     pack_bf2(o[4], o[5]), pack_bf2(o[6], o[7]));
                    } else {
                        *reinterpret_cast<uint2 *>(dst) = make_uint2(pack_bf2(o[0], o[1]), pack_bf2(o[2], o[3]));
                    }
                } else if (EPI == 1) {
                    float *dst = static_cast<float *>(out) + px * COUT + ch0;
#pragma unroll
                    for (int t = 0; t < NT; t++)
                        *reinterpret_cast<float4 *>(dst + 4 * t) =
                            make_float4(acc[t][0], acc[t][1], acc[t][2], acc[t][3]);
                } else if (kg == 0) {
                    const uint32_t c = color[px];
                    uint32_t o = c & 0xff000000u;
#pragma unroll
                    for (int ch = 0; ch < 3; ch++) {
                        float f = (float)((c >> (8 * ch)) & 0xff) * (1.0f / 255.0f) + acc[0][ch];
                        f = f < 0.0f ? 0.0f : (f > 1.0f ? 1.0f : f);
                        o |= (uint32_t)(f * 255.0f + 0.5f) << (8 * ch);
                    }
                    static_cast<uint32_t *>(out)[px] = o;
                }
            }
        }
        if (jn >= nband) break;
        __syncthreads();  // every wave is done reading this tile
        if (FEAT) {
            put_features(jn);
        } else {
            put(v);
        }
        __syncthreads();
        j = jn;
    }
}

// ---------------------------------------------------------------------------
// Two layers in one kernel (inference): layer A (CIN_A -> 32, ReLU) and
// layer B (32 -> COUT_B, epilogue EPI_B as in k_nn_conv) on an F2_W x F2_H
// output tile.  The input tile (two-pixel halo) and layer A's output (one-
// pixel halo, zero outside the image: the padding layer B sees) stay in
// LDS, so that the intermediate activation never goes to HBM.  One
// persistent workgroup of 8 waves per CU; both layers' weights sit in LDS
// (rows padded to F2_WS bytes) and reach registers from there, once per
// kernel where the register budget allows, else once per phase.
// Each wave computes a strip of rows of one 16-pixel column block, reading
// every input row's three shifted fragments once for the up to three
// output rows that use it.  The multiply-accumulate order per output is
// bias, then taps 0..8, as in k_nn_conv: both kernels give the same bits.

constexpr int F2_W = 62, F2_H = 8;                  // output tile
constexpr int F2_IW = F2_W + 4, F2_IH = F2_H + 4;   // input tile
constexpr int F2_MW = F2_W + 4, F2_MH = F2_H + 2;   // intermediate: 64 computed columns + 2 padding
constexpr int F2_S = 80, F2_WS = 80;
constexpr int F2_WG_PER_XCD = 32;                   // one workgroup per CU

template <int CIN_A, int COUT_B, int EPI_B, bool FEAT>
__global__ __launch_bounds__(512) void k_nn_conv2(const uint16_t *in, const uint16_t *wa, const float *bias_a,
                                                  const uint16_t *wb, const float *bias_b, int32_t H, int32_t W,
                                                  void *out, const uint32_t *color, const double *depth,
                                                  const double *range, int32_t nrange) {
    static_assert(CIN_A == 16 || CIN_A == 32, "CIN_A");
    static_assert((COUT_B == 32 && EPI_B == 0) || (COUT_B == 16 && (EPI_B == 1 || EPI_B == 3)), "layer B");
    static_assert(!FEAT || CIN_A == 16, "features are 16 channels");
    constexpr int CH = CIN_A / 8, NTB = COUT_B / 16;
    constexpr int IN_BYTES = F2_IW * F2_IH * F2_S, MID_BYTES = F2_MW * F2_MH * F2_S;
    constexpr int WA_S = CIN_A == 16 ? 48 : F2_WS;
    constexpr int WA_BYTES = 9 * 32 * WA_S, WB_BYTES = 9 * COUT_B * F2_WS;
    constexpr int ITEMS = F2_IW * F2_IH * CH, ITER = (ITEMS + 511) / 512;
    __shared__ __attribute__((aligned(16))) uint8_t lds[IN_BYTES + MID_BYTES + WA_BYTES + WB_BYTES + 64 * 4];
    uint8_t *const lin = lds, *const lmid = lds + IN_BYTES, *const lwa = lmid + MID_BYTES, *const lwb = lwa + WA_BYTES;
    float *const lbias = reinterpret_cast<float *>(lwb + WB_BYTES);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int pm = lane & 15, kg = lane >> 4;
    const int blk = wave & 3, half = wave >> 2;

    const int ntx = (W + F2_W - 1) / F2_W, nty = (H + F2_H - 1) / F2_H;
    const int band = blockIdx.x & 7, g = gridDim.x >> 3;
    const int by0 = (band * nty) >> 3, by1 = ((band + 1) * nty) >> 3;
    const int nband = (by1 - by0) * ntx;
    int j = blockIdx.x >> 3;
    if (j >= nband) return;

    // both layers' weights and biases into LDS, once
    for (int q = tid; q < 9 * 32 * CH; q += 512)
        *reinterpret_cast<uint4 *>(lwa + (q / CH) * WA_S + (q % CH) * 16) = reinterpret_cast<const uint4 *>(wa)[q];
    for (int q = tid; q < 9 * COUT_B * 4; q += 512)
        *reinterpret_cast<uint4 *>(lwb + (q >> 2) * F2_WS + (q & 3) * 16) = reinterpret_cast<const uint4 *>(wb)[q];
    if (tid < 32) lbias[tid] = bias_a[tid];
    if (tid < COUT_B) lbias[32 + tid] = bias_b[tid];
    // channels the input does not have stay zero for every tile
    for (int p = tid; p < F2_IW * F2_IH; p += 512) {
        if (FEAT) *reinterpret_cast<uint4 *>(lin + p * F2_S + 16) = make_uint4(0, 0, 0, 0);
        if (CIN_A == 16) {
            *reinterpret_cast<uint4 *>(lin + p * F2_S + 32) = make_uint4(0, 0, 0, 0);
            *reinterpret_cast<uint4 *>(lin + p * F2_S + 48) = make_uint4(0, 0, 0, 0);
        }
    }

    auto fetch = [&](int jj, uint4 *v) {
        const int x0 = (jj % ntx) * F2_W, y0 = (by0 + jj / ntx) * F2_H;
#pragma unroll
        for (int it = 0; it < ITER; it++) {
            const int idx = tid + it * 512;
            const int p = (idx < ITEMS ? idx : ITEMS - 1) / CH, c = idx % CH;
            const int ry = p / F2_IW, rx = p - ry * F2_IW;
            const int gy = y0 - 2 + ry, gx = x0 - 2 + rx;
            const int cy = gy < 0 ? 0 : (gy >= H ? H - 1 : gy), cx = gx < 0 ? 0 : (gx >= W ? W - 1 : gx);
            v[it] = *reinterpret_cast<const uint4 *>(in + ((int64_t)cy * W + cx) * CIN_A + c * 8);
            if (gy != cy || gx != cx) v[it] = make_uint4(0, 0, 0, 0);
        }
    };
    auto put = [&](const uint4 *v) {
#pragma unroll
        for (int it = 0; it < ITER; it++) {
            const int idx = tid + it * 512;
            if (idx < ITEMS) *reinterpret_cast<uint4 *>(lin + (idx / CH) * F2_S + (idx % CH) * 16) = v[it];
        }
    };
    // layer 0 of inference: each tile pixel's depth and color samples are
    // loaded before the current tile is computed and turned into features
    // after it (k_nn_features' values, bit for bit)
    constexpr int FITER = (F2_IW * F2_IH + 511) / 512;
    // depth range: range = {lo, hi}, or nrange per-block {lo, hi} partials
    // of k_depth_minmax reduced here (k_nn_range's result, exactly)
    __shared__ double srange[2][8];
    double rlo = 0.0, rhi = 0.0;
    if (FEAT) {
        if (nrange > 0) {
            double lo = kMaxF64, hi = -kMaxF64;
            for (int q = tid; q < nrange; q += 512) {
                const double l = range[2 * q], h = range[2 * q + 1];
                lo = l < lo ? l : lo;
                hi = h > hi ? h : hi;
            }
            for (int off = 32; off > 0; off >>= 1) {
                const double l = __shfl_xor(lo, off), h = __shfl_xor(hi, off);
                lo = l < lo ? l : lo;
                hi = h > hi ? h : hi;
            }
            if (lane == 0) {
                srange[0][wave] = lo;
                srange[1][wave] = hi;
            }
            __syncthreads();
            lo = srange[0][0];
            hi = srange[1][0];
            for (int w = 1; w < 8; w++) {
                lo = srange[0][w] < lo ? srange[0][w] : lo;
                hi = srange[1][w] > hi ? srange[1][w] : hi;
            }
            rlo = lo;
            rhi = hi;
        } else {
            rlo = range[0];
            rhi = range[1];
        }
    }
    auto fetch_features = [&](int jj, FeatIn *raw) {
        const int x0 = (jj % ntx) * F2_W, y0 = (by0 + jj / ntx) * F2_H;
#pragma unroll
        for (int it = 0; it < FITER; it++) {
            const int p0 = tid + it * 512, p = p0 < F2_IW * F2_IH ? p0 : F2_IW * F2_IH - 1;
            const int ry = p / F2_IW, rx = p - ry * F2_IW;
            const int gy = y0 - 2 + ry, gx = x0 - 2 + rx;
            const int cy = gy < 0 ? 0 : (gy >= H ? H - 1 : gy), cx = gx < 0 ? 0 : (gx >= W ? W - 1 : gx);
            raw[it] = feat_load(color, depth, H, W, cx, cy);
        }
    };
    auto put_features = [&](int jj, const FeatIn *raw) {
        const int x0 = (jj % ntx) * F2_W, y0 = (by0 + jj / ntx) * F2_H;
#pragma unroll
        for (int it = 0; it < FITER; it++) {
            const int p = tid + it * 512;
            if (p >= F2_IW * F2_IH) continue;
            const int ry = p / F2_IW, rx = p - ry * F2_IW;
            const int gy = y0 - 2 + ry, gx = x0 - 2 + rx;
            uint4 f = make_uint4(0, 0, 0, 0);
            if (gy >= 0 && gy < H && gx >= 0 && gx < W) f = feat_eval(raw[it], rlo, rhi);
            *reinterpret_cast<uint4 *>(lin + p * F2_S) = f;
        }
    };

    __syncthreads();  // weights staged before the first tile goes in
    if (FEAT) {
        FeatIn raw0[FITER];
        fetch_features(j, raw0);
        put_features(j, raw0);
    } else {
        uint4 v0[ITER];
        fetch(j, v0);
        put(v0);
    }
    __syncthreads();

    // weight fragments: lane (pm, kg) holds w[tap][row pm of tile t][8 kg .. 8 kg + 7]
    auto load_a = [&](bf16x8 (&A)[9][2]) {
#pragma unroll
        for (int tap = 0; tap < 9; tap++)
#pragma unroll
            for (int t = 0; t < 2; t++) {
                const int co = 8 * (pm >> 2) + 4 * t + (pm & 3);
                if (CIN_A == 32 || kg < 2)
                    A[tap][t] = *reinterpret_cast<const bf16x8 *>(lwa + (tap * 32 + co) * WA_S + 16 * kg);
                else
                    A[tap][t] = bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
            }
    };
    auto load_b = [&](bf16x8 (&B)[9][NTB]) {
#pragma unroll
        for (int tap = 0; tap < 9; tap++)
#pragma unroll
            for (int t = 0; t < NTB; t++) {
                const int co = NTB == 2 ? 8 * (pm >> 2) + 4 * t + (pm & 3) : pm;
                B[tap][t] = *reinterpret_cast<const bf16x8 *>(lwb + (tap * COUT_B + co) * F2_WS + 16 * kg);
            }
    };
    // Both layers' fragments stay in registers across tiles where that fits
    // the register budget (a 16-channel side makes one layer small);
    // otherwise each phase reloads its layer's fragments from LDS.
    constexpr bool HOIST = CIN_A == 16 || COUT_B == 16;
    bf16x8 A[9][2], B[9][NTB];
    if (HOIST) {
        load_a(A);
        load_b(B);
    }

    for (;;) {
        const int x0 = (j % ntx) * F2_W, y0 = (by0 + j / ntx) * F2_H;
        const int jn = j + g;
        uint4 v[FEAT ? 1 : ITER];
        FeatIn raw[FEAT ? FITER : 1];
        if (jn < nband) {
            if (FEAT) fetch_features(jn, raw);
            else fetch(jn, v);
        }

        // ---- layer A: intermediate rows 5 * half .. + 4, columns 16 * blk .. + 15
        {
            if (!HOIST) load_a(A);
            const int r0 = 5 * half;
            f32x4 acc[5][2];
#pragma unroll
            for (int q = 0; q < 5; q++)
#pragma unroll
                for (int t = 0; t < 2; t++)
                    acc[q][t] = f32x4{lbias[8 * kg + 4 * t], lbias[8 * kg + 4 * t + 1], lbias[8 * kg + 4 * t + 2],
                                      lbias[8 * kg + 4 * t + 3]};
            // input row r0 + k feeds output row r0 + k - dy with tap row dy;
            // for each output row the taps arrive in order 0..8
#pragma unroll
            for (int k = 0; k < 7; k++)
#pragma unroll
                for (int dx = 0; dx < 3; dx++) {
                    const bf16x8 f = *reinterpret_cast<const bf16x8 *>(
                        lin + ((r0 + k) * F2_IW + 16 * blk + pm + dx) * F2_S + 16 * kg);
#pragma unroll
                    for (int dy = 0; dy < 3; dy++) {
                        const int q = k - dy;
                        if (q < 0 || q > 4) continue;
#pragma unroll
                        for (int t = 0; t < 2; t++)
                            acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[3 * dy + dx][t], f, acc[q][t], 0, 0, 0);
                    }
                }
#pragma unroll
            for (int q = 0; q < 5; q++) {
                const int iy = r0 + q, ix = 16 * blk + pm;
                const int gy = y0 - 1 + iy, gx = x0 - 1 + ix;
                uint4 o = make_uint4(0, 0, 0, 0);
                if (gy >= 0 && gy < H && gx >= 0 && gx < W)
                    o = make_uint4(pack_bf2(relu(acc[q][0][0]), relu(acc[q][0][1])),
                                   pack_bf2(relu(acc[q][0][2]), relu(acc[q][0][3])),
                                   pack_bf2(relu(acc[q][1][0]), relu(acc[q][1][1])),
                                   pack_bf2(relu(acc[q][1][2]), relu(acc[q][1][3])));
                *reinterpret_cast<uint4 *>(lmid + (iy * F2_MW + ix) * F2_S + 16 * kg) = o;
            }
        }
        __syncthreads();

        // ---- layer B: output rows 4 * half .. + 3, columns 16 * blk .. + 15
        {
            if (!HOIST) load_b(B);
            const int ch0 = NTB == 2 ? 8 * kg : 4 * kg;
            const int r0 = 4 * half;
            f32x4 acc[4][NTB];
#pragma unroll
            for (int q = 0; q < 4; q++)
#pragma unroll
                for (int t = 0; t < NTB; t++)
                    acc[q][t] = f32x4{lbias[32 + ch0 + 4 * t], lbias[32 + ch0 + 4 * t + 1], lbias[32 + ch0 + 4 * t + 2],
                                      lbias[32 + ch0 + 4 * t + 3]};
#pragma unroll
            for (int k = 0; k < 6; k++)
#pragma unroll
                for (int dx = 0; dx < 3; dx++) {
                    const bf16x8 f = *reinterpret_cast<const bf16x8 *>(
                        lmid + ((r0 + k) * F2_MW + 16 * blk + pm + dx) * F2_S + 16 * kg);
#pragma unroll
                    for (int dy = 0; dy < 3; dy++) {
                        const int q = k - dy;
                        if (q < 0 || q > 3) continue;
#pragma unroll
                        for (int t = 0; t < NTB; t++)
                            acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(B[3 * dy + dx][t], f, acc[q][t], 0, 0, 0);
                    }
                }
            const int ox = 16 * blk + pm, x = x0 + ox;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int y = y0 + r0 + q;
                if (ox >= F2_W || x >= W || y >= H) continue;
                const int64_t px = (int64_t)y * W + x;
                if (EPI_B == 0) {
                    *reinterpret_cast<uint4 *>(static_cast<uint16_t *>(out) + px * COUT_B + ch0) =
                        make_uint4(pack_bf2(relu(acc[q][0][0]), relu(acc[q][0][1])),
                                   pack_bf2(relu(acc[q][0][2]), relu(acc[q][0][3])),
                                   pack_bf2(relu(acc[q][NTB - 1][0]), relu(acc[q][NTB - 1][1])),
                                   pack_bf2(relu(acc[q][NTB - 1][2]), relu(acc[q][NTB - 1][3])));
                } else if (EPI_B == 1) {
                    *reinterpret_cast<float4 *>(static_cast<float *>(out) + px * COUT_B + ch0) =
                        make_float4(acc[q][0][0], acc[q][0][1], acc[q][0][2], acc[q][0][3]);
                } else if (kg == 0) {
                    const uint32_t c = color[px];
                    uint32_t o = c & 0xff000000u;
#pragma unroll
                    for (int ch = 0; ch < 3; ch++) {
                        float f = (float)((c >> (8 * ch)) & 0xff) * (1.0f / 255.0f) + acc[q][0][ch];
                        f = f < 0.0f ? 0.0f : (f > 1.0f ? 1.0f : f);
                        o |= (uint32_t)(f * 255.0f + 0.5f) << (8 * ch);
                    }
                    static_cast<uint32_t *>(out)[px] = o;
                }
            }
        }
        if (jn >= nband) break;
        // every wave is past layer A, so the input buffer is free
        if (FEAT) {
            put_features(jn, raw);
        } else {
            put(v);
        }
        __syncthreads();  // next input in place; every wave is done with the intermediate
        j = jn;
    }
}

// ---------------------------------------------------------------------------
// All six layers in one kernel (inference), streamed down vertical strips.
//
// A workgroup owns a strip of NS_OW output columns and a band of output
// rows.  Every layer is computed on NS_CW = NS_OW + 12 columns (the six
// layers' one-pixel halos; a layer's result is exact on one column fewer at
// each side than its input), and the band is walked one row per step: at
// step t the feature stage writes feature row y0 - 6 + t, layer l writes its
// output row y0 - 8 - 2 l + t.  Layer l's input lives in a ring of four rows
// in LDS: the three rows its taps read at step t were written at steps
// before, and the fourth slot takes the row its producer writes in the same
// step, so one barrier per step orders everything.  No activation goes to
// HBM: the kernel reads the 1x color and depth and writes the final color
// (or the raw output).
//
// Each wave owns one stage and keeps that layer's weight fragments in
// registers: waves 0-3 layers 1-4 (a whole row: NS_NB blocks of 16 pixels,
// both 16-channel output tiles), wave 4 layer 0, waves 5 and 7 half of
// layer 5's blocks each, wave 6 the features.  Rows outside the image are
// written as zeros and columns outside it are zeroed, which is the zero
// padding every layer sees; the ring columns left and right of the strip
// stay zero.  Layers 1-5 accumulate bias, then taps 0..8 (k_nn_conv's
// order, the same bits).  Layer 0 has 8 used input channels: one MFMA
// covers four taps (lane group kg holds tap 4 g + kg), three MFMAs a tile,
// so its sums are grouped differently from k_nn_conv's (within rounding).
//
// Ring pixel p holds local column p - 1.  A row is four planes, one per
// group of 8 channels, of 16 bytes per pixel: the 16 lanes that read (or
// write) one channel group of 16 consecutive pixels touch 256 contiguous
// bytes, whatever the pixel offset, so no access has a bank conflict.

constexpr int NS_NB = 6, NS_CW = 16 * NS_NB, NS_OW = NS_CW - 12, NS_RW = NS_CW + 2, NS_ARW = NS_RW;
constexpr int NS_ACT_ROW = NS_ARW * 64, NS_FEAT_ROW = NS_RW * 16;
constexpr int NS_ACT_RING = 4 * NS_ACT_ROW, NS_FEAT_RING = 4 * NS_FEAT_ROW;
constexpr int NS_COLOR_ROW = NS_CW * 4, NS_COLOR_RING = 16 * NS_COLOR_ROW;
constexpr int NS_LDS = 5 * NS_ACT_RING + NS_FEAT_RING + NS_COLOR_RING;
constexpr int NS_FEAT_PER_LANE = (NS_CW + 63) / 64;

FGL_D int ns_chunk(int p, int kg) { return (kg * NS_ARW + p) * 16; }

// The nine tap fragments of block b from the three ring rows (tap =
// 3 dy + dx reads pixel 16 b + pm + dx of row dy).
FGL_D void ns_load9(const uint8_t *const (&rows)[3], int b, int pm, int kg, bf16x8 (&f)[9]) {
#pragma unroll
    for (int dy = 0; dy < 3; dy++)
#pragma unroll
        for (int dx = 0; dx < 3; dx++)
            f[3 * dy + dx] = *reinterpret_cast<const bf16x8 *>(rows[dy] + ns_chunk(16 * b + pm + dx, kg));
}

// Scheduling of one block: the next block's nine fragment loads spread
// between this block's MFMAs (one load, then two MFMAs), so that their LDS
// latency hides behind the matrix work.
FGL_D void ns_interleave() {
#pragma unroll
    for (int i = 0; i < 9; i++) {
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // DS read
        __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);  // MFMA
    }
}

template <int EPI>
__global__ __launch_bounds__(512) void k_nn_stream(const uint32_t *color, const double *depth, const double *range,
                                                   int32_t nrange, const uint16_t *w0, const uint16_t *w1,
                                                   const uint16_t *w2, const uint16_t *w3, const uint16_t *w4,
                                                   const uint16_t *w5, const float *b0, const float *b1,
                                                   const float *b2, const float *b3, const float *b4, const float *b5,
                                                   int32_t H, int32_t W, int32_t band_h, int32_t nstrips, void *out) {
    static_assert(EPI == 1 || EPI == 3, "EPI");
    static_assert(NS_LDS <= 160 * 1024, "rings fit the LDS");
    __shared__ __attribute__((aligned(16))) uint8_t lds[NS_LDS];
    // ring k (k = 1..5) is layer k's input; the feature ring is layer 0's
    uint8_t *const feat_ring = lds + 5 * NS_ACT_RING;
    uint8_t *const color_ring = feat_ring + NS_FEAT_RING;  // input colors, 16 rows (EPI 3)
    auto act_ring = [&](int k) { return lds + (k - 1) * NS_ACT_RING; };

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int pm = lane & 15, kg = lane >> 4;
    const int strip = blockIdx.x % nstrips, band = blockIdx.x / nstrips;
    const int x0 = strip * NS_OW - 6;  // global x of local column 0
    const int y0 = band * band_h, y1 = min(H, y0 + band_h);
    if (y0 >= H) return;
    const int nsteps = (y1 - y0) + 18;

    for (int q = tid; q < NS_LDS / 16; q += 512) reinterpret_cast<uint4 *>(lds)[q] = make_uint4(0, 0, 0, 0);

    // Every stage runs its own loop of nsteps steps with one barrier each, so
    // that the registers of one stage (a layer's weights, the feature
    // stage's prefetched rows) are not live in another's code.
    const int layer = wave < 4 ? wave + 1 : (wave == 4 ? 0 : (wave == 6 ? -1 : 5));

    if (layer == -1) {
        // features: depth range, then the raw samples (color and five depths
        // per pixel) of the rows three steps ahead, in four register sets used
        // in rotation, so that three rows' loads are in flight while a fourth
        // is turned into features
        double lo = kMaxF64, hi = -kMaxF64;
        for (int q = lane; q < nrange; q += 64) {
            const double l = range[2 * q], h = range[2 * q + 1];
            lo = l < lo ? l : lo;
            hi = h > hi ? h : hi;
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double l = __shfl_xor(lo, off), h = __shfl_xor(hi, off);
            lo = l < lo ? l : lo;
            hi = h > hi ? h : hi;
        }
        const double rlo = lo, rhi = hi;
        FeatIn raw0[NS_FEAT_PER_LANE], raw1[NS_FEAT_PER_LANE], raw2[NS_FEAT_PER_LANE], raw3[NS_FEAT_PER_LANE];
        // unconditional (rows past the band load clamped samples nobody
        // uses): a load that may be skipped would merge old and new values of
        // a set, and the copies that merge costs wait for the loads
        auto fetch = [&](int fy, FeatIn (&raw)[NS_FEAT_PER_LANE]) {
#pragma unroll
            for (int i = 0; i < NS_FEAT_PER_LANE; i++) {
                const int gx = x0 + lane + 64 * i;
                const int cy = fy < 0 ? 0 : (fy >= H ? H - 1 : fy), cx = gx < 0 ? 0 : (gx >= W ? W - 1 : gx);
                raw[i] = feat_load(color, depth, H, W, cx, cy);
            }
        };
        // row fy into the feature ring, its colors into the color ring (layer
        // 5 reads them twelve steps later for the residual)
        auto put = [&](int fy, const FeatIn (&raw)[NS_FEAT_PER_LANE]) {
            if (fy >= y1 + 6) return;
            uint8_t *row = feat_ring + (fy & 3) * NS_FEAT_ROW;
            uint32_t *crow = reinterpret_cast<uint32_t *>(color_ring + (fy & 15) * NS_COLOR_ROW);
#pragma unroll
            for (int i = 0; i < NS_FEAT_PER_LANE; i++) {
                const int c = lane + 64 * i;
                if (c >= NS_CW) continue;
                const int gx = x0 + c;
                uint4 f = make_uint4(0, 0, 0, 0);
                if (fy >= 0 && fy < H && gx >= 0 && gx < W) f = feat_eval(raw[i], rlo, rhi);
                *reinterpret_cast<uint4 *>(row + (c + 1) * 16) = f;
                if (EPI == 3) crow[c] = raw[i].c;
            }
        };
        fetch(y0 - 6, raw0);
        fetch(y0 - 5, raw1);
        fetch(y0 - 4, raw2);
        __syncthreads();  // rings zeroed
        // four steps per iteration, each set in a fixed role: step t loads
        // the row of step t + 3 into the set step t - 1 used
        for (int t = 0; t < nsteps; t += 4) {
            const int fy = y0 - 6 + t;
            fetch(fy + 3, raw3);
            put(fy, raw0);
            __syncthreads();
            fetch(fy + 4, raw0);
            put(fy + 1, raw1);
            if (t + 1 < nsteps) __syncthreads();
            fetch(fy + 5, raw1);
            put(fy + 2, raw2);
            if (t + 2 < nsteps) __syncthreads();
            fetch(fy + 6, raw2);
            put(fy + 3, raw3);
            if (t + 3 < nsteps) __syncthreads();
        }
        return;
    }

    if (layer == 0) {
        // 8 used input channels: A[g][t] holds w[4 g + kg][co(pm, t)][0 .. 7]
        bf16x8 A[3][2];
        f32x4 bias[2];
#pragma unroll
        for (int g = 0; g < 3; g++)
#pragma unroll
            for (int t2 = 0; t2 < 2; t2++) {
                const int co = 8 * (pm >> 2) + 4 * t2 + (pm & 3), tap = 4 * g + kg;
                A[g][t2] = tap < 9 ? *reinterpret_cast<const bf16x8 *>(w0 + (tap * 32 + co) * 16)
                                   : bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
            }
#pragma unroll
        for (int t2 = 0; t2 < 2; t2++)
            bias[t2] = f32x4{b0[8 * kg + 4 * t2], b0[8 * kg + 4 * t2 + 1], b0[8 * kg + 4 * t2 + 2],
                             b0[8 * kg + 4 * t2 + 3]};
        // lane (pm, kg) reads tap 4 g + kg of its pixel.  Taps 9..11 (g = 2,
        // kg > 0) have zero weights; they read tap kg - 1 instead (row r - 1,
        // complete before this step), so each product is an exact zero, the
        // sums are those of a zero operand up to the sign of a zero, and the
        // ReLU makes every zero +0.  (Tap kg - 1 is in the same 3x3 window: a
        // non-finite feature there reaches the sum through its own tap too.)
        int trow[3], tcol[3];
#pragma unroll
        for (int g = 0; g < 3; g++) {
            const int tap = 4 * g + kg < 9 ? 4 * g + kg : kg - 1;
            trow[g] = tap / 3;
            tcol[g] = (pm + tap % 3) * 16;
        }
        __syncthreads();  // rings zeroed
        for (int t = 0; t < nsteps; t++) {
            const int r = y0 - 8 + t;
            if (r >= y0 - 5 && r < y1 + 5) {
                uint8_t *dst = act_ring(1) + (r & 3) * NS_ACT_ROW;
                if (r < 0 || r >= H) {
                    for (int q = lane; q < NS_CW * 4; q += 64)
                        *reinterpret_cast<uint4 *>(dst + ns_chunk(1 + (q >> 2), q & 3)) = make_uint4(0, 0, 0, 0);
                } else {
                    const uint8_t *base[3];
#pragma unroll
                    for (int g = 0; g < 3; g++) base[g] = feat_ring + ((r - 1 + trow[g]) & 3) * NS_FEAT_ROW + tcol[g];
                    // block b + 1's three loads go out before block b's MFMAs
                    bf16x8 fa[3], fb[3];
#pragma unroll
                    for (int g = 0; g < 3; g++) fa[g] = *reinterpret_cast<const bf16x8 *>(base[g]);
#pragma unroll
                    for (int b = 0; b < NS_NB; b++) {
                        bf16x8(&cur)[3] = (b & 1) ? fb : fa;
                        bf16x8(&nxt)[3] = (b & 1) ? fa : fb;
                        if (b + 1 < NS_NB)
#pragma unroll
                            for (int g = 0; g < 3; g++)
                                nxt[g] = *reinterpret_cast<const bf16x8 *>(base[g] + 256 * (b + 1));
                        f32x4 acc[2] = {bias[0], bias[1]};
#pragma unroll
                        for (int g = 0; g < 3; g++)
#pragma unroll
                            for (int t2 = 0; t2 < 2; t2++)
                                acc[t2] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[g][t2], cur[g], acc[t2], 0, 0, 0);
                        const int c = 16 * b + pm, gx = x0 + c;
                        uint4 o = make_uint4(0, 0, 0, 0);
                        if (gx >= 0 && gx < W)
                            o = make_uint4(pack_relu_bf2(acc[0][0], acc[0][1]), pack_relu_bf2(acc[0][2], acc[0][3]),
                                           pack_relu_bf2(acc[1][0], acc[1][1]), pack_relu_bf2(acc[1][2], acc[1][3]));
                        *reinterpret_cast<uint4 *>(dst + ns_chunk(c + 1, kg)) = o;
                    }
                }
            }
            __syncthreads();
        }
        return;
    }

    if (layer <= 4) {
        // layers 1-4: lane (pm, kg) holds w[tap][co(pm, t)][8 kg .. 8 kg + 7]
        const uint16_t *wl = layer == 1 ? w1 : layer == 2 ? w2 : layer == 3 ? w3 : w4;
        const float *bl = layer == 1 ? b1 : layer == 2 ? b2 : layer == 3 ? b3 : b4;
        bf16x8 A[9][2];
        f32x4 bias[2];
#pragma unroll
        for (int tap = 0; tap < 9; tap++)
#pragma unroll
            for (int t2 = 0; t2 < 2; t2++) {
                const int co = 8 * (pm >> 2) + 4 * t2 + (pm & 3);
                A[tap][t2] = *reinterpret_cast<const bf16x8 *>(wl + (tap * 32 + co) * 32 + 8 * kg);
            }
#pragma unroll
        for (int t2 = 0; t2 < 2; t2++)
            bias[t2] = f32x4{bl[8 * kg + 4 * t2], bl[8 * kg + 4 * t2 + 1], bl[8 * kg + 4 * t2 + 2],
                             bl[8 * kg + 4 * t2 + 3]};
        const uint8_t *src = act_ring(layer);
        uint8_t *const dring = act_ring(layer + 1);
        __syncthreads();  // rings zeroed
        for (int t = 0; t < nsteps; t++) {
            const int r = y0 - 8 - 2 * layer + t;
            if (r >= y0 - (5 - layer) && r < y1 + (5 - layer)) {
                uint8_t *dst = dring + (r & 3) * NS_ACT_ROW;
                if (r < 0 || r >= H) {
                    for (int q = lane; q < NS_CW * 4; q += 64)
                        *reinterpret_cast<uint4 *>(dst + ns_chunk(1 + (q >> 2), q & 3)) = make_uint4(0, 0, 0, 0);
                } else {
                    const uint8_t *const rows[3] = {src + ((r - 1) & 3) * NS_ACT_ROW, src + (r & 3) * NS_ACT_ROW,
                                                    src + ((r + 1) & 3) * NS_ACT_ROW};
                    // double-buffered fragments: block b + 1's loads go out
                    // between block b's MFMAs
                    bf16x8 fa[9], fb[9];
                    ns_load9(rows, 0, pm, kg, fa);
                    // block 0's nine loads go out back to back; without this
                    // group the interleave below would pair each of them with
                    // the two MFMAs that wait for it
                    __builtin_amdgcn_sched_group_barrier(0x100, 9, 0);
#pragma unroll
                    for (int b = 0; b < NS_NB; b++) {
                        bf16x8(&cur)[9] = (b & 1) ? fb : fa;
                        bf16x8(&nxt)[9] = (b & 1) ? fa : fb;
                        if (b + 1 < NS_NB) ns_load9(rows, b + 1, pm, kg, nxt);
                        f32x4 acc[2] = {bias[0], bias[1]};
#pragma unroll
                        for (int tap = 0; tap < 9; tap++)
#pragma unroll
                            for (int t2 = 0; t2 < 2; t2++)
                                acc[t2] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[tap][t2], cur[tap], acc[t2], 0, 0, 0);
                        if (b + 1 < NS_NB) ns_interleave();
                        const int c = 16 * b + pm, gx = x0 + c;
                        uint4 o = make_uint4(0, 0, 0, 0);
                        if (gx >= 0 && gx < W)
                            o = make_uint4(pack_relu_bf2(acc[0][0], acc[0][1]), pack_relu_bf2(acc[0][2], acc[0][3]),
                                           pack_relu_bf2(acc[1][0], acc[1][1]), pack_relu_bf2(acc[1][2], acc[1][3]));
                        *reinterpret_cast<uint4 *>(dst + ns_chunk(c + 1, kg)) = o;
                    }
                }
            }
            __syncthreads();
        }
        return;
    }

    // layer 5 (waves 5 and 7: blocks blk0 .. blk0 + 2): 16 output channels,
    // lane (pm, kg) holds w[tap][pm][8 kg .. 8 kg + 7]
    const int blk0 = wave == 7 ? NS_NB / 2 : 0;
    bf16x8 A[9];
#pragma unroll
    for (int tap = 0; tap < 9; tap++) A[tap] = *reinterpret_cast<const bf16x8 *>(w5 + (tap * 16 + pm) * 32 + 8 * kg);
    const f32x4 bias = f32x4{b5[4 * kg], b5[4 * kg + 1], b5[4 * kg + 2], b5[4 * kg + 3]};
    const uint8_t *src = act_ring(5);
    __syncthreads();  // rings zeroed
    for (int t = 0; t < nsteps; t++) {
        const int r = y0 - 18 + t;
        if (r >= y0 && r < y1) {
            const uint8_t *const rows[3] = {src + ((r - 1) & 3) * NS_ACT_ROW, src + (r & 3) * NS_ACT_ROW,
                                            src + ((r + 1) & 3) * NS_ACT_ROW};
            const uint32_t *crow = reinterpret_cast<const uint32_t *>(color_ring + (r & 15) * NS_COLOR_ROW);
            bf16x8 fa[9], fb[9];
            ns_load9(rows, blk0, pm, kg, fa);
            __builtin_amdgcn_sched_group_barrier(0x100, 9, 0);  // the first block's loads back to back
#pragma unroll
            for (int k = 0; k < NS_NB / 2; k++) {
                const int b = blk0 + k;
                bf16x8(&cur)[9] = (k & 1) ? fb : fa;
                bf16x8(&nxt)[9] = (k & 1) ? fa : fb;
                if (k + 1 < NS_NB / 2) ns_load9(rows, b + 1, pm, kg, nxt);
                f32x4 acc = bias;
#pragma unroll
                for (int tap = 0; tap < 9; tap++) acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[tap], cur[tap], acc, 0, 0, 0);
                const int c = 16 * b + pm, x = x0 + c;
                if (c < 6 || c >= 6 + NS_OW || x >= W) continue;
                const int64_t px = (int64_t)r * W + x;
                if (EPI == 1) {
                    *reinterpret_cast<float4 *>(static_cast<float *>(out) + px * 16 + 4 * kg) =
                        make_float4(acc[0], acc[1], acc[2], acc[3]);
                } else if (kg == 0) {
                    const uint32_t cc = crow[c];
                    uint32_t o = cc & 0xff000000u;
#pragma unroll
                    for (int ch = 0; ch < 3; ch++) {
                        float v = (float)((cc >> (8 * ch)) & 0xff) * (1.0f / 255.0f) + acc[ch];
                        v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
                        o |= (uint32_t)(v * 255.0f + 0.5f) << (8 * ch);
                    }
                    static_cast<uint32_t *>(out)[px] = o;
                }
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// weight gradient: dw[tap][co][ci] = sum over pixels p of
// x[p + d(tap)][ci] * dy[p][co]; each wave sums NN_DW_ROWS rows of 16
// pixels and writes its own partial (9 * Cout * Cin floats) plus the bias
// gradient partial (Cout floats).

__global__ __launch_bounds__(64) void k_nn_dw(const uint16_t *x, int32_t Cin, const uint16_t *dy, int32_t Cout,
                                              int32_t H, int32_t W, float *partial, float *bpartial) {
    const int lane = threadIdx.x;
    const int x0 = blockIdx.x * 16;
    const int y0 = blockIdx.y * NN_DW_ROWS;
    const int pm = lane & 15, kg = lane >> 4;
    const int64_t wave_id = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
    const int nco = Cout >> 4, nci = Cin >> 4;
    float bsum = 0.0f;  // bias gradient for channel (co tile, pm) accumulated by kg == 0 lanes
    for (int cot = 0; cot < nco; cot++) {
        for (int cit = 0; cit < nci; cit++) {
            f32x4 acc[9];
#pragma unroll
            for (int t = 0; t < 9; t++) acc[t] = f32x4{0, 0, 0, 0};
            for (int r = 0; r < NN_DW_ROWS; r++) {
                const int y = y0 + r;
                if (y >= H) break;
                // A = dy^T: row co = cot*16 + pm, k = pixels x0 + 4*kg + j
                bf16x4 a;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int px = x0 + 4 * kg + j;
                    a[j] = px < W ? (short)dy[((int64_t)y * W + px) * Cout + cot * 16 + pm] : (short)0;
                    if (cit == 0 && px < W) bsum += bf2f((uint16_t)a[j]);
                }
#pragma unroll
                for (int tap = 0; tap < 9; tap++) {
                    const int sy = y + tap / 3 - 1;
                    bf16x4 b;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const int px = x0 + 4 * kg + j;
                        const int sx = px + tap % 3 - 1;
                        const bool ok = px < W && sy >= 0 && sy < H && sx >= 0 && sx < W;
                        b[j] = ok ? (short)x[((int64_t)sy * W + sx) * Cin + cit * 16 + pm] : (short)0;
                    }
                    acc[tap] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[tap], 0, 0, 0);
                }
            }
            // lane holds rows co = cot*16 + 4*kg + j, column ci = cit*16 + pm
            float *pw = partial + wave_id * 9 * Cout * Cin;
#pragma unroll
            for (int tap = 0; tap < 9; tap++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    pw[((int64_t)tap * Cout + cot * 16 + 4 * kg + j) * Cin + cit * 16 + pm] = acc[tap][j];
        }
        // bias: lanes with the same pm hold partial sums over their pixels
        float s = bsum;
        s += __shfl_xor(s, 16);
        s += __shfl_xor(s, 32);
        if (kg == 0) bpartial[wave_id * Cout + cot * 16 + pm] = s;
        bsum = 0.0f;
    }
}

// out[k2][i] = sum over k = k2, k2 + ks, ... < n of part[k * len + i]
// (grid: len / 256 x ks); a second pass with n = ks sums the ks rows
__global__ void k_nn_reduce(const float *part, int64_t n, int64_t len, float *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    const int64_t ks = gridDim.y;
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
    int64_t k = blockIdx.y;
    for (; k + 3 * ks < n; k += 4 * ks) {
        s0 += part[k * len + i];
        s1 += part[(k + ks) * len + i];
        s2 += part[(k + 2 * ks) * len + i];
        s3 += part[(k + 3 * ks) * len + i];
    }
    for (; k < n; k += ks) s0 += part[k * len + i];
    out[(int64_t)blockIdx.y * len + i] = (s0 + s1) + (s2 + s3);
}

// ---------------------------------------------------------------------------
// loss: prediction = input color + out[0..2]; mean squared error against the
// target over the three channels; gradient into gout (bf16, 16 channels).

__global__ __launch_bounds__(256) void k_nn_loss(const float *out, const uint32_t *color, const uint32_t *target,
                                                 int32_t H, int32_t W, uint16_t *gout, float *lpart) {
    __shared__ float ws[4];
    const int64_t n = (int64_t)W * H;
    const float sc = 2.0f / (3.0f * (float)n);
    float l = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t c = color[i], t = target[i];
        uint16_t g[NN_OUT_C];
        for (int k = 0; k < NN_OUT_C; k++) g[k] = 0;
        for (int ch = 0; ch < 3; ch++) {
            const float pred = (float)((c >> (8 * ch)) & 0xff) * (1.0f / 255.0f) + out[i * NN_OUT_C + ch];
            const float tg = (float)((t >> (8 * ch)) & 0xff) * (1.0f / 255.0f);
            const float d = pred - tg;
            l += d * d;
            g[ch] = f2bf(sc * d);
        }
        for (int k = 0; k < NN_OUT_C; k++) gout[i * NN_OUT_C + k] = g[k];
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) lpart[blockIdx.x] = (ws[0] + ws[1] + ws[2] + ws[3]) / (3.0f * (float)n);
}

// ---------------------------------------------------------------------------
// parameters

__global__ void k_nn_adam(float *p, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2,
                          float eps, float c1, float c2) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// f32 weights [9][Cout][Cin] -> bf16 forward copy and the flipped,
// transposed copy [9][Cin][Cout] used for the input gradient
__global__ void k_nn_pack(const float *wf32, int32_t Cin, int32_t Cout, uint16_t *wf, uint16_t *wt) {
    const int64_t n = 9LL * Cout * Cin;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int tap = (int)(i / ((int64_t)Cout * Cin));
        const int co = (int)((i / Cin) % Cout), ci = (int)(i % Cin);
        const uint16_t b = f2bf(wf32[i]);
        wf[i] = b;
        if (wt) wt[((int64_t)(8 - tap) * Cin + ci) * Cout + co] = b;
    }
}

// ---------------------------------------------------------------------------

static unsigned nn_grid(int64_t n) {
    int64_t b = (n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

hipError_t launch_nn_range(const double *partial, int32_t nblocks, double *range, hipStream_t s) {
    hipLaunchKernelGGL(k_nn_range, dim3(1), dim3(64), 0, s, partial, nblocks, range);
    return hipGetLastError();
}
hipError_t launch_nn_features(const uint32_t *color, const double *depth, const double *range, int32_t H, int32_t W,
                              uint16_t *feat, hipStream_t s) {
    hipLaunchKernelGGL(k_nn_features, dim3(nn_grid((int64_t)W * H)), dim3(256), 0, s, color, depth, range, H, W, feat);
    return hipGetLastError();
}
hipError_t launch_nn_conv(int epi, const uint16_t *in, int32_t Cin, const uint16_t *w, const float *bias, int32_t H,
                          int32_t W, int32_t Cout, void *out, const uint16_t *mask, const uint32_t *color,
                          const double *depth, const double *range, hipStream_t s) {
    const int64_t ntiles = (int64_t)((W + CT_W - 1) / CT_W) * ((H + CT_H - 1) / CT_H);
    const int64_t per_xcd = (ntiles + 7) / 8;
    const unsigned grid = 8u * (unsigned)(per_xcd < CT_WG_PER_XCD ? per_xcd : CT_WG_PER_XCD);
    const bool feat = in == nullptr;
#define FGL_NN_CONV(CI, CO, E, F)                                                                               \
    if (Cin == CI && Cout == CO && epi == E && feat == F) {                                                     \
        hipLaunchKernelGGL((k_nn_conv<CI, CO, E, F>), dim3(grid), dim3(256), 0, s, in, w, bias, H, W, out, mask, \
                           color, depth, range);                                                                \
        return hipGetLastError();                                                                               \
    }
    FGL_NN_CONV(16, 32, 0, true)
    FGL_NN_CONV(16, 32, 0, false)
    FGL_NN_CONV(32, 32, 0, false)
    FGL_NN_CONV(32, 16, 1, false)
    FGL_NN_CONV(32, 16, 3, false)
    FGL_NN_CONV(16, 32, 2, false)
    FGL_NN_CONV(32, 32, 2, false)
#undef FGL_NN_CONV
    return hipErrorInvalidValue;
}
hipError_t launch_nn_conv2(int cin_a, int cout_b, int epi_b, const uint16_t *in, const uint16_t *wa, const float *ba,
                           const uint16_t *wb, const float *bb, int32_t H, int32_t W, void *out,
                           const uint32_t *color, const double *depth, const double *range, int32_t nrange,
                           hipStream_t s) {
    const int64_t ntiles = (int64_t)((W + F2_W - 1) / F2_W) * ((H + F2_H - 1) / F2_H);
    const int64_t per_xcd = (ntiles + 7) / 8;
    const unsigned grid = 8u * (unsigned)(per_xcd < F2_WG_PER_XCD ? per_xcd : F2_WG_PER_XCD);
    const bool feat = in == nullptr;
#define FGL_NN_CONV2(CA, CB, E, F)                                                                              \
    if (cin_a == CA && cout_b == CB && epi_b == E && feat == F) {                                               \
        hipLaunchKernelGGL((k_nn_conv2<CA, CB, E, F>), dim3(grid), dim3(512), 0, s, in, wa, ba, wb, bb, H, W,    \
                           out, color, depth, range, nrange);                                                   \
        return hipGetLastError();                                                                               \
    }
    FGL_NN_CONV2(16, 32, 0, true)
    FGL_NN_CONV2(32, 32, 0, false)
    FGL_NN_CONV2(32, 16, 3, false)
    FGL_NN_CONV2(32, 16, 1, false)
#undef FGL_NN_CONV2
    return hipErrorInvalidValue;
}
void nn_stream_grid(int32_t H, int32_t W, int32_t *band_h, int32_t *nstrips, int32_t *nbands) {
    // one workgroup per CU (the rings take most of its LDS): about 256
    // strips x bands, bands at least 16 rows so that the 18-step pipeline
    // fill stays a minor part of each band
    *nstrips = (W + NS_OW - 1) / NS_OW;
    int32_t nb = 256 / *nstrips;
    if (nb < 1) nb = 1;
    int32_t bh = (H + nb - 1) / nb;
    if (bh < 16) bh = 16;
    *band_h = bh;
    *nbands = (H + bh - 1) / bh;
}
hipError_t launch_nn_stream(int epi, const uint32_t *color, const double *depth, const double *range,
                            int32_t nrange, const uint16_t *const *w, const float *const *b, int32_t H, int32_t W,
                            void *out, hipStream_t s) {
    int32_t band_h, nstrips, nbands;
    nn_stream_grid(H, W, &band_h, &nstrips, &nbands);
    const dim3 grid((unsigned)(nstrips * nbands));
    if (epi == 1)
        hipLaunchKernelGGL(k_nn_stream<1>, grid, dim3(512), 0, s, color, depth, range, nrange, w[0], w[1], w[2], w[3],
                           w[4], w[5], b[0], b[1], b[2], b[3], b[4], b[5], H, W, band_h, nstrips, out);
    else if (epi == 3)
        hipLaunchKernelGGL(k_nn_stream<3>, grid, dim3(512), 0, s, color, depth, range, nrange, w[0], w[1], w[2], w[3],
                           w[4], w[5], b[0], b[1], b[2], b[3], b[4], b[5], H, W, band_h, nstrips, out);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
    return hipGetLastError();
}
int64_t nn_dw_waves(int32_t H, int32_t W) { return (int64_t)((W + 15) / 16) * ((H + NN_DW_ROWS - 1) / NN_DW_ROWS); }
hipError_t launch_nn_dw(const uint16_t *x, int32_t Cin, const uint16_t *dy, int32_t Cout, int32_t H, int32_t W,
                        float *partial, float *bpartial, hipStream_t s) {
    dim3 grid((W + 15) / 16, (H + NN_DW_ROWS - 1) / NN_DW_ROWS);
    hipLaunchKernelGGL(k_nn_dw, grid, dim3(64), 0, s, x, Cin, dy, Cout, H, W, partial, bpartial);
    return hipGetLastError();
}
hipError_t launch_nn_reduce(const float *part, int64_t n, int64_t len, float *out, float *tmp, hipStream_t s) {
    const int64_t ks = n < NN_REDUCE_SPLIT ? n : NN_REDUCE_SPLIT;
    const unsigned gx = (unsigned)((len + 255) / 256);
    hipLaunchKernelGGL(k_nn_reduce, dim3(gx, (unsigned)ks), dim3(256), 0, s, part, n, len, tmp);
    hipLaunchKernelGGL(k_nn_reduce, dim3(gx, 1), dim3(256), 0, s, tmp, ks, len, out);
    return hipGetLastError();
}
hipError_t launch_nn_loss(const float *out, const uint32_t *color, const uint32_t *target, int32_t H, int32_t W,
                          uint16_t *gout, float *lpart, int32_t nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_nn_loss, dim3(nblocks), dim3(256), 0, s, out, color, target, H, W, gout, lpart);
    return hipGetLastError();
}
hipError_t launch_nn_adam(float *p, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2,
                          float eps, float c1, float c2, hipStream_t s) {
    hipLaunchKernelGGL(k_nn_adam, dim3(nn_grid(n)), dim3(256), 0, s, p, g, m, v, n, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}
hipError_t launch_nn_pack(const float *wf32, int32_t Cin, int32_t Cout, uint16_t *wf, uint16_t *wt, hipStream_t s) {
    hipLaunchKernelGGL(k_nn_pack, dim3(nn_grid(9LL * Cout * Cin)), dim3(256), 0, s, wf32, Cin, Cout, wf, wt);
    return hipGetLastError();
}

}  // namespace fgl
