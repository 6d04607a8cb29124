// fgl_nn.hip -- learned anti-aliasing: a small convolutional network maps a
// 1x render (color, depth and depth-derived normals) to an estimate of the
// 4x4-supersampled image.  Six 3x3 convolutions (features -> 32 -> 32 -> 32
// -> 32 -> 32 -> 3, ReLU between), the output added to the input color.
// No reference counterpart: this replaces 16x SSAA + resolve after the draw.
//
// Tensors are pixel-major with channels contiguous (H x W x C), BF16 for
// convolution operands, FP32 accumulation (MFMA 16x16x16 bf16).  A wave
// computes 16 consecutive pixels of one row times every output channel:
// the A operand is 16 pixels x 16 input channels of one tap, the B operand
// 16 input x 16 output channels of that tap's weights.  Training runs the
// backward pass with the same convolution kernel (input gradient: flipped,
// transposed weights, ReLU mask in the epilogue), a weight-gradient kernel
// (MFMA over pixels) with per-wave partial sums reduced without atomics,
// and Adam on FP32 master weights.
#include "fgl_internal.h"

namespace fgl {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round-to-nearest-even FP32 -> BF16 (finite inputs)
FGL_D uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
FGL_D float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

constexpr double kMaxF64 = 1.79769313486231570814527423731704356798070e+308;

// ---------------------------------------------------------------------------
// input features: R, G, B (NRGBA / 255), depth t in [0, 1] (DepthImage's
// normalisation; 1 where cleared), the normal of the depth surface from
// central differences (nx, ny, nz), coverage mask; channels 8..15 zero.

__global__ void k_nn_range(const double *partial, int32_t nblocks, double *range) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double lo = kMaxF64, hi = -kMaxF64;
    for (int32_t b = 0; b < nblocks; b++) {
        const double l = partial[2 * b], h = partial[2 * b + 1];
        if (l < lo) lo = l;
        if (h > hi) hi = h;
    }
    range[0] = lo;
    range[1] = hi;
}

FGL_D float depth_t(const double *depth, int64_t i, double lo, double hi) {
    const double d = depth[i];
    if (d == kMaxF64 || !(hi > lo)) return 1.0f;
    return (float)((d - lo) / (hi - lo));
}

__global__ __launch_bounds__(256) void k_nn_features(const uint32_t *color, const double *depth, const double *range,
                                                     int32_t H, int32_t W, uint16_t *feat) {
    const int64_t n = (int64_t)W * H;
    const double lo = range[0], hi = range[1];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t x = (int32_t)(i % W), y = (int32_t)(i / W);
        const uint32_t c = color[i];
        const float t = depth_t(depth, i, lo, hi);
        const int32_t xl = x > 0 ? x - 1 : x, xr = x + 1 < W ? x + 1 : x;
        const int32_t yu = y > 0 ? y - 1 : y, yd = y + 1 < H ? y + 1 : y;
        const float gx = depth_t(depth, (int64_t)y * W + xr, lo, hi) - depth_t(depth, (int64_t)y * W + xl, lo, hi);
        const float gy = depth_t(depth, (int64_t)yd * W + x, lo, hi) - depth_t(depth, (int64_t)yu * W + x, lo, hi);
        const float sx = -64.0f * gx, sy = -64.0f * gy;
        const float rn = 1.0f / sqrtf(sx * sx + sy * sy + 1.0f);
        float f[8];
        f[0] = (float)(c & 0xff) * (1.0f / 255.0f);
        f[1] = (float)((c >> 8) & 0xff) * (1.0f / 255.0f);
        f[2] = (float)((c >> 16) & 0xff) * (1.0f / 255.0f);
        f[3] = t;
        f[4] = sx * rn;
        f[5] = sy * rn;
        f[6] = rn;
        f[7] = depth[i] == kMaxF64 ? 0.0f : 1.0f;
        uint4 lo4, hi4 = make_uint4(0, 0, 0, 0);
        lo4.x = f2bf(f[0]) | ((uint32_t)f2bf(f[1]) << 16);
        lo4.y = f2bf(f[2]) | ((uint32_t)f2bf(f[3]) << 16);
        lo4.z = f2bf(f[4]) | ((uint32_t)f2bf(f[5]) << 16);
        lo4.w = f2bf(f[6]) | ((uint32_t)f2bf(f[7]) << 16);
        uint4 *o = reinterpret_cast<uint4 *>(feat + i * NN_FEAT_C);
        o[0] = lo4;
        o[1] = hi4;
    }
}

// ---------------------------------------------------------------------------
// 3x3 convolution, zero padding.  in: H x W x Cin (bf16), w: 9 x Cout x Cin
// (bf16, tap = 3 * (dy + 1) + (dx + 1)), Cin and Cout multiples of 16,
// Cout <= NN_C.  EPI 0: + bias, ReLU -> bf16; 1: + bias -> f32;
// 2: times (mask > 0) -> bf16 (the ReLU derivative of the layer below).

template <int EPI>
__global__ __launch_bounds__(256) void k_nn_conv(const uint16_t *in, int32_t Cin, const uint16_t *w,
                                                 const float *bias, int32_t H, int32_t W, int32_t Cout, void *out,
                                                 const uint16_t *mask) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int y = blockIdx.y * 4 + wave;
    const int x0 = blockIdx.x * 16;
    if (y >= H) return;
    const int pm = lane & 15, kg = lane >> 4;
    const int nt = Cout >> 4;
    f32x4 acc[NN_C / 16];
#pragma unroll
    for (int t = 0; t < NN_C / 16; t++) acc[t] = f32x4{0, 0, 0, 0};
    for (int tap = 0; tap < 9; tap++) {
        const int sy = y + tap / 3 - 1, sx = x0 + pm + tap % 3 - 1;
        const bool inb = sy >= 0 && sy < H && sx >= 0 && sx < W;
        const uint16_t *ap = in + ((int64_t)(inb ? sy : 0) * W + (inb ? sx : 0)) * Cin + 4 * kg;
        const uint16_t *bp = w + ((int64_t)tap * Cout + pm) * Cin + 4 * kg;
        for (int c0 = 0; c0 < Cin; c0 += 16) {
            uint2 av = inb ? *reinterpret_cast<const uint2 *>(ap + c0) : make_uint2(0, 0);
            bf16x4 a;
            a[0] = (short)(av.x & 0xffff);
            a[1] = (short)(av.x >> 16);
            a[2] = (short)(av.y & 0xffff);
            a[3] = (short)(av.y >> 16);
#pragma unroll
            for (int t = 0; t < NN_C / 16; t++) {
                if (t < nt) {
                    const uint2 bv = *reinterpret_cast<const uint2 *>(bp + (int64_t)t * 16 * Cin + c0);
                    bf16x4 b;
                    b[0] = (short)(bv.x & 0xffff);
                    b[1] = (short)(bv.x >> 16);
                    b[2] = (short)(bv.y & 0xffff);
                    b[3] = (short)(bv.y >> 16);
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
                }
            }
        }
    }
    // lane holds rows (pixels) 4 * kg + j, column (channel) t * 16 + pm
#pragma unroll
    for (int t = 0; t < NN_C / 16; t++) {
        if (t >= nt) continue;
        const int ch = t * 16 + pm;
        const float bb = (EPI == 2) ? 0.0f : bias[ch];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int x = x0 + 4 * kg + j;
            if (x >= W) continue;
            const int64_t o = ((int64_t)y * W + x) * Cout + ch;
            float v = acc[t][j] + bb;
            if (EPI == 0) {
                v = v > 0.0f ? v : 0.0f;
                static_cast<uint16_t *>(out)[o] = f2bf(v);
            } else if (EPI == 1) {
                static_cast<float *>(out)[o] = v;
            } else {
                const float m = bf2f(mask[o]);
                static_cast<uint16_t *>(out)[o] = f2bf(m > 0.0f ? v : 0.0f);
            }
        }
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

// out[i] = sum over k < n of part[k * len + i]
__global__ void k_nn_reduce(const float *part, int64_t n, int64_t len, float *out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x) {
        float s = 0.0f;
        for (int64_t k = 0; k < n; k++) s += part[k * len + i];
        out[i] = s;
    }
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

// prediction -> NRGBA8 (alpha from the input)
__global__ void k_nn_output(const float *out, const uint32_t *color, int32_t H, int32_t W, uint32_t *dst) {
    const int64_t n = (int64_t)W * H;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t c = color[i];
        uint32_t o = c & 0xff000000u;
        for (int ch = 0; ch < 3; ch++) {
            float v = (float)((c >> (8 * ch)) & 0xff) * (1.0f / 255.0f) + out[i * NN_OUT_C + ch];
            v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
            o |= (uint32_t)(v * 255.0f + 0.5f) << (8 * ch);
        }
        dst[i] = o;
    }
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
                          int32_t W, int32_t Cout, void *out, const uint16_t *mask, hipStream_t s) {
    dim3 grid((W + 15) / 16, (H + 3) / 4), block(256);
    if (epi == 0) hipLaunchKernelGGL(k_nn_conv<0>, grid, block, 0, s, in, Cin, w, bias, H, W, Cout, out, mask);
    else if (epi == 1) hipLaunchKernelGGL(k_nn_conv<1>, grid, block, 0, s, in, Cin, w, bias, H, W, Cout, out, mask);
    else hipLaunchKernelGGL(k_nn_conv<2>, grid, block, 0, s, in, Cin, w, bias, H, W, Cout, out, mask);
    return hipGetLastError();
}
int64_t nn_dw_waves(int32_t H, int32_t W) { return (int64_t)((W + 15) / 16) * ((H + NN_DW_ROWS - 1) / NN_DW_ROWS); }
hipError_t launch_nn_dw(const uint16_t *x, int32_t Cin, const uint16_t *dy, int32_t Cout, int32_t H, int32_t W,
                        float *partial, float *bpartial, hipStream_t s) {
    dim3 grid((W + 15) / 16, (H + NN_DW_ROWS - 1) / NN_DW_ROWS);
    hipLaunchKernelGGL(k_nn_dw, grid, dim3(64), 0, s, x, Cin, dy, Cout, H, W, partial, bpartial);
    return hipGetLastError();
}
hipError_t launch_nn_reduce(const float *part, int64_t n, int64_t len, float *out, hipStream_t s) {
    hipLaunchKernelGGL(k_nn_reduce, dim3(nn_grid(len)), dim3(256), 0, s, part, n, len, out);
    return hipGetLastError();
}
hipError_t launch_nn_loss(const float *out, const uint32_t *color, const uint32_t *target, int32_t H, int32_t W,
                          uint16_t *gout, float *lpart, int32_t nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_nn_loss, dim3(nblocks), dim3(256), 0, s, out, color, target, H, W, gout, lpart);
    return hipGetLastError();
}
hipError_t launch_nn_output(const float *out, const uint32_t *color, int32_t H, int32_t W, uint32_t *dst,
                            hipStream_t s) {
    hipLaunchKernelGGL(k_nn_output, dim3(nn_grid((int64_t)W * H)), dim3(256), 0, s, out, color, H, W, dst);
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
