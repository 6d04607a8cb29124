// fgl_internal.h -- kernel argument blocks and launch entry points shared by
// the HIP translation units and the host runtime (fgl_host.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>
#include <vector>

#include "fgl_device.h"

namespace fgl {

enum PrimKind : int32_t { PRIM_TRI = 0, PRIM_LINE = 1 };

// Input soup on the device (SoA, see fgl_prims in include/fgl.h).
struct DevPrims {
    int64_t n;
    const double *pos, *nrm, *uv, *col;
};

// Everything a draw's kernels share.
struct DrawArgs {
    int32_t W, H;          // framebuffer size
    int32_t ntx, nty;      // tiles per row / column
    int32_t ntiles;        // ntx * nty
    int32_t kind;          // PrimKind
    DevPrims in;
    const double *zeros;   // 16 zero doubles standing in for a missing attribute array
    DevState st;
    DevShader sh;
    double S[16];          // Screen(W, H) (matrix.go:119-128)

    // vertex stage outputs
    PrimRec *rec_main;     // in.n records (triangles drawn without clipping)
    PrimRec *side_rec;     // raster triangles from clipping, wireframe and lines
    uint32_t side_rec_cap;
    double *side_vtx;      // clipped vertices, SIDE_VTX_DOUBLES each
    uint32_t side_vtx_cap;
    uint32_t *ntl;         // per input primitive: (tile, triangle) pairs it emits
    uint32_t *nrast;       // per input primitive: raster triangles (bit 31: side records)
    uint32_t *sstart;      // per input primitive: first side record
    int4 *tbox;            // per input triangle: its main record's box (x0, y0, x1, y1), for the pair emitter
    uint64_t *blocksum;    // per vertex-stage block: sum of ntl
    uint64_t *blockoff;    // exclusive prefix of blocksum
    int32_t nvsblocks;
    uint32_t *deferred;    // input triangles that need clipping (in.n slots)

    // binning
    uint32_t *keys[2], *vals[2];  // ping-pong (tile id, triangle ref) pairs
    uint32_t pair_cap;
    uint32_t *hist;        // radix histograms, 256 * nrsblocks
    uint32_t *dtot;        // per-digit totals of the current radix pass
    int32_t nrsblocks;
    int32_t key_bits;
    uint32_t *tstart, *tend;  // per tile range in the sorted pairs
    int32_t sorted_buf;    // which of keys/vals holds the sorted result

    // coverage spans
    uint32_t *pspans;           // spans per sorted pair
    uint16_t *prows;            // per sorted pair: its rows that have spans (ROWS_ALL: every row)
    uint64_t *span_blocksum;    // per SPAN_THREADS pairs
    uint64_t *span_blockoff;
    uint32_t *poff;             // first span of each sorted pair (pair_cap + 1 entries)
    Span *spans;
    uint32_t span_cap;

    // shaded fragments
    uint64_t *frag_blocksum;    // fragments per SPAN_THREADS spans (accumulated by k_spans)
    uint64_t *frag_blockoff;    // offset of each span block within its group of 1024 blocks
    uint64_t *frag_gsum;        // fragments per group of 1024 span blocks
    uint32_t *foff;             // first fragment of each span (span_cap + 1 entries)
    Frag *frags;
    uint32_t frag_cap;

    // framebuffer and deferred clears
    uint32_t *color;       // NRGBA8 as little-endian u32
    double *depth;
    int32_t clear_color_pending, clear_depth_pending;
    uint32_t clear_color;
    double clear_depth;
    // per CB_W x CB_H pixel block: CB_COLOR / CB_DEPTH set while the block's
    // color / depth is the context's clear value and has not been written
    uint8_t *cblk;
    int32_t cbw;           // blocks per row

    DrawCounters *ctr;
};

constexpr int VS_THREADS = 256;
constexpr int SPAN_THREADS = 256;
// Spans per fragment-count block: the shading kernel's waves each take one
// such block (one span per lane) and need only its fragment offset.
constexpr int FRAG_BLOCK = 64;
constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_BLOCK = RS_THREADS * RS_ITEMS;

// The launches of one draw, in order (the three radix stages once per pass).
enum DrawStage : int {
    ST_ZERO,
    ST_VS,
    ST_VS_DEFERRED,
    ST_SCAN_TOP,
    ST_EMIT,
    ST_RADIX_HIST,
    ST_RADIX_SCAN,
    ST_RADIX_SCATTER,
    ST_SPAN_COUNT,
    ST_SPAN_SCAN,
    ST_SPANS,
    ST_FRAG_SCAN,
    ST_SHADE,
    ST_ROP,
    ST_COUNT
};

// One launch of a stage for many draws (the views of a batch), each with its
// own buffers: launch block b runs block b - first[v] of draw v = view_of[b]
// (first: one entry per draw, then the total).  The draws of one launch use
// the same tile shape and kernel variant (primitive kind, wireframe, shader
// kind and texture); `pass` is the radix pass of the radix stages.
struct MultiArgs {
    const DrawArgs *draws;
    const uint32_t *view_of;
    const uint32_t *first;
    int32_t pass;
};

// Kernel launchers (stream-ordered).  Each returns the hipError_t of the launch.

// The draw pipeline's launchers, one set per tile shape (see FGL_TILE_NS).
namespace FGL_TILE_NS {
hipError_t launch_zero(const DrawArgs &a, hipStream_t s);
hipError_t launch_vs(const DrawArgs &a, hipStream_t s);
hipError_t launch_vs_deferred(const DrawArgs &a, hipStream_t s);
hipError_t launch_scan_top(const DrawArgs &a, hipStream_t s);
hipError_t launch_emit(const DrawArgs &a, hipStream_t s);
hipError_t launch_radix_hist(const DrawArgs &a, int pass, hipStream_t s);
hipError_t launch_radix_scan(const DrawArgs &a, hipStream_t s);
hipError_t launch_radix_scatter(const DrawArgs &a, int pass, hipStream_t s);
hipError_t launch_span_count(const DrawArgs &a, hipStream_t s);
hipError_t launch_span_scan(const DrawArgs &a, hipStream_t s);
hipError_t launch_spans(const DrawArgs &a, hipStream_t s);
hipError_t launch_frag_scan(const DrawArgs &a, hipStream_t s);
hipError_t launch_shade(const DrawArgs &a, hipStream_t s);
hipError_t launch_tile(const DrawArgs &a, hipStream_t s);
// workgroups of one stage's launch for draw a (radix stages: per pass);
// `many`: the launch covers many draws
uint32_t stage_blocks(int stage, const DrawArgs &a, bool many);
// one launch of a stage over nblocks workgroups of many draws; rep is one
// of the draws (it selects the kernel variant)
hipError_t launch_multi(int stage, const MultiArgs &m, uint32_t nblocks, const DrawArgs &rep, hipStream_t s);
// the pieces the two above are made of, one per stage
uint32_t blocks_zero(const DrawArgs &a);
uint32_t blocks_vs_deferred(const DrawArgs &a, bool many);
uint32_t span_blocks(const DrawArgs &a);
using MultiLaunch = hipError_t (*)(const MultiArgs &, uint32_t, const DrawArgs &, hipStream_t);
hipError_t launch_zero_multi(const MultiArgs &m, uint32_t nb, const DrawArgs &rep, hipStream_t s);
hipError_t launch_vs_multi(const MultiArgs &m, uint32_t nb, const DrawArgs &rep, hipStream_t s);
hipError_t launch_vs_deferred_multi(const MultiArgs &m, uint32_t nb, const DrawArgs &rep, hipStream_t s);
hipError_t launch_scan_top_multi(const MultiArgs &m, uint32_t nb, const DrawArgs &rep, hipStream_t s);
hipError_t launch_emit_multi(const MultiArgs &m, uint32_t nb, const DrawArgs &rep, hipStream_t s);
hipError_t launch_radix_hist_multi(const MultiArgs &m, uint32_t nb, const DrawArgs &rep, hipStream_t s);
hipError_t launch_radix_scan_multi(const MultiArgs &m, uint32_t nb, const DrawArgs &rep, hipStream_t s);
hipError_t launch_radix_scatter_multi(const MultiArgs &m, uint32_t nb, const DrawArgs &rep, hipStream_t s);
hipError_t launch_span_count_multi(const MultiArgs &m, uint32_t nb, const DrawArgs &rep, hipStream_t s);
hipError_t launch_span_scan_multi(const MultiArgs &m, uint32_t nb, const DrawArgs &rep, hipStream_t s);
hipError_t launch_spans_multi(const MultiArgs &m, uint32_t nb, const DrawArgs &rep, hipStream_t s);
}  // namespace FGL_TILE_NS

using DrawLaunch = hipError_t (*)(const DrawArgs &, hipStream_t);
using PassLaunch = hipError_t (*)(const DrawArgs &, int, hipStream_t);

// One tile shape's launchers, as the host calls them.
struct TileOps {
    int32_t tw, th;
    DrawLaunch zero, vs, vs_deferred, scan_top, emit;
    PassLaunch radix_hist;
    DrawLaunch radix_scan;
    PassLaunch radix_scatter;
    DrawLaunch span_count, span_scan, spans, frag_scan, shade, rop;
    uint32_t (*stage_blocks)(int stage, const DrawArgs &a, bool many);
    hipError_t (*multi)(int stage, const MultiArgs &m, uint32_t nblocks, const DrawArgs &rep, hipStream_t s);
};

// The compiled shapes (width, height); 32 x 32 is the default.
#define FGL_TILE_SHAPES(X) X(32, 32) X(64, 32) X(32, 16) X(64, 64) X(128, 64)
#define FGL_DECLARE_TILE_OPS(w, h) const TileOps &tile_ops_##w##x##h();
FGL_TILE_SHAPES(FGL_DECLARE_TILE_OPS)
#undef FGL_DECLARE_TILE_OPS

hipError_t launch_copy_words(const uint32_t *src, uint32_t *dst, int32_t n, hipStream_t s);
// count records of n words: record k from src[k] to dst + slot[k] * n
hipError_t launch_gather_words(const uint32_t *const *src, const uint32_t *slot, uint32_t *dst, int32_t n,
                               int32_t count, hipStream_t s);
// Deferred clears per block.  A draw whose clear is still pending leaves the
// blocks of the tiles no primitive reaches unwritten and marks them here; the
// ROP starts a later tile's marked blocks from the clear value; a reader of
// the whole buffer first fills the marked blocks (launch_cblk_fill).  Every
// tile shape is a whole number of blocks.
constexpr int CB_W = 32, CB_H = 16;
constexpr uint8_t CB_COLOR = 1, CB_DEPTH = 2;
// fills the marked blocks' channels in `chans` with the clear values and
// unmarks them
hipError_t launch_cblk_fill(uint32_t *color, double *depth, int32_t W, int32_t H, uint8_t *cblk, int32_t cbw,
                            int32_t cbh, uint32_t clear_color, double clear_depth, uint8_t chans, hipStream_t s);
// unmarks `bits` in every block (the channel was overwritten as a whole)
hipError_t launch_cblk_drop(uint8_t *cblk, int64_t n, uint8_t bits, hipStream_t s);
hipError_t launch_fill_color(uint32_t *color, int64_t n, uint32_t v, hipStream_t s);
hipError_t launch_fill_depth(double *depth, int64_t n, double v, hipStream_t s);
// The 4x tent when every output row of both weight tables is the same eight
// 8-bit weights with a power-of-two sum and offsets 4*i - 2 (the host checks
// the tables): weights packed four to a dword for v_dot4_u32_u8.  wx holds
// twelve taps from a 16-byte-aligned staged column (two zero weights each
// side), wy eight.
struct ResolveTent4 {
    int32_t ok;
    uint32_t wx[3], wy[2];
    int32_t shx, shy;  // log2 of each table's weight sum
};

// The 4x tent path reads the source's deferred-clear marks (cblk, or null
// when there are none) and takes the clear color for marked blocks; the other
// paths need the source's color materialised.
bool resolve_reads_marks(int32_t filter, int32_t sw, int32_t sh, int32_t dw, int32_t dh, const ResolveTent4 &t4);
hipError_t launch_resolve(const uint32_t *src, int32_t sw, int32_t sh, uint32_t *tmp, uint32_t *dst, int32_t dw,
                          int32_t dh, int32_t filter, const int16_t *wx, const int32_t *ox, int32_t flx,
                          const int16_t *wy, const int32_t *oy, int32_t fly, const ResolveTent4 &t4,
                          const uint8_t *cblk, int32_t cbw, uint32_t clear_color, hipStream_t s);
hipError_t launch_depth_minmax(const double *depth, int64_t n, double *partial, int32_t nblocks, hipStream_t s);
hipError_t launch_depth_image(const double *depth, int64_t n, const double *partial, int32_t nblocks,
                              uint16_t *out, hipStream_t s);
// the allocation-free DepthImage's per-device scratch (fgl_misc.hip): room
// for kDepthScratchBlocks blocks of (min, max); *ptr = this device's copy
constexpr int32_t kDepthScratchBlocks = 512;
constexpr int32_t kMaxDevices = 64;
hipError_t depth_scratch(int32_t device, double **ptr);

// learned anti-aliasing (fgl_nn.hip)
constexpr int NN_C = 32;         // hidden channels (widest layer)
constexpr int NN_FEAT_C = 16;    // input feature channels (8 used)
constexpr int NN_OUT_C = 16;     // output channels (3 used)
constexpr int NN_DW_ROWS = 32;   // rows of 16 pixels per weight-gradient wave
constexpr int NN_REDUCE_SPLIT = 64;  // first-pass rows of the partial-sum reduction
constexpr int NN_LAYERS = 6;

hipError_t launch_nn_range(const double *partial, int32_t nblocks, double *range, hipStream_t s);
hipError_t launch_nn_features(const uint32_t *color, const double *depth, const double *range, int32_t H, int32_t W,
                              uint16_t *feat, hipStream_t s);
// in == nullptr: layer 0 with the features computed from color, depth and
// range (Cin 16)
hipError_t launch_nn_conv(int epi, const uint16_t *in, int32_t Cin, const uint16_t *w, const float *bias, int32_t H,
                          int32_t W, int32_t Cout, void *out, const uint16_t *mask, const uint32_t *color,
                          const double *depth, const double *range, hipStream_t s);
// mesh ingest (fgl_ingest.hip)
hipError_t launch_stl_decode(const uint8_t *file, int64_t count, double *pos, double *nrm, hipStream_t s);
hipError_t launch_fix_normals(const double *pos, double *nrm, int64_t ntri, hipStream_t s);
double ingest_parse_float(const char *s, size_t n);
int64_t ingest_parse_int(const char *s, size_t n);
void ingest_fields(const char *a, const char *b, std::vector<std::pair<const char *, size_t>> &out);

// PNG encode (fgl_png.hip)
hipError_t launch_png_opaque(const uint32_t *color, int64_t n, uint32_t *flag, hipStream_t s);
hipError_t launch_png_filter(const uint32_t *color, int32_t W, int32_t H, int32_t bpp, uint8_t *out, hipStream_t s);
hipError_t launch_png_deflate(const uint8_t *stream, int64_t n, int32_t chunk, int32_t cap_words, int64_t npieces,
                              uint32_t *out, uint32_t *sizes, uint32_t *adl, hipStream_t s);
hipError_t launch_png_pack(const uint32_t *pieces, const uint32_t *sizes, const uint64_t *offs, int32_t cap_words,
                           int64_t npieces, uint8_t *dst, hipStream_t s);

// mesh preparation (fgl_mesh.hip)
hipError_t launch_mesh_first_nonzero(const double *pos, int64_t nprim, int32_t nv, int64_t *block_first,
                                     int32_t nblocks, hipStream_t s);
hipError_t launch_mesh_box(const double *pos, int64_t nprim, int32_t nv, int64_t from, double *block_box,
                           int32_t nblocks, hipStream_t s);
hipError_t launch_mesh_transform(double *pos, double *nrm, int64_t nvert, const double m[16], hipStream_t s);
int32_t mesh_sort_blocks(int64_t n);
hipError_t launch_mesh_group(const double *pos, int64_t nvert, uint32_t *keys[2], uint32_t *vals[2], uint32_t *hist,
                             uint32_t *dtot, uint32_t *where, hipStream_t s);
hipError_t launch_mesh_smooth(bool threshold_mode, const double *pos, const double *nrm, int64_t nvert,
                              const uint32_t *skeys, const uint32_t *svals, const uint32_t *where, double threshold,
                              double *out, hipStream_t s);

// two fused layers (inference): cin_a -> 32 (ReLU) -> cout_b with epilogue
// epi_b (0: ReLU -> bf16, 1: f32, 3: final color); in == nullptr: layer A
// reads the features computed from color, depth and range ({lo, hi}, or
// nrange > 0 per-block {lo, hi} partials)
// all six layers in one launch (inference): epi 1 raw output (H x W x 16
// floats), 3 the final color; w / b the six layers' bf16 weights and biases
hipError_t launch_nn_stream(int epi, const uint32_t *color, const double *depth, const double *range,
                            int32_t nrange, const uint16_t *const *w, const float *const *b, int32_t H, int32_t W,
                            void *out, hipStream_t s);
hipError_t launch_nn_conv2(int cin_a, int cout_b, int epi_b, const uint16_t *in, const uint16_t *wa, const float *ba,
                           const uint16_t *wb, const float *bb, int32_t H, int32_t W, void *out,
                           const uint32_t *color, const double *depth, const double *range, int32_t nrange,
                           hipStream_t s);
int64_t nn_dw_waves(int32_t H, int32_t W);
hipError_t launch_nn_dw(const uint16_t *x, int32_t Cin, const uint16_t *dy, int32_t Cout, int32_t H, int32_t W,
                        float *partial, float *bpartial, hipStream_t s);
hipError_t launch_nn_reduce(const float *part, int64_t n, int64_t len, float *out, float *tmp, hipStream_t s);
hipError_t launch_nn_loss(const float *out, const uint32_t *color, const uint32_t *target, int32_t H, int32_t W,
                          uint16_t *gout, float *lpart, int32_t nblocks, hipStream_t s);
hipError_t launch_nn_adam(float *p, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2,
                          float eps, float c1, float c2, hipStream_t s);
hipError_t launch_nn_pack(const float *wf32, int32_t Cin, int32_t Cout, uint16_t *wf, uint16_t *wt, hipStream_t s);

}  // namespace fgl
