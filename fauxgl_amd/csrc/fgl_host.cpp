// fgl_host.cpp -- the C ABI (include/fgl.h): contexts, device buffers,
// deferred clears, the per-draw kernel sequence and its capacity retries.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/fgl.h"
#include "fgl_internal.h"

using namespace fgl;

namespace {

thread_local std::string g_err = "";
thread_local int g_device = -1;

void set_err(std::string &dst, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    dst = buf;
}

// Advances whenever the library frees or replaces a device allocation
// (buffers, textures).  A recorded graph bakes device addresses into its
// kernel nodes; a batch replays its graph only while this is unchanged.
std::atomic<uint64_t> g_dev_epoch{0};

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    ~DevBuf() {
        if (p) {
            (void)hipFree(p);
            g_dev_epoch++;
        }
    }
    hipError_t ensure(size_t need) {
        if (need <= bytes && p) return hipSuccess;
        if (p) {
            (void)hipFree(p);
            g_dev_epoch++;
            p = nullptr;
            bytes = 0;
        }
        size_t want = need < 256 ? 256 : need;
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        bytes = want;
        return hipSuccess;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Page-locked host memory (hipHostMalloc), so a device-to-host copy into it
// needs no staging.
struct HostBuf {
    void *p = nullptr;
    size_t bytes = 0;
    ~HostBuf() {
        if (p) (void)hipHostFree(p);
    }
    hipError_t ensure(size_t need) {
        if (need <= bytes && p) return hipSuccess;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = hipHostMalloc(&p, need, hipHostMallocDefault);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        bytes = need;
        return hipSuccess;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Waits for a stream by polling it.  A blocking wait spins only briefly and
// then sleeps until an interrupt, which wakes the host tens of microseconds
// after the GPU has finished: a large share of a draw that takes a few
// hundred.  Polling gives up after kPollMs and blocks from then on, so a
// long wait does not hold a core.
constexpr int kPollMs = 20;
hipError_t wait_stream(hipStream_t s) {
    const auto t0 = std::chrono::steady_clock::now();
    hipError_t q;
    for (;;) {
        q = hipStreamQuery(s);
        if (q != hipErrorNotReady) break;
        if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(kPollMs)) {
            q = hipStreamSynchronize(s);
            break;
        }
    }
    (void)hipGetLastError();  // a not-ready query is no error of a later call
    return q;
}

struct Tex {
    void *d = nullptr;
    int32_t w = 0, h = 0, kind = 0;
};

struct KStat {
    uint64_t launches = 0;
    double ms = 0;
};

struct EvPair {
    std::string name;
    hipEvent_t a, b;
};

}  // namespace

struct fgl_ctx {
    int device = 0;
    int32_t W = 0, H = 0;
    hipStream_t stream = nullptr;
    DevBuf color, depth;
    bool pend_color = false, pend_depth = false;
    uint32_t clear_color = 0;
    double clear_depth = 0;
    // blocks whose pending clear a draw left unwritten (CB_COLOR / CB_DEPTH
    // per CB_W x CB_H block); cblk_any: some block may be marked
    DevBuf cblk;
    int32_t cbw = 0, cbh = 0;
    bool cblk_any = false;
    std::string err;
    std::map<int64_t, Tex> tex;
    int64_t next_tex = 1;

    DevBuf ctr, ntl, nrast, sstart, blocksum, blockoff, rec_main, tbox, side_rec, side_vtx;
    HostBuf hctr;  // the draw counters as the host reads them back
    DevBuf keys0, vals0, keys1, vals1, hist, dtot, tstart, tend, deferred;
    DevBuf pspans, prows, span_bsum, span_boff, poff, spans;
    DevBuf frag_bsum, frag_boff, frag_gsum, foff, frags;
    uint32_t span_cap = 0, frag_cap = 0;
    DevBuf in_pos, in_nrm, in_uv, in_col;
    DevBuf zeros;
    DevBuf rs_tmp, rs_wx, rs_ox, rs_wy, rs_oy;
    int32_t rs_key[4] = {0, 0, 0, 0};  // source and destination sizes the tent tables are for
    ResolveTent4 rs_t4{};              // the tables as packed uniform 4x weights, when they are
    hipEvent_t xev = nullptr;          // orders another context's stream after this one
    DevBuf dpart, dimg;
    uint32_t side_rec_cap = 0, side_vtx_cap = 0, pair_cap = 0;
    DrawCounters last{};

    int32_t tile_w = 0, tile_h = 0;  // tile shape set for this context (0: from the tuning table)

    // Failure injection: capacities (side records, side vertices, pairs,
    // spans, fragments; 0 = as computed) that the next draw's first pass
    // starts with.  Its kernels then overflow, set their bits and exit,
    // and the draw grows the buffers and runs again, as for a real overflow.
    uint32_t force_caps[5] = {0, 0, 0, 0, 0};
    uint32_t last_overflow = 0;  // overflow bits seen by the last draw's passes
    uint32_t last_reruns = 0;    // passes the last draw ran again after an overflow

    bool prof = false;
    std::string prof_only;  // time only launches of this kernel (empty: all)
    std::map<std::string, KStat> kstats;
    std::vector<EvPair> pending;
    std::vector<hipEvent_t> pool;

    ~fgl_ctx() {
        for (auto &kv : tex)
            if (kv.second.d) {
                (void)hipFree(kv.second.d);
                g_dev_epoch++;
            }
        for (auto &p : pending) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        for (auto e : pool) (void)hipEventDestroy(e);
        if (xev) (void)hipEventDestroy(xev);
    }
};

struct fgl_mesh {
    fgl_ctx *ctx = nullptr;  // home context: its stream and error slot serve the mesh's own calls
    int device = -1;         // the device the arrays live on (fgl_mesh_destroy needs no context)
    int64_t ntri = 0, nline = 0;
    DevBuf tpos, tnrm, tuv, tcol, lpos, lnrm, luv, lcol;
    bool t_has[4] = {false, false, false, false};
    bool l_has[4] = {false, false, false, false};
};

namespace {

#define HIPCHK(ctx, call)                                                                        \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) {                                                                  \
            set_err((ctx)->err, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                    __LINE__);                                                                   \
            return FGL_E_DEVICE;                                                                 \
        }                                                                                        \
    } while (0)

hipEvent_t ev_get(fgl_ctx *c) {
    if (!c->pool.empty()) {
        hipEvent_t e = c->pool.back();
        c->pool.pop_back();
        return e;
    }
    // Timing only: without the system-scope release/acquire an event record
    // does not write back and invalidate the caches between kernels.
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
    return e;
}

// Collects the durations of finished profiled launches (in launch order;
// stops at the first one still running).
void prof_collect(fgl_ctx *c) {
    size_t k = 0;
    for (; k < c->pending.size(); k++) {
        EvPair &p = c->pending[k];
        if (hipEventQuery(p.b) == hipErrorNotReady) break;
        float ms = 0;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            KStat &st = c->kstats[p.name];
            st.launches++;
            st.ms += ms;
        }
        c->pool.push_back(p.a);
        c->pool.push_back(p.b);
    }
    c->pending.erase(c->pending.begin(), c->pending.begin() + (long)k);
}

// Runs one launch, bracketed by events when profiling.
template <class F>
int launch(fgl_ctx *c, const char *name, F &&f) {
    hipEvent_t a = nullptr, b = nullptr;
    const bool timed = c->prof && (c->prof_only.empty() || c->prof_only == name);
    if (timed) {
        a = ev_get(c);
        b = ev_get(c);
        if (a) (void)hipEventRecord(a, c->stream);
    }
    // an error left by an earlier call whose result was ignored (a destroy
    // path, a timing event) is not this launch's
    (void)hipGetLastError();
    hipError_t e = f();
    if (e != hipSuccess) {
        set_err(c->err, "launch of %s failed: %s", name, hipGetErrorString(e));
        return FGL_E_DEVICE;
    }
    if (timed && a && b) {
        (void)hipEventRecord(b, c->stream);
        c->pending.push_back(EvPair{name, a, b});
    }
    return FGL_OK;
}

int select_device(std::string &err) {
    if (g_device >= 0) {
        if (hipSetDevice(g_device) != hipSuccess) {
            set_err(err, "hipSetDevice(%d) failed", g_device);
            return FGL_E_DEVICE;
        }
        return FGL_OK;
    }
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n < 1) {
        set_err(err, "no HIP device available (%s)", hipGetErrorString(e));
        return FGL_E_DEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
        set_err(err, "hipGetDeviceProperties failed");
        return FGL_E_DEVICE;
    }
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_err(err, "device 0 is %s; this library is built for gfx950", prop.gcnArchName);
        return FGL_E_DEVICE;
    }
    g_device = 0;
    if (hipSetDevice(0) != hipSuccess) {
        set_err(err, "hipSetDevice(0) failed");
        return FGL_E_DEVICE;
    }
    return FGL_OK;
}

// Screen(w, h) (matrix.go:119-128)
void screen_matrix(int32_t w, int32_t h, double *S) {
    double w2 = (double)w / 2;
    double h2 = (double)h / 2;
    const double m[16] = {w2, 0, 0, w2, 0, -h2, 0, h2, 0, 0, 0.5, 0.5, 0, 0, 0, 1};
    std::memcpy(S, m, sizeof m);
}

// Makes waiter's stream wait for the work queued so far on signaller's.
int stream_after(fgl_ctx *waiter, fgl_ctx *signaller) {
    if (waiter->stream == signaller->stream) return FGL_OK;
    if (!signaller->xev) HIPCHK(signaller, hipEventCreateWithFlags(&signaller->xev, hipEventDisableTiming));
    HIPCHK(signaller, hipEventRecord(signaller->xev, signaller->stream));
    HIPCHK(waiter, hipStreamWaitEvent(waiter->stream, signaller->xev, 0));
    return FGL_OK;
}

// Writes the pending clears and the blocks a draw left at the clear values
// into the channels `chans` (CB_COLOR / CB_DEPTH) of the framebuffer, for a
// reader of those channels.
int materialize(fgl_ctx *c, uint8_t chans = CB_COLOR | CB_DEPTH) {
    int64_t n = (int64_t)c->W * c->H;
    if (c->pend_color && (chans & CB_COLOR)) {
        int r = launch(c, "fill_color", [&] { return launch_fill_color(c->color.as<uint32_t>(), n, c->clear_color, c->stream); });
        if (r) return r;
        c->pend_color = false;
    }
    if (c->pend_depth && (chans & CB_DEPTH)) {
        int r = launch(c, "fill_depth", [&] { return launch_fill_depth(c->depth.as<double>(), n, c->clear_depth, c->stream); });
        if (r) return r;
        c->pend_depth = false;
    }
    if (c->cblk_any) {  // the blocks a draw left at their clear values
        int r = launch(c, "fill_blocks", [&] {
            return launch_cblk_fill(c->color.as<uint32_t>(), c->depth.as<double>(), c->W, c->H, c->cblk.as<uint8_t>(),
                                    c->cbw, c->cbh, c->clear_color, c->clear_depth, chans, c->stream);
        });
        if (r) return r;
        if (chans == (CB_COLOR | CB_DEPTH)) c->cblk_any = false;  // else marks of the other channel may remain
    }
    return FGL_OK;
}

// The channel `bits` of the whole framebuffer was just overwritten (a host
// write, a resolve or network output): no block of it holds a deferred clear.
int cblk_drop(fgl_ctx *c, uint8_t bits) {
    if (!c->cblk_any) return FGL_OK;
    return launch(c, "drop_blocks", [&] {
        return launch_cblk_drop(c->cblk.as<uint8_t>(), (int64_t)c->cbw * c->cbh, bits, c->stream);
    });
}

int key_bits_for(int32_t ntiles) {
    int b = 1;
    while ((1ll << b) < ntiles) b++;
    return b;
}

int fill_shader(fgl_ctx *c, const fgl_shader *sh, DevShader &d) {
    std::memset(&d, 0, sizeof d);
    d.kind = sh->kind;
    if (sh->kind != FGL_SHADER_SOLID && sh->kind != FGL_SHADER_TEXTURE && sh->kind != FGL_SHADER_PHONG) {
        set_err(c->err, "unknown shader kind %d (custom shaders run on the CPU path)", sh->kind);
        return FGL_E_UNSUPPORTED;
    }
    std::memcpy(d.M, sh->matrix, sizeof d.M);
    std::memcpy(d.color, sh->color, sizeof d.color);
    std::memcpy(d.light, sh->light_dir, sizeof d.light);
    std::memcpy(d.cam, sh->camera_pos, sizeof d.cam);
    std::memcpy(d.obj, sh->object_color, sizeof d.obj);
    std::memcpy(d.amb, sh->ambient, sizeof d.amb);
    std::memcpy(d.dif, sh->diffuse, sizeof d.dif);
    std::memcpy(d.spc, sh->specular, sizeof d.spc);
    d.power = sh->specular_power;
    {
        // the y-only part of math.Pow(x, y), y = SpecularPower
        const double ay = std::fabs(d.power);
        double yi = std::trunc(ay);
        double yf = ay - yi;
        d.pw_frac = yf != 0;
        d.pw_big = yi >= 9223372036854775808.0;
        if (yf != 0 && yf > 0.5) {
            yf = yf - 1;
            yi = yi + 1;
        }
        d.pw_yf = yf;
        d.pw_yi = (std::isfinite(yi) && yi < 9223372036854775808.0) ? (uint64_t)yi : 0;
        d.pw_niter = 0;
        for (uint64_t v = d.pw_yi; v; v >>= 1) d.pw_niter++;
        // |log2 x| <= L keeps x^k within 2^+-1000 for every k <= yi + 1/2
        // (the squares, the partial products, x^yf and the reciprocal)
        const double L = std::min(1000.0 / ((double)d.pw_yi + 0.5), 1000.0);
        d.pw_lo = std::exp2(-L);
        d.pw_hi = std::exp2(L);
    }
    // ObjectColor != Discard compares all four channels with 0 (shader.go:76)
    d.obj_is_discard = (sh->object_color[0] == 0 && sh->object_color[1] == 0 && sh->object_color[2] == 0 &&
                        sh->object_color[3] == 0);
    bool want_tex = sh->kind == FGL_SHADER_TEXTURE || (sh->kind == FGL_SHADER_PHONG && sh->texture > 0);
    if (want_tex) {
        auto it = c->tex.find(sh->texture);
        if (it == c->tex.end()) {
            set_err(c->err, "texture handle %lld is not a texture of this context", (long long)sh->texture);
            return FGL_E_INVALID;
        }
        d.has_tex = 1;
        d.tex_kind = it->second.kind;
        d.tex_w = it->second.w;
        d.tex_h = it->second.h;
        d.tex_pix = it->second.d;
    }
    return FGL_OK;
}

constexpr int64_t kMaxPrims = (int64_t)1 << 26;
// Side vertices (from clipping) per draw; keeps the shading kernels'
// attribute offsets within 32 bits.
constexpr uint64_t kMaxSideVtx = (uint64_t)1 << 28;

// capture_ctr: when not null, the launches are being recorded into a graph:
// they are enqueued once with the context's current capacities (buffers
// already sized by an earlier draw of the same inputs), the draw counters
// are copied to capture_ctr by a kernel, and nothing is read back.
// ---------------------------------------------------------------------------
// Tile shape and starting capacities per workload.  The autotuner
// (fgl_autotune_mesh) times every compiled shape on a real draw and records
// the fastest, with the pair / span / fragment capacities that draw needed,
// under (width, height, primitive kind, bit length of the primitive count);
// draws look the table up.  Every shape gives the same images: the shape
// only changes how the framebuffer is cut into tiles.

const TileOps *const kShapes[] = {
#define FGL_SHAPE_PTR(w, h) &tile_ops_##w##x##h(),
    FGL_TILE_SHAPES(FGL_SHAPE_PTR)
#undef FGL_SHAPE_PTR
};
constexpr int kNumShapes = (int)(sizeof kShapes / sizeof kShapes[0]);

const TileOps *find_shape(int32_t w, int32_t h) {
    for (const TileOps *t : kShapes)
        if (t->tw == w && t->th == h) return t;
    return nullptr;
}

struct TuneKey {
    int32_t W, H, kind, size_class;
    bool operator<(const TuneKey &o) const {
        if (W != o.W) return W < o.W;
        if (H != o.H) return H < o.H;
        if (kind != o.kind) return kind < o.kind;
        return size_class < o.size_class;
    }
};
struct TuneEntry {
    int32_t tw = 32, th = 32;
    uint32_t pair_cap = 0, span_cap = 0, frag_cap = 0;
    double ms = 0;  // the timed draw (informative)
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneEntry> g_tune;

int32_t size_class(int64_t n) {
    int32_t b = 0;
    while (b < 62 && ((int64_t)1 << b) <= n) b++;
    return b;
}

// The shape a draw of n primitives uses: the context's own setting, else
// the table's entry (whose capacities are reserved up front), else 32 x 32.
const TileOps *draw_ops(fgl_ctx *c, int kind, int64_t n) {
    if (c->tile_w) return find_shape(c->tile_w, c->tile_h);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(TuneKey{c->W, c->H, kind, size_class(n)});
    if (it == g_tune.end()) return kShapes[0];
    const TuneEntry &e = it->second;
    c->pair_cap = std::max(c->pair_cap, e.pair_cap);
    c->span_cap = std::max(c->span_cap, e.span_cap);
    c->frag_cap = std::max(c->frag_cap, e.frag_cap);
    const TileOps *t = find_shape(e.tw, e.th);
    return t ? t : kShapes[0];
}

// UpdatedPixels of a draw: the sum of the framebuffer update's partials
uint64_t updated_pixels(const DrawCounters &h) {
    uint64_t s = 0;
    for (int k = 0; k < UPD_PARTS; k++) s += h.upd[k];
    return s;
}

uint32_t cap_for(uint32_t lo, uint32_t hi) {
    const uint64_t v = ((uint64_t)hi << 32 | lo) * 5 / 4 + 65536;
    return (uint32_t)std::min<uint64_t>(v, 0xfffff000ull);
}

// The draw-wide fields of a draw's kernel arguments: framebuffer, tiles,
// state (context.go:47-55), shader and input; the buffers come later.
void draw_setup(fgl_ctx *c, const fgl_state *st, const TileOps &T, int kind, const DevPrims &in, DrawArgs &a) {
    a.W = c->W;
    a.H = c->H;
    a.ntx = (c->W + T.tw - 1) / T.tw;
    a.nty = (c->H + T.th - 1) / T.th;
    a.ntiles = a.ntx * a.nty;
    a.kind = kind;
    a.in = in;
    a.st.read_depth = st->read_depth != 0;
    a.st.write_depth = st->write_depth != 0;
    a.st.write_color = st->write_color != 0;
    a.st.alpha_blend = st->alpha_blend != 0;
    a.st.wireframe = st->wireframe != 0;
    a.st.front_face = st->front_face;
    a.st.cull = st->cull;
    a.st.line_width = st->line_width;
    a.st.depth_bias = st->depth_bias;
    // Rejecting bz > depth before shading equals the reference's locked test
    // only while stored depth cannot increase.
    a.st.early_z = a.st.read_depth && (!a.st.write_depth || st->depth_bias >= 0);
    screen_matrix(c->W, c->H, a.S);
    a.color = c->color.as<uint32_t>();
    a.depth = c->depth.as<double>();
    a.nvsblocks = (int32_t)((in.n + VS_THREADS - 1) / VS_THREADS);
    a.key_bits = key_bits_for(a.ntiles);
    a.sorted_buf = ((a.key_bits + 7) / 8) & 1;
    a.clear_color_pending = c->pend_color;
    a.clear_depth_pending = c->pend_depth;
    a.clear_color = c->clear_color;
    a.clear_depth = c->clear_depth;
    a.cblk = c->cblk.as<uint8_t>();
    a.cbw = c->cbw;
}

// One pass of a draw: the kernel sequence on the context's stream.
int draw_pass(fgl_ctx *c, const TileOps &T, const DrawArgs &a, int kind) {
    int r;
    const int npass = (a.key_bits + 7) / 8;
    if ((r = launch(c, "zero", [&] { return T.zero(a, c->stream); }))) return r;
    if ((r = launch(c, "vertex", [&] { return T.vs(a, c->stream); }))) return r;
    if (kind == PRIM_TRI && !a.st.wireframe)
        if ((r = launch(c, "vertex_clip", [&] { return T.vs_deferred(a, c->stream); }))) return r;
    if ((r = launch(c, "scan", [&] { return T.scan_top(a, c->stream); }))) return r;
    if ((r = launch(c, "emit", [&] { return T.emit(a, c->stream); }))) return r;
    for (int p = 0; p < npass; p++) {
        if ((r = launch(c, "radix_hist", [&] { return T.radix_hist(a, p, c->stream); }))) return r;
        if ((r = launch(c, "radix_scan", [&] { return T.radix_scan(a, c->stream); }))) return r;
        if ((r = launch(c, "radix_scatter", [&] { return T.radix_scatter(a, p, c->stream); }))) return r;
    }
    if ((r = launch(c, "span_count", [&] { return T.span_count(a, c->stream); }))) return r;
    if ((r = launch(c, "span_scan", [&] { return T.span_scan(a, c->stream); }))) return r;
    if ((r = launch(c, "spans", [&] { return T.spans(a, c->stream); }))) return r;
    if ((r = launch(c, "frag_scan", [&] { return T.frag_scan(a, c->stream); }))) return r;
    if ((r = launch(c, "shade", [&] { return T.shade(a, c->stream); }))) return r;
    return launch(c, "rop", [&] { return T.rop(a, c->stream); });
}

// The first part of a draw: checks, the shader, the framebuffer and state
// fields, the starting capacities and the buffers whose size does not depend
// on them.  An empty draw returns FGL_OK with a.in.n == 0 and T unset.
int draw_begin(fgl_ctx *c, const fgl_state *st, const fgl_shader *sh, int kind, const DevPrims &in, const TileOps *&T,
               DrawArgs &a) {
    if (in.n < 0 || in.n > kMaxPrims) {
        set_err(c->err, "primitive count %lld outside [0, %lld]", (long long)in.n, (long long)kMaxPrims);
        return FGL_E_INVALID;
    }
    std::memset(&a, 0, sizeof a);
    int r = fill_shader(c, sh, a.sh);
    if (r) return r;
    if (in.n == 0) return FGL_OK;
    if (hipSetDevice(c->device) != hipSuccess) {
        set_err(c->err, "hipSetDevice failed");
        return FGL_E_DEVICE;
    }
    T = draw_ops(c, kind, in.n);
    draw_setup(c, st, *T, kind, in, a);
    const int64_t n = in.n;

    // starting capacities; grown on overflow
    uint64_t want_side = kind == PRIM_LINE ? (uint64_t)2 * n : (st->wireframe ? (uint64_t)6 * n : (uint64_t)n / 8);
    want_side += 4096;
    uint64_t want_vtx = kind == PRIM_LINE ? 256 : (uint64_t)n / 4 + 4096;
    uint64_t want_pairs = (uint64_t)2 * n + (uint64_t)a.ntiles + 65536;
    if (c->side_rec_cap < want_side) c->side_rec_cap = (uint32_t)std::min<uint64_t>(want_side, 0x7fffffffu);
    if (c->side_vtx_cap < want_vtx) c->side_vtx_cap = (uint32_t)std::min<uint64_t>(want_vtx, kMaxSideVtx);
    if (c->pair_cap < want_pairs) c->pair_cap = (uint32_t)std::min<uint64_t>(want_pairs, 0xfffff000u);
    uint64_t want_spans = (uint64_t)3 * c->pair_cap;
    if (c->span_cap < want_spans) c->span_cap = (uint32_t)std::min<uint64_t>(want_spans, 0xfffff000u);
    uint64_t want_frags = (uint64_t)2 * c->span_cap;
    if (c->frag_cap < want_frags) c->frag_cap = (uint32_t)std::min<uint64_t>(want_frags, 0xfffff000u);

    HIPCHK(c, c->ctr.ensure(sizeof(DrawCounters)));
    if (!c->zeros.p) {
        HIPCHK(c, c->zeros.ensure(16 * sizeof(double)));
        HIPCHK(c, hipMemsetAsync(c->zeros.p, 0, 16 * sizeof(double), c->stream));
    }
    HIPCHK(c, c->ntl.ensure((size_t)n * 4));
    HIPCHK(c, c->nrast.ensure((size_t)n * 4));
    HIPCHK(c, c->sstart.ensure((size_t)n * 4));
    HIPCHK(c, c->blocksum.ensure((size_t)a.nvsblocks * 8));
    HIPCHK(c, c->blockoff.ensure((size_t)a.nvsblocks * 8));
    if (kind == PRIM_TRI) HIPCHK(c, c->rec_main.ensure((size_t)n * sizeof(PrimRec)));
    if (kind == PRIM_TRI) HIPCHK(c, c->tbox.ensure((size_t)n * sizeof(int4)));
    if (kind == PRIM_TRI) HIPCHK(c, c->deferred.ensure((size_t)n * 4));
    HIPCHK(c, c->dtot.ensure(256 * 4));
    HIPCHK(c, c->tstart.ensure((size_t)a.ntiles * 4));
    HIPCHK(c, c->tend.ensure((size_t)a.ntiles * 4));
    return FGL_OK;
}

// Sizes the buffers that depend on the capacities (again after they grow)
// and points a at every buffer of the context.
int draw_buffers(fgl_ctx *c, DrawArgs &a) {
    HIPCHK(c, c->side_rec.ensure((size_t)c->side_rec_cap * sizeof(PrimRec)));
    HIPCHK(c, c->side_vtx.ensure((size_t)c->side_vtx_cap * SIDE_VTX_DOUBLES * 8));
    HIPCHK(c, c->keys0.ensure((size_t)c->pair_cap * 4));
    HIPCHK(c, c->vals0.ensure((size_t)c->pair_cap * 4));
    HIPCHK(c, c->keys1.ensure((size_t)c->pair_cap * 4));
    HIPCHK(c, c->vals1.ensure((size_t)c->pair_cap * 4));
    a.nrsblocks = (int32_t)(((uint64_t)c->pair_cap + RS_BLOCK - 1) / RS_BLOCK);
    HIPCHK(c, c->hist.ensure((size_t)256 * a.nrsblocks * 4));
    const size_t nsb = ((size_t)c->pair_cap + SPAN_THREADS - 1) / SPAN_THREADS + 1;
    HIPCHK(c, c->pspans.ensure((size_t)c->pair_cap * 4));
    HIPCHK(c, c->prows.ensure((size_t)c->pair_cap * 2));
    HIPCHK(c, c->span_bsum.ensure(nsb * 8));
    HIPCHK(c, c->span_boff.ensure(nsb * 8));
    HIPCHK(c, c->poff.ensure(((size_t)c->pair_cap + 1) * 4));
    HIPCHK(c, c->spans.ensure((size_t)c->span_cap * sizeof(Span)));
    a.pspans = c->pspans.as<uint32_t>();
    a.prows = c->prows.as<uint16_t>();
    a.span_blocksum = c->span_bsum.as<uint64_t>();
    a.span_blockoff = c->span_boff.as<uint64_t>();
    a.poff = c->poff.as<uint32_t>();
    a.spans = c->spans.as<Span>();
    a.span_cap = c->span_cap;
    const size_t nfb = ((size_t)c->span_cap + FRAG_BLOCK - 1) / FRAG_BLOCK + 1;
    HIPCHK(c, c->frag_bsum.ensure(nfb * 8));
    HIPCHK(c, c->frag_boff.ensure(nfb * 8));
    HIPCHK(c, c->frag_gsum.ensure((nfb / 1024 + 2) * 8));
    HIPCHK(c, c->foff.ensure(((size_t)c->span_cap + 1) * 4));
    HIPCHK(c, c->frags.ensure((size_t)c->frag_cap * sizeof(Frag)));
    a.frag_blocksum = c->frag_bsum.as<uint64_t>();
    a.frag_blockoff = c->frag_boff.as<uint64_t>();
    a.frag_gsum = c->frag_gsum.as<uint64_t>();
    a.foff = c->foff.as<uint32_t>();
    a.frags = c->frags.as<Frag>();
    a.frag_cap = c->frag_cap;

    a.rec_main = c->rec_main.as<PrimRec>();
    a.tbox = c->tbox.as<int4>();
    a.side_rec = c->side_rec.as<PrimRec>();
    a.side_rec_cap = c->side_rec_cap;
    a.side_vtx = c->side_vtx.as<double>();
    a.side_vtx_cap = c->side_vtx_cap;
    a.ntl = c->ntl.as<uint32_t>();
    a.nrast = c->nrast.as<uint32_t>();
    a.sstart = c->sstart.as<uint32_t>();
    a.blocksum = c->blocksum.as<uint64_t>();
    a.blockoff = c->blockoff.as<uint64_t>();
    a.keys[0] = c->keys0.as<uint32_t>();
    a.vals[0] = c->vals0.as<uint32_t>();
    a.keys[1] = c->keys1.as<uint32_t>();
    a.vals[1] = c->vals1.as<uint32_t>();
    a.pair_cap = c->pair_cap;
    a.hist = c->hist.as<uint32_t>();
    a.dtot = c->dtot.as<uint32_t>();
    a.deferred = c->deferred.as<uint32_t>();
    a.tstart = c->tstart.as<uint32_t>();
    a.tend = c->tend.as<uint32_t>();
    a.ctr = c->ctr.as<DrawCounters>();
    a.zeros = c->zeros.as<double>();
    return FGL_OK;
}

int run_draw(fgl_ctx *c, const fgl_state *st, const fgl_shader *sh, int kind, const DevPrims &in, uint64_t info[2],
             uint32_t *capture_ctr = nullptr) {
    if (info) info[0] = info[1] = 0;
    DrawArgs a;
    const TileOps *Tp = nullptr;
    int r = draw_begin(c, st, sh, kind, in, Tp, a);
    if (r || in.n == 0) return r;
    const TileOps &T = *Tp;

    if (!capture_ctr) {
        c->last_overflow = 0;
        c->last_reruns = 0;
    }
    for (int attempt = 0; attempt < 6; attempt++) {
        if ((r = draw_buffers(c, a))) return r;
        if (attempt == 0 && !capture_ctr) {
            // failure injection: the kernels see smaller capacities than the
            // buffers hold, so an overflow check that is off cannot write
            // outside them
            const uint32_t *f = c->force_caps;
            if (f[0]) a.side_rec_cap = std::min(a.side_rec_cap, f[0]);
            if (f[1]) a.side_vtx_cap = std::min(a.side_vtx_cap, f[1]);
            if (f[2]) a.pair_cap = std::min(a.pair_cap, f[2]);
            if (f[3]) a.span_cap = std::min(a.span_cap, f[3]);
            if (f[4]) a.frag_cap = std::min(a.frag_cap, f[4]);
            std::memset(c->force_caps, 0, sizeof c->force_caps);
        }

        if ((r = draw_pass(c, T, a, kind))) return r;
        if (capture_ctr)
            return launch(c, "counters", [&] {
                return launch_copy_words(reinterpret_cast<const uint32_t *>(a.ctr), capture_ctr,
                                         (int32_t)(sizeof(DrawCounters) / 4), c->stream);
            });

        HIPCHK(c, c->hctr.ensure(sizeof(DrawCounters)));
        HIPCHK(c, hipMemcpyAsync(c->hctr.p, a.ctr, sizeof(DrawCounters), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, wait_stream(c->stream));
        const DrawCounters h = *c->hctr.as<DrawCounters>();
        if (c->prof) prof_collect(c);
        c->last = h;
        c->last_overflow |= h.overflow;
        if (h.overflow) c->last_reruns++;
        if (h.overflow == 0) {
            if (info) {
                info[0] = h.total;
                info[1] = updated_pixels(h);
            }
            // the ROP marked the blocks of fragment-free tiles with the
            // clears it consumed
            if (c->pend_color || c->pend_depth) c->cblk_any = true;
            c->pend_color = false;
            c->pend_depth = false;
            return FGL_OK;
        }
        if (h.overflow & 16u) {
            uint64_t need = ((uint64_t)h.spans_hi << 32 | h.spans_lo) * 5 / 4 + 65536;
            if (need > 0xfffff000ull) {
                set_err(c->err, "draw needs %llu coverage spans, more than 2^32", (unsigned long long)need);
                return FGL_E_UNSUPPORTED;
            }
            c->span_cap = (uint32_t)need;
        }
        if (h.overflow & 32u) {
            uint64_t need = ((uint64_t)h.frags_hi << 32 | h.frags_lo) * 5 / 4 + 65536;
            if (need > 0xfffff000ull) {
                set_err(c->err, "draw needs %llu fragments, more than 2^32", (unsigned long long)need);
                return FGL_E_UNSUPPORTED;
            }
            c->frag_cap = (uint32_t)need;
        }
        if (h.overflow & 1u) c->side_rec_cap = (uint32_t)std::min<uint64_t>((uint64_t)h.side_rec * 5 / 4 + 4096, 0x7fffffffu);
        if (h.overflow & 2u) {
            if (c->side_vtx_cap >= kMaxSideVtx) {
                set_err(c->err, "draw needs more than %llu clipped vertices", (unsigned long long)kMaxSideVtx);
                return FGL_E_UNSUPPORTED;
            }
            c->side_vtx_cap = (uint32_t)std::min<uint64_t>((uint64_t)h.side_vtx * 5 / 4 + 4096, kMaxSideVtx);
        }
        if (h.overflow & 4u) {
            uint64_t need = ((uint64_t)h.pairs_hi << 32 | h.pairs_lo) * 5 / 4 + 65536;
            if (need > 0xfffff000ull) {
                set_err(c->err, "draw needs %llu (tile, triangle) pairs, more than 2^32", (unsigned long long)need);
                return FGL_E_UNSUPPORTED;
            }
            c->pair_cap = (uint32_t)need;
        }
    }
    set_err(c->err, "draw did not fit after repeated capacity growth");
    return FGL_E_NOMEM;
}

int upload(fgl_ctx *c, DevBuf &b, const double *src, size_t count, const double *&dst) {
    if (!src) {
        dst = nullptr;
        return FGL_OK;
    }
    HIPCHK(c, b.ensure(count * 8));
    HIPCHK(c, hipMemcpyAsync(b.p, src, count * 8, hipMemcpyHostToDevice, c->stream));
    dst = b.as<double>();
    return FGL_OK;
}

int check_ctx(const fgl_ctx *c) {
    if (!c) {
        g_err = "null context";
        return FGL_E_INVALID;
    }
    return FGL_OK;
}

}  // namespace

// ===========================================================================
// C ABI

extern "C" {

int fgl_init(int32_t device) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n < 1) {
        set_err(g_err, "no HIP device available (%s)", hipGetErrorString(e));
        return FGL_E_DEVICE;
    }
    if (device < 0 || device >= n) {
        set_err(g_err, "device %d out of range [0, %d)", device, n);
        return FGL_E_INVALID;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        set_err(g_err, "hipGetDeviceProperties failed");
        return FGL_E_DEVICE;
    }
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_err(g_err, "device %d is %s; this library is built for gfx950", device, prop.gcnArchName);
        return FGL_E_DEVICE;
    }
    if (hipSetDevice(device) != hipSuccess) {
        set_err(g_err, "hipSetDevice(%d) failed", device);
        return FGL_E_DEVICE;
    }
    g_device = device;
    return FGL_OK;
}

const char *fgl_last_error(const fgl_ctx *ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

int fgl_ctx_create(int32_t width, int32_t height, fgl_ctx **out) {
    if (!out || width <= 0 || height <= 0 || (int64_t)width * height > ((int64_t)1 << 31)) {
        set_err(g_err, "fgl_ctx_create: invalid size %d x %d", width, height);
        return FGL_E_INVALID;
    }
    *out = nullptr;
    int r = select_device(g_err);
    if (r) return r;
    fgl_ctx *c = new (std::nothrow) fgl_ctx;
    if (!c) {
        g_err = "out of host memory";
        return FGL_E_NOMEM;
    }
    c->W = width;
    c->H = height;
    (void)hipGetDevice(&c->device);
    size_t npix = (size_t)width * height;
    c->cbw = (width + CB_W - 1) / CB_W;
    c->cbh = (height + CB_H - 1) / CB_H;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess || c->color.ensure(npix * 4) ||
        c->depth.ensure(npix * 8) || c->cblk.ensure((size_t)c->cbw * c->cbh)) {
        set_err(g_err, "device allocation of a %d x %d framebuffer failed", width, height);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return FGL_E_NOMEM;
    }
    // image.NewNRGBA starts zeroed; NewContext clears depth to MaxFloat64.
    if (hipMemsetAsync(c->color.p, 0, npix * 4, c->stream) != hipSuccess ||
        hipMemsetAsync(c->cblk.p, 0, (size_t)c->cbw * c->cbh, c->stream) != hipSuccess) {
        g_err = "hipMemsetAsync failed";
        (void)hipStreamDestroy(c->stream);
        delete c;
        return FGL_E_DEVICE;
    }
    c->pend_depth = true;
    c->clear_depth = DBL_MAX;
    *out = c;
    return FGL_OK;
}

void fgl_ctx_destroy(fgl_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    hipStream_t s = ctx->stream;
    delete ctx;
    (void)hipStreamDestroy(s);
    (void)hipGetLastError();  // the calls above report nothing to the next launch
}

int fgl_ctx_size(const fgl_ctx *ctx, int32_t *width, int32_t *height) {
    if (check_ctx(ctx)) return FGL_E_INVALID;
    if (width) *width = ctx->W;
    if (height) *height = ctx->H;
    return FGL_OK;
}

int fgl_clear_color(fgl_ctx *ctx, const uint8_t nrgba[4]) {
    if (check_ctx(ctx)) return FGL_E_INVALID;
    if (!nrgba) {
        ctx->err = "fgl_clear_color: null color";
        return FGL_E_INVALID;
    }
    ctx->clear_color = (uint32_t)nrgba[0] | ((uint32_t)nrgba[1] << 8) | ((uint32_t)nrgba[2] << 16) |
                       ((uint32_t)nrgba[3] << 24);
    ctx->pend_color = true;
    return FGL_OK;
}

int fgl_clear_depth(fgl_ctx *ctx, double value) {
    if (check_ctx(ctx)) return FGL_E_INVALID;
    ctx->clear_depth = value;
    ctx->pend_depth = true;
    return FGL_OK;
}

int fgl_texture_create(fgl_ctx *ctx, int32_t width, int32_t height, int32_t kind, const void *texels,
                       int64_t *handle) {
    if (check_ctx(ctx)) return FGL_E_INVALID;
    if (!texels || !handle || width <= 0 || height <= 0 ||
        (kind != FGL_TEX_NRGBA8 && kind != FGL_TEX_RGBA8 && kind != FGL_TEX_RGBA16)) {
        ctx->err = "fgl_texture_create: invalid argument";
        return FGL_E_INVALID;
    }
    size_t bytes = (size_t)width * height * (kind == FGL_TEX_RGBA16 ? 8 : 4);
    Tex t;
    t.w = width;
    t.h = height;
    t.kind = kind;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipMalloc(&t.d, bytes));
    hipError_t e = hipMemcpy(t.d, texels, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(t.d);
        set_err(ctx->err, "texture upload failed: %s", hipGetErrorString(e));
        return FGL_E_DEVICE;
    }
    int64_t h = ctx->next_tex++;
    ctx->tex[h] = t;
    *handle = h;
    return FGL_OK;
}

int fgl_texture_destroy(fgl_ctx *ctx, int64_t handle) {
    if (check_ctx(ctx)) return FGL_E_INVALID;
    auto it = ctx->tex.find(handle);
    if (it == ctx->tex.end()) {
        ctx->err = "fgl_texture_destroy: unknown handle";
        return FGL_E_INVALID;
    }
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(it->second.d);
    g_dev_epoch++;
    ctx->tex.erase(it);
    return FGL_OK;
}

static int draw_host(fgl_ctx *ctx, const fgl_state *state, const fgl_shader *shader, const fgl_prims *p, int kind,
                     uint64_t info[2]) {
    if (check_ctx(ctx)) return FGL_E_INVALID;
    if (!state || !shader || !p) {
        ctx->err = "draw: null state, shader or primitives";
        return FGL_E_INVALID;
    }
    if (p->count < 0 || p->count > kMaxPrims) {
        set_err(ctx->err, "draw: primitive count %lld outside [0, %lld]", (long long)p->count, (long long)kMaxPrims);
        return FGL_E_INVALID;
    }
    if (p->count > 0 && !p->position) {
        ctx->err = "draw: position array is required";
        return FGL_E_INVALID;
    }
    if (hipSetDevice(ctx->device) != hipSuccess) {
        ctx->err = "hipSetDevice failed";
        return FGL_E_DEVICE;
    }
    const size_t nv = (size_t)p->count * (kind == PRIM_TRI ? 3 : 2);
    DevPrims in;
    in.n = p->count;
    int r;
    if ((r = upload(ctx, ctx->in_pos, p->position, nv * 3, in.pos))) return r;
    if ((r = upload(ctx, ctx->in_nrm, p->normal, nv * 3, in.nrm))) return r;
    if ((r = upload(ctx, ctx->in_uv, p->texture, nv * 2, in.uv))) return r;
    if ((r = upload(ctx, ctx->in_col, p->color, nv * 4, in.col))) return r;
    return run_draw(ctx, state, shader, kind, in, info);
}

int fgl_draw_triangles(fgl_ctx *ctx, const fgl_state *state, const fgl_shader *shader, const fgl_prims *tris,
                       uint64_t info[2]) {
    return draw_host(ctx, state, shader, tris, PRIM_TRI, info);
}

int fgl_draw_lines(fgl_ctx *ctx, const fgl_state *state, const fgl_shader *shader, const fgl_prims *lines,
                   uint64_t info[2]) {
    return draw_host(ctx, state, shader, lines, PRIM_LINE, info);
}

static int mesh_upload(fgl_ctx *c, DevBuf &b, const double *src, size_t count, bool &has) {
    has = false;
    if (!src || count == 0) return FGL_OK;
    HIPCHK(c, b.ensure(count * 8));
    HIPCHK(c, hipMemcpy(b.p, src, count * 8, hipMemcpyHostToDevice));
    has = true;
    return FGL_OK;
}

int fgl_mesh_create(fgl_ctx *ctx, const fgl_prims *tris, const fgl_prims *lines, fgl_mesh **out) {
    if (check_ctx(ctx)) return FGL_E_INVALID;
    if (!out) {
        ctx->err = "fgl_mesh_create: null output";
        return FGL_E_INVALID;
    }
    *out = nullptr;
    int64_t nt = tris ? tris->count : 0, nl = lines ? lines->count : 0;
    if (nt < 0 || nl < 0 || nt > kMaxPrims || nl > kMaxPrims || (nt > 0 && !tris->position) ||
        (nl > 0 && !lines->position)) {
        ctx->err = "fgl_mesh_create: invalid primitives";
        return FGL_E_INVALID;
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    fgl_mesh *m = new (std::nothrow) fgl_mesh;
    if (!m) {
        ctx->err = "out of host memory";
        return FGL_E_NOMEM;
    }
    m->ctx = ctx;
    m->device = ctx->device;
    m->ntri = nt;
    m->nline = nl;
    int r = FGL_OK;
    if (nt > 0) {
        size_t nv = (size_t)nt * 3;
        if (!r) r = mesh_upload(ctx, m->tpos, tris->position, nv * 3, m->t_has[0]);
        if (!r) r = mesh_upload(ctx, m->tnrm, tris->normal, nv * 3, m->t_has[1]);
        if (!r) r = mesh_upload(ctx, m->tuv, tris->texture, nv * 2, m->t_has[2]);
        if (!r) r = mesh_upload(ctx, m->tcol, tris->color, nv * 4, m->t_has[3]);
    }
    if (nl > 0) {
        size_t nv = (size_t)nl * 2;
        if (!r) r = mesh_upload(ctx, m->lpos, lines->position, nv * 3, m->l_has[0]);
        if (!r) r = mesh_upload(ctx, m->lnrm, lines->normal, nv * 3, m->l_has[1]);
        if (!r) r = mesh_upload(ctx, m->luv, lines->texture, nv * 2, m->l_has[2]);
        if (!r) r = mesh_upload(ctx, m->lcol, lines->color, nv * 4, m->l_has[3]);
    }
    if (r) {
        delete m;
        return r;
    }
    *out = m;
    return FGL_OK;
}

// The home context may already be gone (a host language's collector may
// free a context before a mesh that refers to it), and any context on the
// device may still draw the mesh: wait for the whole device, not for one
// stream, and read nothing through mesh->ctx.
void fgl_mesh_destroy(fgl_mesh *mesh) {
    if (!mesh) return;
    (void)hipSetDevice(mesh->device);
    (void)hipDeviceSynchronize();
    delete mesh;
    (void)hipGetLastError();  // the calls above report nothing to the next launch
}

int fgl_draw_mesh(fgl_ctx *ctx, const fgl_state *state, const fgl_shader *shader, const fgl_mesh *mesh,
                  uint64_t info[2]) {
    if (check_ctx(ctx)) return FGL_E_INVALID;
    if (!state || !shader || !mesh) {
        ctx->err = "fgl_draw_mesh: null argument";
        return FGL_E_INVALID;
    }
    if (mesh->ctx->device != ctx->device) {
        ctx->err = "fgl_draw_mesh: mesh lives on another device";
        return FGL_E_INVALID;
    }
    uint64_t i1[2] = {0, 0}, i2[2] = {0, 0};
    DevPrims t;
    t.n = mesh->ntri;
    t.pos = mesh->t_has[0] ? mesh->tpos.as<double>() : nullptr;
    t.nrm = mesh->t_has[1] ? mesh->tnrm.as<double>() : nullptr;
    t.uv = mesh->t_has[2] ? mesh->tuv.as<double>() : nullptr;
    t.col = mesh->t_has[3] ? mesh->tcol.as<double>() : nullptr;
    int r = run_draw(ctx, state, shader, PRIM_TRI, t, i1);
    if (r) return r;
    DevPrims l;
    l.n = mesh->nline;
    l.pos = mesh->l_has[0] ? mesh->lpos.as<double>() : nullptr;
    l.nrm = mesh->l_has[1] ? mesh->lnrm.as<double>() : nullptr;
    l.uv = mesh->l_has[2] ? mesh->luv.as<double>() : nullptr;
    l.col = mesh->l_has[3] ? mesh->lcol.as<double>() : nullptr;
    r = run_draw(ctx, state, shader, PRIM_LINE, l, i2);
    if (r) return r;
    if (info) {
        info[0] = i1[0] + i2[0];
        info[1] = i1[1] + i2[1];
    }
    return FGL_OK;
}

int fgl_tile_shapes(int32_t *wh, int32_t cap, int32_t *n) {
    if (!n || (cap > 0 && !wh)) {
        g_err = "fgl_tile_shapes: null argument";
        return FGL_E_INVALID;
    }
    *n = kNumShapes;
    for (int i = 0; i < kNumShapes && i < cap; i++) {
        wh[2 * i] = kShapes[i]->tw;
        wh[2 * i + 1] = kShapes[i]->th;
    }
    return FGL_OK;
}

int fgl_ctx_set_tile(fgl_ctx *ctx, int32_t tile_w, int32_t tile_h) {
    if (check_ctx(ctx)) return FGL_E_INVALID;
    if ((tile_w || tile_h) && !find_shape(tile_w, tile_h)) {
        set_err(ctx->err, "fgl_ctx_set_tile: %d x %d is not a compiled tile shape", tile_w, tile_h);
        return FGL_E_INVALID;
    }
    ctx->tile_w = tile_w;
    ctx->tile_h = tile_h;
    return FGL_OK;
}

int fgl_ctx_get_tile(fgl_ctx *ctx, int64_t ntri, int32_t out_wh[2]) {
    if (check_ctx(ctx)) return FGL_E_INVALID;
    if (!out_wh || ntri < 0) {
        ctx->err = "fgl_ctx_get_tile: invalid argument";
        return FGL_E_INVALID;
    }
    const int32_t sw = ctx->tile_w, sh = ctx->tile_h;
    const uint32_t caps[3] = {ctx->pair_cap, ctx->span_cap, ctx->frag_cap};
    const TileOps *t = draw_ops(ctx, PRIM_TRI, ntri);
    ctx->pair_cap = caps[0];  // a query reserves nothing
    ctx->span_cap = caps[1];
    ctx->frag_cap = caps[2];
    ctx->tile_w = sw;
    ctx->tile_h = sh;
    out_wh[0] = t->tw;
    out_wh[1] = t->th;
    return FGL_OK;
}

int fgl_autotune_mesh(fgl_ctx *ctx, const fgl_state *state, const fgl_shader *shader, const fgl_mesh *mesh,
                      int32_t reps, int32_t out_wh[2]) {
    if (check_ctx(ctx)) return FGL_E_INVALID;
    if (!state || !shader || !mesh || reps < 1 || reps > 100) {
        ctx->err = "fgl_autotune_mesh: invalid argument";
        return FGL_E_INVALID;
    }
    if (hipSetDevice(ctx->device) != hipSuccess) {
        ctx->err = "hipSetDevice failed";
        return FGL_E_DEVICE;
    }
    // the context's framebuffer, pending clears, shape and counters are put
    // back afterwards: tuning leaves no trace but the table entry
    const size_t npix = (size_t)ctx->W * ctx->H;
    DevBuf keep_color, keep_depth;
    HIPCHK(ctx, keep_color.ensure(npix * 4));
    HIPCHK(ctx, keep_depth.ensure(npix * 8));
    HIPCHK(ctx, hipMemcpyAsync(keep_color.p, ctx->color.p, npix * 4, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(keep_depth.p, ctx->depth.p, npix * 8, hipMemcpyDeviceToDevice, ctx->stream));
    DevBuf keep_cblk;
    const size_t ncb = (size_t)ctx->cbw * ctx->cbh;
    HIPCHK(ctx, keep_cblk.ensure(ncb));
    HIPCHK(ctx, hipMemcpyAsync(keep_cblk.p, ctx->cblk.p, ncb, hipMemcpyDeviceToDevice, ctx->stream));
    const bool pc = ctx->pend_color, pd = ctx->pend_depth, cba = ctx->cblk_any;
    const int32_t sw = ctx->tile_w, sh = ctx->tile_h;
    const DrawCounters last = ctx->last;
    const bool prof = ctx->prof;
    ctx->prof = false;
    TuneEntry best;
    best.ms = -1;
    int r = FGL_OK;
    for (int k = 0; k < kNumShapes && r == FGL_OK; k++) {
        ctx->tile_w = kShapes[k]->tw;
        ctx->tile_h = kShapes[k]->th;
        double t_min = 1e300;
        for (int i = 0; i <= reps && r == FGL_OK; i++) {  // the first draw grows the buffers
            ctx->pend_color = ctx->pend_depth = true;
            HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
            const auto t0 = std::chrono::steady_clock::now();
            r = fgl_draw_mesh(ctx, state, shader, mesh, nullptr);
            const auto t1 = std::chrono::steady_clock::now();
            if (i > 0) t_min = std::min(t_min, std::chrono::duration<double, std::milli>(t1 - t0).count());
        }
        if (r) break;
        if (best.ms < 0 || t_min < best.ms) {
            best.tw = kShapes[k]->tw;
            best.th = kShapes[k]->th;
            best.ms = t_min;
            // the counters are the last draw's: the triangles' when the mesh
            // has no lines (a lines draw would overwrite them; then no
            // capacities are recorded)
            const bool tris_last = mesh->nline == 0;
            best.pair_cap = tris_last ? cap_for(ctx->last.pairs_lo, ctx->last.pairs_hi) : 0;
            best.span_cap = tris_last ? cap_for(ctx->last.spans_lo, ctx->last.spans_hi) : 0;
            best.frag_cap = tris_last ? cap_for(ctx->last.frags_lo, ctx->last.frags_hi) : 0;
        }
    }
    ctx->tile_w = sw;
    ctx->tile_h = sh;
    ctx->prof = prof;
    HIPCHK(ctx, hipMemcpyAsync(ctx->color.p, keep_color.p, npix * 4, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(ctx->depth.p, keep_depth.p, npix * 8, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(ctx->cblk.p, keep_cblk.p, ncb, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->pend_color = pc;
    ctx->pend_depth = pd;
    ctx->cblk_any = cba;
    ctx->last = last;
    if (r) return r;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune[TuneKey{ctx->W, ctx->H, PRIM_TRI, size_class(mesh->ntri)}] = best;
    }
    if (out_wh) {
        out_wh[0] = best.tw;
        out_wh[1] = best.th;
    }
    return FGL_OK;
}

static const char kTuneHeader[] = "fauxgl_amd tile tuning 1";

int fgl_tune_save(const char *path) {
    if (!path) {
        g_err = "fgl_tune_save: null path";
        return FGL_E_INVALID;
    }
    FILE *f = std::fopen(path, "w");
    if (!f) {
        set_err(g_err, "fgl_tune_save: cannot open %s", path);
        return FGL_E_IO;
    }
    std::fprintf(f, "%s\n# width height kind size_class tile_w tile_h pair_cap span_cap frag_cap ms\n", kTuneHeader);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        for (const auto &kv : g_tune)
            std::fprintf(f, "%d %d %d %d %d %d %u %u %u %.6f\n", kv.first.W, kv.first.H, kv.first.kind,
                         kv.first.size_class, kv.second.tw, kv.second.th, kv.second.pair_cap, kv.second.span_cap,
                         kv.second.frag_cap, kv.second.ms);
    }
    const bool ok = std::fclose(f) == 0;
    if (!ok) {
        set_err(g_err, "fgl_tune_save: writing %s failed", path);
        return FGL_E_IO;
    }
    return FGL_OK;
}

int fgl_tune_load(const char *path) {
    if (!path) {
        g_err = "fgl_tune_load: null path";
        return FGL_E_INVALID;
    }
    FILE *f = std::fopen(path, "r");
    if (!f) {
        set_err(g_err, "fgl_tune_load: cannot open %s", path);
        return FGL_E_IO;
    }
    char line[512];
    std::map<TuneKey, TuneEntry> got;
    int lineno = 0, r = FGL_OK;
    while (std::fgets(line, sizeof line, f)) {
        lineno++;
        if (lineno == 1) {
            if (std::strncmp(line, kTuneHeader, sizeof kTuneHeader - 1) != 0) {
                set_err(g_err, "fgl_tune_load: %s is not a tuning file", path);
                r = FGL_E_INVALID;
                break;
            }
            continue;
        }
        if (line[0] == '#' || line[0] == '\n' || line[0] == 0) continue;
        TuneKey k{};
        TuneEntry e;
        if (std::sscanf(line, "%d %d %d %d %d %d %u %u %u %lf", &k.W, &k.H, &k.kind, &k.size_class, &e.tw, &e.th,
                        &e.pair_cap, &e.span_cap, &e.frag_cap, &e.ms) != 10 ||
            !find_shape(e.tw, e.th) || k.W <= 0 || k.H <= 0 || (k.kind != PRIM_TRI && k.kind != PRIM_LINE) ||
            e.pair_cap > 0xfffff000u || e.span_cap > 0xfffff000u || e.frag_cap > 0xfffff000u) {
            set_err(g_err, "fgl_tune_load: %s line %d is malformed or names an unknown tile shape", path, lineno);
            r = FGL_E_INVALID;
            break;
        }
        got[k] = e;
    }
    std::fclose(f);
    if (r == FGL_OK && lineno == 0) {
        set_err(g_err, "fgl_tune_load: %s is empty", path);
        r = FGL_E_INVALID;
    }
    if (r) return r;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (const auto &kv : got) g_tune[kv.first] = kv.second;
    return FGL_OK;
}

int fgl_tune_clear(void) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune.clear();
    return FGL_OK;
}

int fgl_read_color(fgl_ctx *ctx, uint8_t *nrgba) {
    if (check_ctx(ctx)) return FGL_E_INVALID;
    if (!nrgba) {
        ctx->err = "fgl_read_color: null buffer";
        return FGL_E_INVALID;
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    int r = materialize(ctx, CB_COLOR);
    if (r) return r;
    HIPCHK(ctx, hipMemcpyAsync(nrgba, ctx->color.p, (size_t)ctx->W * ctx->H * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->prof) prof_collect(ctx);
    return FGL_OK;
}

int fgl_read_depth(fgl_ctx *ctx, double *depth) {
    if (check_ctx(ctx)) return FGL_E_INVALID;
    if (!depth) {
        ctx->err = "fgl_read_depth: null buffer";
        return FGL_E_INVALID;
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    int r = materialize(ctx, CB_DEPTH);
    if (r) return r;
    HIPCHK(ctx, hipMemcpyAsync(depth, ctx->depth.p, (size_t)ctx->W * ctx->H * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->prof) prof_collect(ctx);
    return FGL_OK;
}

int fgl_write_color(fgl_ctx *ctx, const uint8_t *nrgba) {
    if (check_ctx(ctx)) return FGL_E_INVALID;
    if (!nrgba) {
        ctx->err = "fgl_write_color: null buffer";
        return FGL_E_INVALID;
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipMemcpyAsync(ctx->color.p, nrgba, (size_t)ctx->W * ctx->H * 4, hipMemcpyHostToDevice, ctx->stream));
    int r = cblk_drop(ctx, CB_COLOR);
    if (r) return r;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->pend_color = false;
    return FGL_OK;
}

int fgl_write_depth(fgl_ctx *ctx, const double *depth) {
    if (check_ctx(ctx)) return FGL_E_INVALID;
    if (!depth) {
        ctx->err = "fgl_write_depth: null buffer";
        return FGL_E_INVALID;
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipMemcpyAsync(ctx->depth.p, depth, (size_t)ctx->W * ctx->H * 8, hipMemcpyHostToDevice, ctx->stream));
    int r = cblk_drop(ctx, CB_DEPTH);
    if (r) return r;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->pend_depth = false;
    return FGL_OK;
}

int fgl_depth_image(fgl_ctx *ctx, uint16_t *gray16) {
    if (check_ctx(ctx)) return FGL_E_INVALID;
    if (!gray16) {
        ctx->err = "fgl_depth_image: null buffer";
        return FGL_E_INVALID;
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    int r = materialize(ctx, CB_DEPTH);
    if (r) return r;
    const int64_t n = (int64_t)ctx->W * ctx->H;
    const int32_t nb = 512;
    HIPCHK(ctx, ctx->dpart.ensure((size_t)nb * 16));
    HIPCHK(ctx, ctx->dimg.ensure((size_t)n * 2));
    r = launch(ctx, "depth_minmax", [&] { return launch_depth_minmax(ctx->depth.as<double>(), n, ctx->dpart.as<double>(), nb, ctx->stream); });
    if (r) return r;
    r = launch(ctx, "depth_image", [&] { return launch_depth_image(ctx->depth.as<double>(), n, ctx->dpart.as<double>(), nb, ctx->dimg.as<uint16_t>(), ctx->stream); });
    if (r) return r;
    HIPCHK(ctx, hipMemcpyAsync(gray16, ctx->dimg.p, (size_t)n * 2, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->prof) prof_collect(ctx);
    return FGL_OK;
}

// fgl_depth_image's kernels with the min/max partials in the library's
// per-device scratch and the image written to the caller's device buffer:
// no allocation, no event, no synchronisation.  A pending clear is applied
// by the fill kernels first, as every reader of the buffers does.
int fgl_depth_image_noalloc_async(fgl_ctx *ctx, uint16_t *gray16_device) {
    if (check_ctx(ctx)) return FGL_E_INVALID;
    if (!gray16_device) {
        ctx->err = "fgl_depth_image_noalloc_async: null buffer";
        return FGL_E_INVALID;
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    double *part = nullptr;
    HIPCHK(ctx, depth_scratch(ctx->device, &part));
    const int64_t n = (int64_t)ctx->W * ctx->H;
    // one partial (min, max) per block of k_depth_minmax, which strides over
    // any n: the scratch bounds the block count, not the image size
    const int32_t nb = 512;
    if (nb > kDepthScratchBlocks) {
        set_err(ctx->err, "fgl_depth_image_noalloc_async: %d blocks exceed the %d-block scratch", nb,
                kDepthScratchBlocks);
        return FGL_E_UNSUPPORTED;
    }
    if (ctx->pend_color) {
        HIPCHK(ctx, launch_fill_color(ctx->color.as<uint32_t>(), n, ctx->clear_color, ctx->stream));
        ctx->pend_color = false;
    }
    if (ctx->pend_depth) {
        HIPCHK(ctx, launch_fill_depth(ctx->depth.as<double>(), n, ctx->clear_depth, ctx->stream));
        ctx->pend_depth = false;
    }
    if (ctx->cblk_any) {
        HIPCHK(ctx, launch_cblk_fill(ctx->color.as<uint32_t>(), ctx->depth.as<double>(), ctx->W, ctx->H,
                                     ctx->cblk.as<uint8_t>(), ctx->cbw, ctx->cbh, ctx->clear_color, ctx->clear_depth,
                                     CB_COLOR | CB_DEPTH, ctx->stream));
        ctx->cblk_any = false;
    }
    HIPCHK(ctx, launch_depth_minmax(ctx->depth.as<double>(), n, part, nb, ctx->stream));
    HIPCHK(ctx, launch_depth_image(ctx->depth.as<double>(), n, part, nb, gray16_device, ctx->stream));
    return FGL_OK;
}

// Tent weights: the linear kernel widened by the scale factor, 8-bit fixed
// point, as the Bilinear resize the reference pipeline applies after
// rendering (README.md:109).
static void tent_weights(int32_t src, int32_t dst, std::vector<int16_t> &w, std::vector<int32_t> &off, int32_t &fl) {
    const double blur = 1.0;
    double scale = (double)src / (double)dst;
    fl = 2 * (int32_t)std::max(std::ceil(blur * scale), 1.0);
    double factor = std::min(1.0 / (blur * scale), 1.0);
    w.assign((size_t)dst * fl, 0);
    off.assign((size_t)dst, 0);
    for (int32_t y = 0; y < dst; y++) {
        double interp = scale * ((double)y + 0.5) - 0.5;
        int32_t start = (int32_t)interp - fl / 2 + 1;
        off[y] = start;
        interp -= (double)start;
        for (int32_t i = 0; i < fl; i++) {
            double in = (interp - (double)i) * factor;
            double a = std::fabs(in);
            double k = a <= 1 ? 1 - a : 0;
            w[(size_t)y * fl + i] = (int16_t)(k * 256);
        }
    }
}

// The packed form of one axis's tent table when the factor is 4 and every
// output has the same eight weights in 0..255, a power-of-two sum and offset
// 4 * i - 2.  pre_zeros zero weights come before the eight in the packing.
static bool tent4_pack(const std::vector<int16_t> &w, const std::vector<int32_t> &off, int32_t fl, int pre_zeros,
                       uint32_t *packed, int npacked, int32_t *shift) {
    if (fl != 8 || off.empty()) return false;
    int32_t sum = 0;
    for (int k = 0; k < 8; k++) {
        if (w[k] < 0 || w[k] > 255) return false;
        sum += w[k];
    }
    if (sum <= 0 || (sum & (sum - 1)) != 0) return false;
    for (size_t i = 0; i < off.size(); i++) {
        if (off[i] != 4 * (int32_t)i - 2) return false;
        for (int k = 0; k < 8; k++)
            if (w[i * 8 + k] != w[k]) return false;
    }
    for (int d = 0; d < npacked; d++) packed[d] = 0;
    for (int k = 0; k < 8; k++) {
        const int b = pre_zeros + k;
        if (b / 4 >= npacked) return false;
        packed[b / 4] |= (uint32_t)w[k] << (8 * (b % 4));
    }
    *shift = __builtin_ctz((uint32_t)sum);
    return true;
}

int fgl_resolve(fgl_ctx *src, fgl_ctx *dst, int32_t filter) {
    if (check_ctx(src) || check_ctx(dst)) return FGL_E_INVALID;
    if (src->device != dst->device) {
        src->err = "fgl_resolve: contexts on different devices";
        return FGL_E_INVALID;
    }
    if (dst->W <= 0 || src->W % dst->W || src->H % dst->H || src->W / dst->W != src->H / dst->H) {
        set_err(src->err, "fgl_resolve: %dx%d is not an integer multiple of %dx%d", src->W, src->H, dst->W, dst->H);
        return FGL_E_INVALID;
    }
    if (filter != FGL_RESOLVE_TENT && filter != FGL_RESOLVE_BOX) {
        src->err = "fgl_resolve: unknown filter";
        return FGL_E_INVALID;
    }
    HIPCHK(src, hipSetDevice(src->device));
    int r = FGL_OK;
    // dst's earlier work finishes before the resolve overwrites its color
    if ((r = stream_after(src, dst))) return r;
    int32_t flx = 0, fly = 0;
    if (filter == FGL_RESOLVE_TENT) {
        const size_t fx = 2 * (size_t)(src->W / dst->W), fy = 2 * (size_t)(src->H / dst->H);
        flx = (int32_t)fx;
        fly = (int32_t)fy;
        const int32_t key[4] = {src->W, src->H, dst->W, dst->H};
        if (std::memcmp(key, src->rs_key, sizeof key) != 0) {
            std::vector<int16_t> wx, wy;
            std::vector<int32_t> ox, oy;
            tent_weights(src->W, dst->W, wx, ox, flx);
            tent_weights(src->H, dst->H, wy, oy, fly);
            HIPCHK(src, src->rs_wx.ensure(wx.size() * 2));
            HIPCHK(src, src->rs_ox.ensure(ox.size() * 4));
            HIPCHK(src, src->rs_wy.ensure(wy.size() * 2));
            HIPCHK(src, src->rs_oy.ensure(oy.size() * 4));
            HIPCHK(src, hipMemcpyAsync(src->rs_wx.p, wx.data(), wx.size() * 2, hipMemcpyHostToDevice, src->stream));
            HIPCHK(src, hipMemcpyAsync(src->rs_ox.p, ox.data(), ox.size() * 4, hipMemcpyHostToDevice, src->stream));
            HIPCHK(src, hipMemcpyAsync(src->rs_wy.p, wy.data(), wy.size() * 2, hipMemcpyHostToDevice, src->stream));
            HIPCHK(src, hipMemcpyAsync(src->rs_oy.p, oy.data(), oy.size() * 4, hipMemcpyHostToDevice, src->stream));
            // the host vectors go out of scope: finish the copies first
            HIPCHK(src, hipStreamSynchronize(src->stream));
            std::memcpy(src->rs_key, key, sizeof key);
            ResolveTent4 t4{};
            t4.ok = tent4_pack(wx, ox, flx, 2, t4.wx, 3, &t4.shx) && tent4_pack(wy, oy, fly, 0, t4.wy, 2, &t4.shy);
            src->rs_t4 = t4;
        }
        HIPCHK(src, src->rs_tmp.ensure((size_t)dst->W * src->H * 4));
    }
    // the 4x tent reads the source's marks of deferred clears itself; the
    // other filters read a materialised color buffer
    const bool marks = resolve_reads_marks(filter, src->W, src->H, dst->W, dst->H, src->rs_t4);
    if (marks && src->pend_color) r = materialize(src, CB_COLOR);
    if (!marks) r = materialize(src, CB_COLOR);
    if (r) return r;
    const uint8_t *cb = marks && src->cblk_any ? src->cblk.as<uint8_t>() : nullptr;
    r = launch(src, "resolve", [&] {
        return launch_resolve(src->color.as<uint32_t>(), src->W, src->H, src->rs_tmp.as<uint32_t>(),
                              dst->color.as<uint32_t>(), dst->W, dst->H, filter, src->rs_wx.as<int16_t>(),
                              src->rs_ox.as<int32_t>(), flx, src->rs_wy.as<int16_t>(), src->rs_oy.as<int32_t>(), fly,
                              src->rs_t4, cb, src->cbw, src->clear_color, src->stream);
    });
    if (r) return r;
    // later work on dst (reads, draws) runs after the resolve
    if ((r = stream_after(dst, src))) return r;
    if ((r = cblk_drop(dst, CB_COLOR))) return r;
    if (src->prof) prof_collect(src);
    dst->pend_color = false;
    return FGL_OK;
}

int fgl_profile_enable(fgl_ctx *ctx, int32_t enable) {
    if (check_ctx(ctx)) return FGL_E_INVALID;
    ctx->prof = enable != 0;
    return FGL_OK;
}

int fgl_profile_only(fgl_ctx *ctx, const char *kernel) {
    if (check_ctx(ctx)) return FGL_E_INVALID;
    ctx->prof_only = kernel ? kernel : "";
    return FGL_OK;
}

int fgl_profile_read(fgl_ctx *ctx, fgl_kernel_time *out, int32_t cap, int32_t *n) {
    if (check_ctx(ctx)) return FGL_E_INVALID;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    prof_collect(ctx);
    int32_t k = 0;
    for (auto &kv : ctx->kstats) {
        if (out && k < cap) {
            std::memset(out[k].name, 0, sizeof out[k].name);
            std::strncpy(out[k].name, kv.first.c_str(), sizeof out[k].name - 1);
            out[k].launches = kv.second.launches;
            out[k].total_ms = kv.second.ms;
        }
        k++;
    }
    if (n) *n = k;
    return FGL_OK;
}

int fgl_profile_reset(fgl_ctx *ctx) {
    if (check_ctx(ctx)) return FGL_E_INVALID;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    prof_collect(ctx);
    ctx->kstats.clear();
    return FGL_OK;
}

int fgl_draw_stats(fgl_ctx *ctx, fgl_draw_stat *out) {
    if (check_ctx(ctx) || !out) return FGL_E_INVALID;
    const DrawCounters &h = ctx->last;
    out->total_pixels = h.total;
    out->updated_pixels = updated_pixels(h);
    out->clipped_triangles = h.ndeferred;
    out->side_triangles = h.side_rec;
    out->pairs = (uint64_t)h.pairs_hi << 32 | h.pairs_lo;
    out->spans = (uint64_t)h.spans_hi << 32 | h.spans_lo;
    out->fragments = (uint64_t)h.frags_hi << 32 | h.frags_lo;
    out->overflow_mask = ctx->last_overflow;
    out->reruns = ctx->last_reruns;
    return FGL_OK;
}

int fgl_ctx_force_capacities(fgl_ctx *ctx, const uint32_t caps[5]) {
    if (check_ctx(ctx)) return FGL_E_INVALID;
    if (!caps) {
        ctx->err = "fgl_ctx_force_capacities: null capacities";
        return FGL_E_INVALID;
    }
    std::memcpy(ctx->force_caps, caps, sizeof ctx->force_caps);
    return FGL_OK;
}

int fgl_ctx_sync(fgl_ctx *ctx) {
    if (check_ctx(ctx)) return FGL_E_INVALID;
    HIPCHK(ctx, wait_stream(ctx->stream));
    if (ctx->prof) prof_collect(ctx);
    return FGL_OK;
}

}  // extern "C"

// ===========================================================================
// Learned anti-aliasing: parameters, forward, backward and Adam on the
// context's device (fgl_nn.hip).

constexpr int32_t kNnRangeBlocks = 2048;  // depth min / max partials (enough blocks to fill the GPU)

struct fgl_nn {
    fgl_ctx *home = nullptr;  // device owner (parameters live with it)
    int device = -1;
    int32_t cin[NN_LAYERS] = {NN_FEAT_C, NN_C, NN_C, NN_C, NN_C, NN_C};
    int32_t cout[NN_LAYERS] = {NN_C, NN_C, NN_C, NN_C, NN_C, NN_OUT_C};
    int64_t woff[NN_LAYERS] = {}, boff[NN_LAYERS] = {};
    int64_t nparam = 0;
    DevBuf p, g, m, v;                        // f32 master weights, gradient, Adam moments
    DevBuf wf[NN_LAYERS], wt[NN_LAYERS];      // bf16 forward / flipped-transposed copies
    DevBuf feat, act[NN_LAYERS - 1], out, gout, dy[2], dwp, dbp, rtmp, lpart, range, dpart;
    int64_t step = 0;
    int32_t H = 0, W = 0;
};

namespace {

uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
double gauss(uint64_t &s) {
    const double u1 = ((double)(splitmix(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    const double u2 = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

int nn_pack(fgl_nn *nn) {
    fgl_ctx *c = nn->home;
    for (int l = 0; l < NN_LAYERS; l++) {
        int r = launch(c, "nn_pack", [&] {
            return launch_nn_pack(nn->p.as<float>() + nn->woff[l], nn->cin[l], nn->cout[l], nn->wf[l].as<uint16_t>(),
                                  nn->wt[l].as<uint16_t>(), c->stream);
        });
        if (r) return r;
    }
    return FGL_OK;
}

int nn_buffers(fgl_nn *nn, int32_t H, int32_t W) {
    fgl_ctx *c = nn->home;
    const size_t px = (size_t)H * W;
    HIPCHK(c, nn->feat.ensure(px * NN_FEAT_C * 2));
    for (int l = 0; l < NN_LAYERS - 1; l++) HIPCHK(c, nn->act[l].ensure(px * NN_C * 2));
    HIPCHK(c, nn->out.ensure(px * NN_OUT_C * 4));
    HIPCHK(c, nn->range.ensure(16));
    HIPCHK(c, nn->dpart.ensure(kNnRangeBlocks * 16));
    nn->H = H;
    nn->W = W;
    return FGL_OK;
}

// features and the six convolutions of src's color and depth; the last
// layer writes the raw output (nn->out) or, with dst, the final color.
// keep_feat: the features go through nn->feat (the weight gradient of
// layer 0 reads them); otherwise layer 0 computes them as it stages tiles.
// FGL_NN_PAIRS=1 runs inference as three launches of two fused layers each
// (the earlier kernel), for comparison with the one-launch stream
bool nn_pairs() {
    static const bool v = [] {
        const char *e = std::getenv("FGL_NN_PAIRS");
        return e && e[0] == '1';
    }();
    return v;
}

int nn_forward(fgl_nn *nn, fgl_ctx *src, fgl_ctx *dst, bool keep_feat) {
    fgl_ctx *c = nn->home;
    int r = materialize(src);
    if (r) return r;
    if ((r = stream_after(c, src))) return r;
    if ((r = nn_buffers(nn, src->H, src->W))) return r;
    const int32_t H = src->H, W = src->W;
    const int64_t n = (int64_t)H * W;
    if ((r = launch(c, "nn_depth_minmax", [&] {
             return launch_depth_minmax(src->depth.as<double>(), n, nn->dpart.as<double>(), kNnRangeBlocks,
                                       c->stream);
         })))
        return r;
    // the fused first pair of inference reduces the partials itself
    if (keep_feat && (r = launch(c, "nn_depth_range", [&] {
                          return launch_nn_range(nn->dpart.as<double>(), kNnRangeBlocks, nn->range.as<double>(),
                                                 c->stream);
                      })))
        return r;
    if (keep_feat && (r = launch(c, "nn_features", [&] {
                          return launch_nn_features(src->color.as<uint32_t>(), src->depth.as<double>(),
                                                    nn->range.as<double>(), H, W, nn->feat.as<uint16_t>(), c->stream);
                      })))
        return r;
    if (dst && (r = stream_after(c, dst))) return r;
    if (!keep_feat && !nn_pairs()) {
        // inference: all six layers in one launch, streamed down strips with
        // every activation in LDS
        const uint16_t *w[NN_LAYERS];
        const float *b[NN_LAYERS];
        for (int l = 0; l < NN_LAYERS; l++) {
            w[l] = nn->wf[l].as<uint16_t>();
            b[l] = nn->p.as<float>() + nn->boff[l];
        }
        return launch(c, "nn_stream", [&] {
            return launch_nn_stream(dst ? 3 : 1, src->color.as<uint32_t>(), src->depth.as<double>(),
                                    nn->dpart.as<double>(), kNnRangeBlocks, w, b, H, W,
                                    dst ? dst->color.p : nn->out.p, c->stream);
        });
    }
    if (!keep_feat) {
        // inference: layers fused in pairs, the odd layers' inputs stay in LDS
        static const char *const kPairName[NN_LAYERS / 2] = {"nn_conv01", "nn_conv23", "nn_conv45"};
        const uint16_t *pin = nullptr;  // pair 0 computes the features itself
        for (int l = 0; l < NN_LAYERS; l += 2) {
            const bool last = l + 2 == NN_LAYERS;
            void *o = last ? (dst ? dst->color.p : nn->out.p) : nn->act[l + 1].p;
            const int epi = last ? (dst ? 3 : 1) : 0;
            if ((r = launch(c, kPairName[l / 2], [&] {
                     return launch_nn_conv2(nn->cin[l], nn->cout[l + 1], epi, pin, nn->wf[l].as<uint16_t>(),
                                            nn->p.as<float>() + nn->boff[l], nn->wf[l + 1].as<uint16_t>(),
                                            nn->p.as<float>() + nn->boff[l + 1], H, W, o, src->color.as<uint32_t>(),
                                            src->depth.as<double>(), nn->dpart.as<double>(), kNnRangeBlocks,
                                            c->stream);
                 })))
                return r;
            pin = nn->act[l + 1].as<uint16_t>();
        }
        return FGL_OK;
    }
    const uint16_t *in = nn->feat.as<uint16_t>();
    for (int l = 0; l < NN_LAYERS; l++) {
        const bool last = l == NN_LAYERS - 1;
        void *o = last ? (dst ? dst->color.p : nn->out.p) : nn->act[l].p;
        const int epi = last ? (dst ? 3 : 1) : 0;
        static const char *const kConvName[NN_LAYERS] = {"nn_conv0", "nn_conv1", "nn_conv2",
                                                          "nn_conv3", "nn_conv4", "nn_conv5"};
        if ((r = launch(c, kConvName[l], [&] {
                 return launch_nn_conv(epi, in, nn->cin[l], nn->wf[l].as<uint16_t>(), nn->p.as<float>() + nn->boff[l],
                                       H, W, nn->cout[l], o, nullptr, src->color.as<uint32_t>(),
                                       src->depth.as<double>(), nn->range.as<double>(), c->stream);
             })))
            return r;
        if (!last) in = nn->act[l].as<uint16_t>();
    }
    return FGL_OK;
}

// loss against target's color, gradient of every parameter into nn->g
int nn_backward(fgl_nn *nn, fgl_ctx *src, fgl_ctx *target, float *loss) {
    fgl_ctx *c = nn->home;
    int r = materialize(target, CB_COLOR);
    if (r) return r;
    if ((r = stream_after(c, target))) return r;
    const int32_t H = nn->H, W = nn->W;
    const size_t px = (size_t)H * W;
    const int32_t nlb = 256;
    HIPCHK(c, nn->gout.ensure(px * NN_OUT_C * 2));
    HIPCHK(c, nn->dy[0].ensure(px * NN_C * 2));
    HIPCHK(c, nn->dy[1].ensure(px * NN_C * 2));
    HIPCHK(c, nn->lpart.ensure(nlb * 4));
    const int64_t nw = nn_dw_waves(H, W);
    HIPCHK(c, nn->dwp.ensure((size_t)nw * 9 * NN_C * NN_C * 4));
    HIPCHK(c, nn->dbp.ensure((size_t)nw * NN_C * 4));
    HIPCHK(c, nn->rtmp.ensure((size_t)NN_REDUCE_SPLIT * 9 * NN_C * NN_C * 4));
    if ((r = launch(c, "nn_loss", [&] {
             return launch_nn_loss(nn->out.as<float>(), src->color.as<uint32_t>(), target->color.as<uint32_t>(), H, W,
                                   nn->gout.as<uint16_t>(), nn->lpart.as<float>(), nlb, c->stream);
         })))
        return r;
    const uint16_t *dy = nn->gout.as<uint16_t>();
    for (int l = NN_LAYERS - 1; l >= 0; l--) {
        const uint16_t *x = l == 0 ? nn->feat.as<uint16_t>() : nn->act[l - 1].as<uint16_t>();
        const int32_t ci = nn->cin[l], co = nn->cout[l];
        if ((r = launch(c, "nn_dw", [&] {
                 return launch_nn_dw(x, ci, dy, co, H, W, nn->dwp.as<float>(), nn->dbp.as<float>(), c->stream);
             })))
            return r;
        if ((r = launch(c, "nn_reduce", [&] {
                 return launch_nn_reduce(nn->dwp.as<float>(), nw, 9LL * co * ci, nn->g.as<float>() + nn->woff[l],
                                         nn->rtmp.as<float>(), c->stream);
             })))
            return r;
        if ((r = launch(c, "nn_reduce", [&] {
                 return launch_nn_reduce(nn->dbp.as<float>(), nw, co, nn->g.as<float>() + nn->boff[l],
                                         nn->rtmp.as<float>(), c->stream);
             })))
            return r;
        if (l == 0) break;
        uint16_t *dx = nn->dy[l & 1].as<uint16_t>();
        if ((r = launch(c, "nn_conv_dx", [&] {
                 return launch_nn_conv(2, dy, co, nn->wt[l].as<uint16_t>(), nullptr, H, W, ci, dx, x, nullptr,
                                       nullptr, nullptr, c->stream);
             })))
            return r;
        dy = dx;
    }
    std::vector<float> lp(nlb);
    HIPCHK(c, hipMemcpyAsync(lp.data(), nn->lpart.p, nlb * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->prof) prof_collect(c);
    double s = 0;
    for (float x : lp) s += x;
    if (loss) *loss = (float)s;
    return FGL_OK;
}

int check_nn(fgl_nn *nn) {
    if (!nn || !nn->home) {
        g_err = "null network";
        return FGL_E_INVALID;
    }
    return FGL_OK;
}

int check_same_size(fgl_ctx *a, fgl_ctx *b, const char *what) {
    if (a->W != b->W || a->H != b->H) {
        set_err(a->err, "%s: contexts differ in size (%dx%d vs %dx%d)", what, a->W, a->H, b->W, b->H);
        return FGL_E_INVALID;
    }
    if (a->device != b->device) {
        set_err(a->err, "%s: contexts on different devices", what);
        return FGL_E_INVALID;
    }
    return FGL_OK;
}

}  // namespace

extern "C" {

int fgl_nn_create(fgl_ctx *ctx, uint64_t seed, fgl_nn **out) {
    if (check_ctx(ctx) || !out) return FGL_E_INVALID;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    fgl_nn *nn = new fgl_nn;
    nn->home = ctx;
    nn->device = ctx->device;
    int64_t off = 0;
    for (int l = 0; l < NN_LAYERS; l++) {
        nn->woff[l] = off;
        off += 9LL * nn->cin[l] * nn->cout[l];
        nn->boff[l] = off;
        off += nn->cout[l];
    }
    nn->nparam = off;
    std::vector<float> h((size_t)off, 0.0f);
    uint64_t s = seed;
    for (int l = 0; l < NN_LAYERS; l++) {
        const int real_in = l == 0 ? 8 : nn->cin[l];
        const int real_out = l == NN_LAYERS - 1 ? 3 : nn->cout[l];
        const double sd = std::sqrt(2.0 / (9.0 * real_in)) * (l == NN_LAYERS - 1 ? 0.1 : 1.0);
        for (int tap = 0; tap < 9; tap++)
            for (int co = 0; co < nn->cout[l]; co++)
                for (int ci = 0; ci < nn->cin[l]; ci++) {
                    const double gv = gauss(s);
                    if (co < real_out && ci < real_in)
                        h[(size_t)(nn->woff[l] + ((int64_t)tap * nn->cout[l] + co) * nn->cin[l] + ci)] = (float)(gv * sd);
                }
    }
    int r = FGL_OK;
    auto fail = [&](const char *m) {
        set_err(ctx->err, "fgl_nn_create: %s", m);
        delete nn;
        return FGL_E_NOMEM;
    };
    if (nn->p.ensure(off * 4) || nn->g.ensure(off * 4) || nn->m.ensure(off * 4) || nn->v.ensure(off * 4))
        return fail("device allocation failed");
    for (int l = 0; l < NN_LAYERS; l++)
        if (nn->wf[l].ensure(9LL * nn->cin[l] * nn->cout[l] * 2) || nn->wt[l].ensure(9LL * nn->cin[l] * nn->cout[l] * 2))
            return fail("device allocation failed");
    if (hipMemcpyAsync(nn->p.p, h.data(), off * 4, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipMemsetAsync(nn->g.p, 0, off * 4, ctx->stream) != hipSuccess ||
        hipMemsetAsync(nn->m.p, 0, off * 4, ctx->stream) != hipSuccess ||
        hipMemsetAsync(nn->v.p, 0, off * 4, ctx->stream) != hipSuccess) {
        set_err(ctx->err, "fgl_nn_create: upload failed");
        delete nn;
        return FGL_E_DEVICE;
    }
    if ((r = nn_pack(nn))) {
        delete nn;
        return r;
    }
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
        set_err(ctx->err, "fgl_nn_create: synchronize failed");
        delete nn;
        return FGL_E_DEVICE;
    }
    *out = nn;
    return FGL_OK;
}

void fgl_nn_destroy(fgl_nn *nn) {
    if (!nn) return;
    // as fgl_mesh_destroy: nothing read through a home context that may be gone
    (void)hipSetDevice(nn->device);
    (void)hipDeviceSynchronize();
    delete nn;
    (void)hipGetLastError();
}

int fgl_nn_param_count(const fgl_nn *nn, int64_t *n) {
    if (!nn || !n) return FGL_E_INVALID;
    *n = nn->nparam;
    return FGL_OK;
}

int fgl_nn_get_params(fgl_nn *nn, float *host, int64_t n) {
    if (check_nn(nn) || !host || n != nn->nparam) return FGL_E_INVALID;
    fgl_ctx *c = nn->home;
    HIPCHK(c, hipMemcpyAsync(host, nn->p.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FGL_OK;
}

int fgl_nn_set_params(fgl_nn *nn, const float *host, int64_t n) {
    if (check_nn(nn) || !host || n != nn->nparam) return FGL_E_INVALID;
    fgl_ctx *c = nn->home;
    HIPCHK(c, hipMemcpyAsync(nn->p.p, host, n * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(nn->m.p, 0, n * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(nn->v.p, 0, n * 4, c->stream));
    nn->step = 0;
    int r = nn_pack(nn);
    if (r) return r;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FGL_OK;
}

int fgl_nn_infer(fgl_nn *nn, fgl_ctx *src, fgl_ctx *dst) {
    if (check_nn(nn) || check_ctx(src) || check_ctx(dst)) return FGL_E_INVALID;
    fgl_ctx *c = nn->home;
    int r = check_same_size(src, dst, "fgl_nn_infer");
    if (r) return r;
    if (src->device != c->device) {
        set_err(src->err, "fgl_nn_infer: context and network on different devices");
        return FGL_E_INVALID;
    }
    HIPCHK(c, hipSetDevice(c->device));
    if ((r = nn_forward(nn, src, dst, false))) return r;
    if ((r = stream_after(dst, c))) return r;
    if ((r = stream_after(src, c))) return r;
    if ((r = cblk_drop(dst, CB_COLOR))) return r;
    dst->pend_color = false;
    return FGL_OK;
}

int fgl_nn_residual(fgl_nn *nn, fgl_ctx *src, float *host, int64_t n) {
    if (check_nn(nn) || check_ctx(src) || !host) return FGL_E_INVALID;
    fgl_ctx *c = nn->home;
    if (src->device != c->device || n != (int64_t)src->W * src->H * NN_OUT_C) {
        set_err(src->err, "fgl_nn_residual: needs %lld floats on the network's device",
                (long long)src->W * src->H * NN_OUT_C);
        return FGL_E_INVALID;
    }
    HIPCHK(c, hipSetDevice(c->device));
    int r = nn_forward(nn, src, nullptr, false);
    if (r) return r;
    HIPCHK(c, hipMemcpyAsync(host, nn->out.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FGL_OK;
}

int fgl_nn_grad(fgl_nn *nn, fgl_ctx *src, fgl_ctx *target, float *loss, float *grad, int64_t n) {
    if (check_nn(nn) || check_ctx(src) || check_ctx(target)) return FGL_E_INVALID;
    int r = check_same_size(src, target, "fgl_nn_grad");
    if (r) return r;
    fgl_ctx *c = nn->home;
    if (src->device != c->device) {
        set_err(src->err, "fgl_nn_grad: context and network on different devices");
        return FGL_E_INVALID;
    }
    if (grad && n != nn->nparam) {
        set_err(src->err, "fgl_nn_grad: gradient buffer holds %lld floats, the network %lld", (long long)n,
                (long long)nn->nparam);
        return FGL_E_INVALID;
    }
    HIPCHK(c, hipSetDevice(c->device));
    if ((r = nn_forward(nn, src, nullptr, true))) return r;
    if ((r = nn_backward(nn, src, target, loss))) return r;
    if (grad) {
        HIPCHK(c, hipMemcpyAsync(grad, nn->g.p, n * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return FGL_OK;
}

int fgl_nn_train_step(fgl_nn *nn, fgl_ctx *src, fgl_ctx *target, float lr, float *loss) {
    int r = fgl_nn_grad(nn, src, target, loss, nullptr, 0);
    if (r) return r;
    fgl_ctx *c = nn->home;
    nn->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - (float)std::pow((double)b1, (double)nn->step);
    const float c2 = 1.0f - (float)std::pow((double)b2, (double)nn->step);
    if ((r = launch(c, "nn_adam", [&] {
             return launch_nn_adam(nn->p.as<float>(), nn->g.as<float>(), nn->m.as<float>(), nn->v.as<float>(),
                                   nn->nparam, lr, b1, b2, 1e-8f, c1, c2, c->stream);
         })))
        return r;
    return nn_pack(nn);
}

}  // extern "C"

// ===========================================================================
// Device-side mesh preparation (fgl_mesh.hip): BoundingBox, Transform,
// FitInside, SmoothNormals, SmoothNormalsThreshold on a resident mesh.

namespace {

double hgo_min(double x, double y) {
    if ((std::isinf(x) && x < 0) || (std::isinf(y) && y < 0)) return -INFINITY;
    if (std::isnan(x) || std::isnan(y)) return NAN;
    if (x == 0 && x == y) return std::signbit(x) ? x : y;
    return x < y ? x : y;
}
double hgo_max(double x, double y) {
    if ((std::isinf(x) && x > 0) || (std::isinf(y) && y > 0)) return INFINITY;
    if (std::isnan(x) || std::isnan(y)) return NAN;
    if (x == 0 && x == y) return std::signbit(x) ? y : x;
    return x > y ? x : y;
}

// Matrix.Mul (matrix.go:188-207), row-major X00..X33
void mat_mul(const double a[16], const double b[16], double m[16]) {
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            m[4 * i + j] = a[4 * i] * b[j] + a[4 * i + 1] * b[4 + j] + a[4 * i + 2] * b[8 + j] + a[4 * i + 3] * b[12 + j];
}

int check_mesh(fgl_mesh *m) {
    if (!m || !m->ctx) {
        g_err = "null mesh";
        return FGL_E_INVALID;
    }
    return FGL_OK;
}

constexpr int32_t kMeshRedBlocks = 1024;

// index of the first primitive with a coordinate != 0, or -1
int mesh_first_nonzero(fgl_ctx *c, const double *pos, int64_t nprim, int32_t nv, int64_t *first) {
    *first = -1;
    if (nprim == 0) return FGL_OK;
    const int32_t nb = (int32_t)std::min<int64_t>(kMeshRedBlocks, (nprim + 255) / 256);
    DevBuf part;
    HIPCHK(c, part.ensure((size_t)nb * 8));
    int r = launch(c, "mesh_box", [&] {
        return launch_mesh_first_nonzero(pos, nprim, nv, part.as<int64_t>(), nb, c->stream);
    });
    if (r) return r;
    std::vector<int64_t> h(nb);
    HIPCHK(c, hipMemcpyAsync(h.data(), part.p, (size_t)nb * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int64_t f = INT64_MAX;
    for (int64_t x : h) f = std::min(f, x);
    if (f != INT64_MAX) *first = f;
    return FGL_OK;
}

// math.Min / Max of every coordinate of primitives p >= from into box
int mesh_box_from(fgl_ctx *c, const double *pos, int64_t nprim, int32_t nv, int64_t from, double box[6]) {
    const int32_t nb = (int32_t)std::min<int64_t>(kMeshRedBlocks, ((nprim - from) * nv + 255) / 256);
    DevBuf part;
    HIPCHK(c, part.ensure((size_t)nb * 6 * 8));
    int r = launch(c, "mesh_box", [&] {
        return launch_mesh_box(pos, nprim, nv, from, part.as<double>(), nb, c->stream);
    });
    if (r) return r;
    std::vector<double> h((size_t)nb * 6);
    HIPCHK(c, hipMemcpyAsync(h.data(), part.p, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int b = 0; b < nb; b++)
        for (int k = 0; k < 3; k++) {
            box[k] = hgo_min(box[k], h[6 * b + k]);
            box[3 + k] = hgo_max(box[3 + k], h[6 * b + 3 + k]);
        }
    return FGL_OK;
}

// Mesh.BoundingBox (mesh.go:153-165): triangles, then lines, each folded
// with Box.Extend (box.go:48-53)
int mesh_bbox(fgl_mesh *m, double box[6]) {
    fgl_ctx *c = m->ctx;
    for (int k = 0; k < 3; k++) {
        box[k] = INFINITY;
        box[3 + k] = -INFINITY;
    }
    int64_t ft = -1, fl = -1;
    int r;
    if (m->t_has[0] && (r = mesh_first_nonzero(c, m->tpos.as<double>(), m->ntri, 3, &ft))) return r;
    if (ft >= 0) {
        if ((r = mesh_box_from(c, m->tpos.as<double>(), m->ntri, 3, ft, box))) return r;
        if (m->l_has[0] && m->nline > 0 && (r = mesh_box_from(c, m->lpos.as<double>(), m->nline, 2, 0, box)))
            return r;
        return FGL_OK;
    }
    if (m->l_has[0] && (r = mesh_first_nonzero(c, m->lpos.as<double>(), m->nline, 2, &fl))) return r;
    if (fl >= 0) return mesh_box_from(c, m->lpos.as<double>(), m->nline, 2, fl, box);
    for (int k = 0; k < 6; k++) box[k] = 0.0;  // EmptyBox
    return FGL_OK;
}

// a zero-filled array for an attribute the mesh was created without
int mesh_materialize(fgl_ctx *c, DevBuf &b, bool &has, int64_t ndoubles) {
    if (has || ndoubles == 0) return FGL_OK;
    HIPCHK(c, b.ensure((size_t)ndoubles * 8));
    HIPCHK(c, hipMemsetAsync(b.p, 0, (size_t)ndoubles * 8, c->stream));
    has = true;
    return FGL_OK;
}

int mesh_transform(fgl_mesh *m, const double M[16]) {
    fgl_ctx *c = m->ctx;
    int r;
    const int64_t ntv = m->ntri * 3, nlv = m->nline * 2;
    if ((r = mesh_materialize(c, m->tpos, m->t_has[0], ntv * 3))) return r;
    if ((r = mesh_materialize(c, m->tnrm, m->t_has[1], ntv * 3))) return r;
    if ((r = mesh_materialize(c, m->lpos, m->l_has[0], nlv * 3))) return r;
    if ((r = mesh_materialize(c, m->lnrm, m->l_has[1], nlv * 3))) return r;
    if (ntv > 0 && (r = launch(c, "mesh_transform", [&] {
                        return launch_mesh_transform(m->tpos.as<double>(), m->tnrm.as<double>(), ntv, M, c->stream);
                    })))
        return r;
    if (nlv > 0 && (r = launch(c, "mesh_transform", [&] {
                        return launch_mesh_transform(m->lpos.as<double>(), m->lnrm.as<double>(), nlv, M, c->stream);
                    })))
        return r;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FGL_OK;
}

int mesh_smooth(fgl_mesh *m, bool threshold_mode, double threshold) {
    fgl_ctx *c = m->ctx;
    const int64_t nv = m->ntri * 3;
    if (nv == 0) return FGL_OK;
    if (nv > (int64_t)UINT32_MAX) {
        set_err(c->err, "mesh smoothing: %lld vertices exceed 2^32", (long long)nv);
        return FGL_E_INVALID;
    }
    int r;
    if ((r = mesh_materialize(c, m->tpos, m->t_has[0], nv * 3))) return r;
    if ((r = mesh_materialize(c, m->tnrm, m->t_has[1], nv * 3))) return r;
    const int32_t nb = mesh_sort_blocks(nv);
    DevBuf k0, k1, v0, v1, hist, dtot, where, out;
    HIPCHK(c, k0.ensure((size_t)nv * 4));
    HIPCHK(c, k1.ensure((size_t)nv * 4));
    HIPCHK(c, v0.ensure((size_t)nv * 4));
    HIPCHK(c, v1.ensure((size_t)nv * 4));
    HIPCHK(c, hist.ensure((size_t)256 * nb * 4));
    HIPCHK(c, dtot.ensure(256 * 4));
    HIPCHK(c, where.ensure((size_t)nv * 4));
    HIPCHK(c, out.ensure((size_t)nv * 3 * 8));
    uint32_t *keys[2] = {k0.as<uint32_t>(), k1.as<uint32_t>()};
    uint32_t *vals[2] = {v0.as<uint32_t>(), v1.as<uint32_t>()};
    if ((r = launch(c, "mesh_group", [&] {
             return launch_mesh_group(m->tpos.as<double>(), nv, keys, vals, hist.as<uint32_t>(), dtot.as<uint32_t>(),
                                      where.as<uint32_t>(), c->stream);
         })))
        return r;
    if ((r = launch(c, "mesh_smooth", [&] {
             return launch_mesh_smooth(threshold_mode, m->tpos.as<double>(), m->tnrm.as<double>(), nv, keys[0],
                                       vals[0], where.as<uint32_t>(), threshold, out.as<double>(), c->stream);
         })))
        return r;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::swap(m->tnrm.p, out.p);
    std::swap(m->tnrm.bytes, out.bytes);
    return FGL_OK;
}

}  // namespace

extern "C" {

int fgl_mesh_counts(const fgl_mesh *mesh, int64_t *ntri, int64_t *nline) {
    if (!mesh || !ntri || !nline) return FGL_E_INVALID;
    *ntri = mesh->ntri;
    *nline = mesh->nline;
    return FGL_OK;
}

int fgl_mesh_read(const fgl_mesh *mesh, int32_t lines, int32_t attr, double *out, int64_t n) {
    if (!mesh || !mesh->ctx || !out || attr < 0 || attr > 3) return FGL_E_INVALID;
    fgl_ctx *c = mesh->ctx;
    static const int kWidth[4] = {3, 3, 2, 4};
    const int64_t nvtx = lines ? mesh->nline * 2 : mesh->ntri * 3;
    const int64_t need = nvtx * kWidth[attr];
    if (n != need) {
        set_err(c->err, "fgl_mesh_read: needs %lld doubles", (long long)need);
        return FGL_E_INVALID;
    }
    if (need == 0) return FGL_OK;
    const DevBuf *bufs[2][4] = {{&mesh->tpos, &mesh->tnrm, &mesh->tuv, &mesh->tcol},
                                {&mesh->lpos, &mesh->lnrm, &mesh->luv, &mesh->lcol}};
    const bool has = lines ? mesh->l_has[attr] : mesh->t_has[attr];
    HIPCHK(c, hipSetDevice(c->device));
    if (!has) {
        for (int64_t i = 0; i < need; i++) out[i] = 0.0;
        return FGL_OK;
    }
    HIPCHK(c, hipMemcpyAsync(out, bufs[lines ? 1 : 0][attr]->p, (size_t)need * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FGL_OK;
}

int fgl_mesh_bounding_box(fgl_mesh *mesh, double box[6]) {
    if (check_mesh(mesh) || !box) return FGL_E_INVALID;
    HIPCHK(mesh->ctx, hipSetDevice(mesh->ctx->device));
    return mesh_bbox(mesh, box);
}

int fgl_mesh_transform(fgl_mesh *mesh, const double matrix[16]) {
    if (check_mesh(mesh) || !matrix) return FGL_E_INVALID;
    HIPCHK(mesh->ctx, hipSetDevice(mesh->ctx->device));
    return mesh_transform(mesh, matrix);
}

int fgl_mesh_fit_inside(fgl_mesh *mesh, const double box[6], const double anchor[3], double matrix_out[16]) {
    if (check_mesh(mesh) || !box || !anchor) return FGL_E_INVALID;
    HIPCHK(mesh->ctx, hipSetDevice(mesh->ctx->device));
    double bb[6];
    int r = mesh_bbox(mesh, bb);
    if (r) return r;
    double size[3], bsize[3], extra[3], q[3];
    for (int k = 0; k < 3; k++) {
        size[k] = bb[3 + k] - bb[k];
        bsize[k] = box[3 + k] - box[k];
        q[k] = bsize[k] / size[k];
    }
    const double scale = hgo_min(hgo_min(q[0], q[1]), q[2]);
    for (int k = 0; k < 3; k++) extra[k] = bsize[k] - size[k] * scale;
    const double I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    double T1[16] = {1, 0, 0, -bb[0], 0, 1, 0, -bb[1], 0, 0, 1, -bb[2], 0, 0, 0, 1};
    double S[16] = {scale, 0, 0, 0, 0, scale, 0, 0, 0, 0, scale, 0, 0, 0, 0, 1};
    double T2[16] = {1, 0, 0, box[0] + extra[0] * anchor[0], 0, 1, 0, box[1] + extra[1] * anchor[1],
                     0, 0, 1, box[2] + extra[2] * anchor[2], 0, 0, 0, 1};
    double m1[16], m2[16], m3[16];
    mat_mul(T1, I, m1);
    mat_mul(S, m1, m2);
    mat_mul(T2, m2, m3);
    if ((r = mesh_transform(mesh, m3))) return r;
    if (matrix_out)
        for (int k = 0; k < 16; k++) matrix_out[k] = m3[k];
    return FGL_OK;
}

int fgl_mesh_smooth_normals(fgl_mesh *mesh) {
    if (check_mesh(mesh)) return FGL_E_INVALID;
    HIPCHK(mesh->ctx, hipSetDevice(mesh->ctx->device));
    return mesh_smooth(mesh, false, 0.0);
}

int fgl_mesh_smooth_normals_threshold(fgl_mesh *mesh, double cos_threshold) {
    if (check_mesh(mesh)) return FGL_E_INVALID;
    HIPCHK(mesh->ctx, hipSetDevice(mesh->ctx->device));
    return mesh_smooth(mesh, true, cos_threshold);
}

}  // extern "C"

// ===========================================================================
// Mesh ingest (fgl_ingest.hip): LoadSTL, LoadOBJ into a resident mesh.

namespace {

// a triangle mesh with device positions and normals (and uv when given)
// from host arrays, then FixNormals on the device
int mesh_from_host(fgl_ctx *c, const std::vector<double> &pos, const std::vector<double> &nrm,
                   const std::vector<double> *uv, fgl_mesh **out) {
    const int64_t nt = (int64_t)(pos.size() / 9);
    fgl_prims t{};
    t.count = nt;
    t.position = pos.data();
    t.normal = nrm.data();
    t.texture = uv ? uv->data() : nullptr;
    fgl_mesh *m = nullptr;
    int r = fgl_mesh_create(c, nt ? &t : nullptr, nullptr, &m);
    if (r) return r;
    if (nt > 0 && (r = launch(c, "fix_normals", [&] {
                       return launch_fix_normals(m->tpos.as<double>(), m->tnrm.as<double>(), nt, c->stream);
                   }))) {
        fgl_mesh_destroy(m);
        return r;
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        fgl_mesh_destroy(m);
        set_err(c->err, "mesh ingest: synchronize failed");
        return FGL_E_DEVICE;
    }
    *out = m;
    return FGL_OK;
}

}  // namespace

extern "C" {

int fgl_mesh_load_stl(fgl_ctx *ctx, const uint8_t *data, int64_t size, fgl_mesh **out) {
    if (check_ctx(ctx) || !out || (!data && size > 0) || size < 0) return FGL_E_INVALID;
    *out = nullptr;
    if (size < 84) {
        ctx->err = "fgl_mesh_load_stl: shorter than the 84-byte header";
        return FGL_E_INVALID;
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const uint32_t count = (uint32_t)data[80] | ((uint32_t)data[81] << 8) | ((uint32_t)data[82] << 16) |
                           ((uint32_t)data[83] << 24);
    if ((int64_t)count * 50 + 84 == size) {
        // binary: the bytes go to the device, records decode there
        if ((int64_t)count > kMaxPrims) {
            set_err(ctx->err, "fgl_mesh_load_stl: %u triangles exceed the limit", count);
            return FGL_E_INVALID;
        }
        fgl_mesh *m = new (std::nothrow) fgl_mesh;
        if (!m) return FGL_E_NOMEM;
        m->ctx = ctx;
        m->device = ctx->device;
        m->ntri = count;
        DevBuf file;
        auto fail = [&](int code) {
            delete m;
            return code;
        };
        if (count == 0) {
            *out = m;
            return FGL_OK;
        }
        if (file.ensure((size_t)size) != hipSuccess || m->tpos.ensure((size_t)count * 72) != hipSuccess ||
            m->tnrm.ensure((size_t)count * 72) != hipSuccess) {
            set_err(ctx->err, "fgl_mesh_load_stl: device allocation failed");
            return fail(FGL_E_NOMEM);
        }
        if (hipMemcpy(file.p, data, (size_t)size, hipMemcpyHostToDevice) != hipSuccess) {
            set_err(ctx->err, "fgl_mesh_load_stl: upload failed");
            return fail(FGL_E_DEVICE);
        }
        int r = launch(ctx, "stl_decode", [&] {
            return launch_stl_decode(file.as<uint8_t>(), count, m->tpos.as<double>(), m->tnrm.as<double>(),
                                     ctx->stream);
        });
        if (r) return fail(r);
        if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
            set_err(ctx->err, "fgl_mesh_load_stl: synchronize failed");
            return fail(FGL_E_DEVICE);
        }
        m->t_has[0] = m->t_has[1] = true;
        *out = m;
        return FGL_OK;
    }
    // ASCII (stl.go:59-84): lines of exactly "vertex x y z"
    std::vector<double> pos, nrm;
    std::vector<std::pair<const char *, size_t>> f;
    const char *t = reinterpret_cast<const char *>(data), *end = t + size;
    for (const char *a = t; a < end;) {
        const char *b = static_cast<const char *>(memchr(a, '\n', (size_t)(end - a)));
        if (!b) b = end;
        ingest_fields(a, b, f);
        if (f.size() == 4 && f[0].second == 6 && !memcmp(f[0].first, "vertex", 6))
            for (int k = 1; k < 4; k++) pos.push_back(ingest_parse_float(f[k].first, f[k].second));
        a = b + 1;
    }
    if ((pos.size() / 3) % 3) {
        ctx->err = "fgl_mesh_load_stl: ASCII vertex count is not a multiple of 3";
        return FGL_E_INVALID;
    }
    nrm.assign(pos.size(), 0.0);
    return mesh_from_host(ctx, pos, nrm, nullptr, out);
}

int fgl_mesh_load_obj(fgl_ctx *ctx, const char *text, int64_t size, fgl_mesh **out) {
    if (check_ctx(ctx) || !out || (!text && size > 0) || size < 0) return FGL_E_INVALID;
    *out = nullptr;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    // 1-based lists with a zero Vector at index 0 (obj.go:25-27)
    std::vector<double> vs(3, 0.0), vts(3, 0.0), vns(3, 0.0);
    std::vector<double> pos, nrm, uv;
    std::vector<std::pair<const char *, size_t>> f;
    std::vector<int64_t> idx;
    const char *end = text + size;
    auto bad = [&](const char *msg) {
        set_err(ctx->err, "fgl_mesh_load_obj: %s", msg);
        return FGL_E_INVALID;
    };
    for (const char *a = text; a < end;) {
        const char *b = static_cast<const char *>(memchr(a, '\n', (size_t)(end - a)));
        if (!b) b = end;
        ingest_fields(a, b, f);
        a = b + 1;
        if (f.empty()) continue;
        const char *kw = f[0].first;
        const size_t kl = f[0].second;
        auto num = [&](size_t k) { return ingest_parse_float(f[k].first, f[k].second); };
        if (kl == 1 && kw[0] == 'v') {
            if (f.size() < 4) return bad("a v line with fewer than 3 numbers");
            vs.insert(vs.end(), {num(1), num(2), num(3)});
        } else if (kl == 2 && kw[0] == 'v' && kw[1] == 't') {
            if (f.size() < 3) return bad("a vt line with fewer than 2 numbers");
            vts.insert(vts.end(), {num(1), num(2), 0.0});
        } else if (kl == 2 && kw[0] == 'v' && kw[1] == 'n') {
            if (f.size() < 4) return bad("a vn line with fewer than 3 numbers");
            vns.insert(vns.end(), {num(1), num(2), num(3)});
        } else if (kl == 1 && kw[0] == 'f') {
            const size_t na = f.size() - 1;
            idx.assign(3 * na, 0);
            const int64_t len[3] = {(int64_t)vs.size() / 3, (int64_t)vts.size() / 3, (int64_t)vns.size() / 3};
            for (size_t k = 0; k < na; k++) {
                // strings.Split(arg + "//", "/")[0..2]
                const char *s = f[1 + k].first, *se = s + f[1 + k].second;
                const char *piece = s;
                for (int q = 0; q < 3; q++) {
                    const char *pe = piece;
                    while (pe < se && *pe != '/') ++pe;
                    int64_t v = piece <= se ? ingest_parse_int(piece, (size_t)(pe - piece)) : 0;
                    if (v < 0) v += len[q];
                    idx[3 * k + q] = v;
                    piece = pe < se ? pe + 1 : se + 1;
                }
            }
            for (size_t i = 1; i + 1 < na; i++) {
                const size_t c[3] = {0, i, i + 1};
                for (int q = 0; q < 3; q++)
                    for (int a3 = 0; a3 < 3; a3++) {
                        const int64_t v = idx[3 * c[q] + a3];
                        if (v < 0 || v >= len[a3]) return bad("index out of range");
                    }
                for (int q = 0; q < 3; q++) {
                    const int64_t iv = idx[3 * c[q]], it = idx[3 * c[q] + 1], in = idx[3 * c[q] + 2];
                    pos.insert(pos.end(), {vs[3 * iv], vs[3 * iv + 1], vs[3 * iv + 2]});
                    nrm.insert(nrm.end(), {vns[3 * in], vns[3 * in + 1], vns[3 * in + 2]});
                    uv.insert(uv.end(), {vts[3 * it], vts[3 * it + 1]});
                }
            }
        }
    }
    return mesh_from_host(ctx, pos, nrm, &uv, out);
}

}  // extern "C"

// ===========================================================================
// SavePNG (util.go:56-63): the color buffer encoded on the device (fgl_png.hip)

namespace {

constexpr int32_t kPngChunk = 4096;                              // filtered bytes per deflate piece
constexpr int32_t kPngCapWords = (9 * kPngChunk + 64) / 32 + 2;  // worst case: all 9-bit literals

uint32_t crc_table_entry(uint32_t n) {
    uint32_t c = n;
    for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
    return c;
}
uint32_t crc32_update(uint32_t crc, const uint8_t *p, size_t n) {
    static uint32_t table[256];
    static bool ready = false;
    if (!ready) {
        for (uint32_t i = 0; i < 256; i++) table[i] = crc_table_entry(i);
        ready = true;
    }
    uint32_t c = crc ^ 0xFFFFFFFFu;
    for (size_t i = 0; i < n; i++) c = table[(c ^ p[i]) & 255u] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}
void put_be32(uint8_t *p, uint32_t v) {
    p[0] = (uint8_t)(v >> 24);
    p[1] = (uint8_t)(v >> 16);
    p[2] = (uint8_t)(v >> 8);
    p[3] = (uint8_t)v;
}

int64_t png_bound(const fgl_ctx *c) {
    const int64_t n = (int64_t)c->H * ((int64_t)c->W * 4 + 1);
    const int64_t pieces = (n + kPngChunk - 1) / kPngChunk;
    return 8 + 25 + 8 + 2 + pieces * kPngCapWords * 4 + 2 + 4 + 4 + 12;
}

}  // namespace

extern "C" {

int fgl_encode_png_bound(fgl_ctx *ctx, int64_t *bound) {
    if (check_ctx(ctx) || !bound) return FGL_E_INVALID;
    *bound = png_bound(ctx);
    return FGL_OK;
}

int fgl_encode_png(fgl_ctx *ctx, uint8_t *out, int64_t cap, int64_t *len) {
    if (check_ctx(ctx) || !out || !len) return FGL_E_INVALID;
    if (cap < png_bound(ctx)) {
        set_err(ctx->err, "fgl_encode_png: the buffer needs fgl_encode_png_bound = %lld bytes",
                (long long)png_bound(ctx));
        return FGL_E_INVALID;
    }
    fgl_ctx *c = ctx;
    HIPCHK(c, hipSetDevice(c->device));
    int r = materialize(c, CB_COLOR);
    if (r) return r;
    const int32_t W = c->W, H = c->H;
    DevBuf flag, stream, pieces, sizes, adl, offs, packed;
    HIPCHK(c, flag.ensure(4));
    HIPCHK(c, hipMemsetAsync(flag.p, 0, 4, c->stream));
    if ((r = launch(c, "png_opaque", [&] {
             return launch_png_opaque(c->color.as<uint32_t>(), (int64_t)W * H, flag.as<uint32_t>(), c->stream);
         })))
        return r;
    uint32_t translucent = 0;
    HIPCHK(c, hipMemcpyAsync(&translucent, flag.p, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int32_t bpp = translucent ? 4 : 3;  // image/png: RGBA unless NRGBA.Opaque()
    const int64_t n = (int64_t)H * ((int64_t)W * bpp + 1);
    const int64_t np = (n + kPngChunk - 1) / kPngChunk;
    HIPCHK(c, stream.ensure((size_t)n));
    HIPCHK(c, pieces.ensure((size_t)np * kPngCapWords * 4));
    HIPCHK(c, sizes.ensure((size_t)np * 4));
    HIPCHK(c, adl.ensure((size_t)np * 8));
    HIPCHK(c, offs.ensure((size_t)np * 8));
    if ((r = launch(c, "png_filter", [&] {
             return launch_png_filter(c->color.as<uint32_t>(), W, H, bpp, stream.as<uint8_t>(), c->stream);
         })))
        return r;
    if ((r = launch(c, "png_deflate", [&] {
             return launch_png_deflate(stream.as<uint8_t>(), n, kPngChunk, kPngCapWords, np, pieces.as<uint32_t>(),
                                       sizes.as<uint32_t>(), adl.as<uint32_t>(), c->stream);
         })))
        return r;
    std::vector<uint32_t> hs((size_t)np), ha((size_t)np * 2);
    HIPCHK(c, hipMemcpyAsync(hs.data(), sizes.p, (size_t)np * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(ha.data(), adl.p, (size_t)np * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<uint64_t> ho((size_t)np);
    uint64_t total = 0;
    uint32_t a = 1, b = 0;  // Adler-32 of the whole stream from the pieces'
    for (int64_t k = 0; k < np; k++) {
        ho[k] = total;
        total += hs[k];
        const uint64_t plen = (uint64_t)std::min<int64_t>(kPngChunk, n - k * kPngChunk);
        b = (uint32_t)((b + (plen % 65521u) * a + ha[2 * k + 1]) % 65521u);
        a = (uint32_t)((a + ha[2 * k]) % 65521u);
    }
    HIPCHK(c, packed.ensure((size_t)total));
    HIPCHK(c, hipMemcpyAsync(offs.p, ho.data(), (size_t)np * 8, hipMemcpyHostToDevice, c->stream));
    if ((r = launch(c, "png_pack", [&] {
             return launch_png_pack(pieces.as<uint32_t>(), sizes.as<uint32_t>(), offs.as<uint64_t>(), kPngCapWords, np,
                                    packed.as<uint8_t>(), c->stream);
         })))
        return r;
    // signature, IHDR, then the IDAT chunk around the zlib stream, IEND
    uint8_t *p = out;
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    memcpy(p, sig, 8);
    p += 8;
    put_be32(p, 13);
    memcpy(p + 4, "IHDR", 4);
    put_be32(p + 8, (uint32_t)W);
    put_be32(p + 12, (uint32_t)H);
    p[16] = 8;                         // bit depth
    p[17] = (uint8_t)(bpp == 4 ? 6 : 2);  // color type: truecolor with / without alpha
    p[18] = p[19] = p[20] = 0;         // deflate, adaptive filters, no interlace
    put_be32(p + 21, crc32_update(0, p + 4, 17));
    p += 25;
    const uint64_t zlen = 2 + total + 2 + 4;
    put_be32(p, (uint32_t)zlen);
    memcpy(p + 4, "IDAT", 4);
    uint8_t *z = p + 8;
    z[0] = 0x78;  // deflate, 32K window
    z[1] = 0x9C;  // default level, header check
    HIPCHK(c, hipMemcpyAsync(z + 2, packed.p, (size_t)total, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    z[2 + total] = 0x03;  // final block: fixed codes, end of block only
    z[3 + total] = 0x00;
    put_be32(z + 4 + total, (b << 16) | a);
    put_be32(z + zlen, crc32_update(0, p + 4, (size_t)(4 + zlen)));
    p = z + zlen + 4;
    static const uint8_t iend[12] = {0, 0, 0, 0, 'I', 'E', 'N', 'D', 0xAE, 0x42, 0x60, 0x82};
    memcpy(p, iend, 12);
    p += 12;
    *len = p - out;
    return FGL_OK;
}

}  // extern "C"

// ===========================================================================
// Batched views: many (context, mesh, state, shader) frames rendered by one
// HIP graph.  The first run draws every view eagerly, which sizes each
// context's buffers (the capacity tuning: every capacity is the need of
// that view's frame plus the usual margin), then records all views into one
// graph: a fork from the batch stream to every context's stream, each
// view's clear + triangle + line pipelines on its context's stream (views
// of different contexts are independent branches; views of one context run
// in order), a kernel copying each pipeline's counters into one array, a
// join back.  Later runs replay the graph and read the counters once.
//
// The graph bakes device addresses: every view's context buffers, mesh
// arrays and textures.  It is replayed only while no library allocation has
// been freed or replaced since it was recorded (g_dev_epoch); otherwise the
// run starts over with the eager pass and records again.  A replay first
// waits for the work already queued on every view's context.  When a view's
// frame outgrew its buffers, that view and every later view on the same
// context are redrawn eagerly, in order, and the graph is recorded again.

struct fgl_batch {
    struct View {
        fgl_ctx *ctx;
        const fgl_mesh *mesh;
        fgl_state st;
        fgl_shader sh;
        uint32_t clear_color;
        double clear_depth;
        uint32_t ndeferred = 0;  // triangles its last direct draw queued for clipping
    };
    std::vector<View> views;
    std::vector<fgl_ctx *> ctxs;  // the distinct contexts, in first-use order
    int device = -1;
    hipStream_t s0 = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> joins;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t epoch = 0;  // g_dev_epoch the recorded addresses are valid for
    DevBuf ctrs;         // 2 DrawCounters per view (triangles, lines)
    HostBuf host;        // the same, read back
    std::string err;

    // Replays launch each stage once for many views instead of replaying a
    // graph with every view's launches.  The views' draws are split into
    // rounds (a context's draws go to successive rounds, in view order) and,
    // within a round, into groups of one tile shape and kernel variant; a
    // group's launches take a device table of its draws' arguments and, per
    // launch, the split of its workgroups among them.
    struct MLaunch {
        int stage, pass;
        uint32_t nblocks;
        size_t view_of, first;  // byte offsets in mtab
    };
    struct MGroup {
        const TileOps *T;
        DrawArgs rep;  // one of the group's draws (the kernel variant)
        size_t draws;  // byte offset of the DrawArgs table in mtab
        std::vector<MLaunch> launches;
    };
    // After each round, its draws' counters are copied to their views' slots
    // of ctrs (a context's counters are reset by its next draw).
    struct MGather {
        size_t groups_end;  // the round's last group + 1
        int32_t count;
        size_t src, slot;   // byte offsets in mtab
    };
    bool multi = false;  // replays use mgroups (else the graph)
    std::vector<MGroup> mgroups;
    std::vector<MGather> mgathers;
    DevBuf mtab;
    ~fgl_batch() {
        if (s0) (void)hipStreamSynchronize(s0);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (hipEvent_t e : joins) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
        if (s0) (void)hipStreamDestroy(s0);
    }
};

namespace {

void mesh_prims(const fgl_mesh *m, bool lines, DevPrims &p) {
    p.n = lines ? m->nline : m->ntri;
    const DevBuf *b[4] = {lines ? &m->lpos : &m->tpos, lines ? &m->lnrm : &m->tnrm, lines ? &m->luv : &m->tuv,
                          lines ? &m->lcol : &m->tcol};
    const bool *has = lines ? m->l_has : m->t_has;
    p.pos = has[0] ? b[0]->as<double>() : nullptr;
    p.nrm = has[1] ? b[1]->as<double>() : nullptr;
    p.uv = has[2] ? b[2]->as<double>() : nullptr;
    p.col = has[3] ? b[3]->as<double>() : nullptr;
}

#define BCHK(b, call)                                                                               \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess) {                                                                     \
            set_err((b)->err, "%s failed: %s", #call, hipGetErrorString(e_));                       \
            return FGL_E_DEVICE;                                                                    \
        }                                                                                           \
    } while (0)

// ClearColorBufferWith + ClearDepthBufferWith + DrawMesh of view i, drawn
// directly (buffers grow as needed).  A view with an empty mesh still gets
// its clears.
int batch_eager(fgl_batch *b, size_t i, uint64_t info[2]) {
    fgl_batch::View &v = b->views[i];
    v.ctx->clear_color = v.clear_color;
    v.ctx->pend_color = true;
    v.ctx->clear_depth = v.clear_depth;
    v.ctx->pend_depth = true;
    DevPrims t, l;
    mesh_prims(v.mesh, false, t);
    mesh_prims(v.mesh, true, l);
    uint64_t i1[2] = {0, 0}, i2[2] = {0, 0};
    int r = run_draw(v.ctx, &v.st, &v.sh, PRIM_TRI, t, i1);
    if (!r && t.n > 0) v.ndeferred = v.ctx->last.ndeferred;
    if (!r) r = run_draw(v.ctx, &v.st, &v.sh, PRIM_LINE, l, i2);
    if (!r && t.n == 0 && l.n == 0) r = materialize(v.ctx);
    if (r) {
        set_err(b->err, "view %zu: %s", i, v.ctx->err.c_str());
        return r;
    }
    info[0] = i1[0] + i2[0];
    info[1] = i1[1] + i2[1];
    return FGL_OK;
}

// Records every view into one graph.  Nothing is allocated while recording
// unless a capacity moved since the eager pass; then the addresses of the
// views recorded earlier may be stale, and the recording is made again.
int batch_record(fgl_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    std::vector<fgl_ctx *> &ctxs = b->ctxs;
    while (b->joins.size() < ctxs.size()) {
        hipEvent_t e;
        BCHK(b, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        b->joins.push_back(e);
    }
    std::vector<bool> prof(ctxs.size());
    for (size_t k = 0; k < ctxs.size(); k++) {
        prof[k] = ctxs[k]->prof;
        ctxs[k]->prof = false;  // no timing events inside the graph
    }
    const uint64_t epoch0 = g_dev_epoch.load();
    BCHK(b, hipStreamBeginCapture(b->s0, hipStreamCaptureModeThreadLocal));
    int r = FGL_OK;
    if (hipEventRecord(b->fork, b->s0) != hipSuccess) r = FGL_E_DEVICE;
    for (size_t k = 0; k < ctxs.size() && !r; k++)
        if (hipStreamWaitEvent(ctxs[k]->stream, b->fork, 0) != hipSuccess) r = FGL_E_DEVICE;
    uint32_t *dst = b->ctrs.as<uint32_t>();
    const int32_t words = (int32_t)(sizeof(DrawCounters) / 4);
    for (size_t i = 0; i < b->views.size() && !r; i++) {
        fgl_batch::View &v = b->views[i];
        v.ctx->clear_color = v.clear_color;
        v.ctx->pend_color = true;
        v.ctx->clear_depth = v.clear_depth;
        v.ctx->pend_depth = true;
        DevPrims t, l;
        mesh_prims(v.mesh, false, t);
        mesh_prims(v.mesh, true, l);
        r = run_draw(v.ctx, &v.st, &v.sh, PRIM_TRI, t, nullptr, dst + (2 * i) * words);
        if (t.n > 0) v.ctx->pend_color = v.ctx->pend_depth = false;  // the triangle pass clears
        if (!r) r = run_draw(v.ctx, &v.st, &v.sh, PRIM_LINE, l, nullptr, dst + (2 * i + 1) * words);
        if (l.n > 0) v.ctx->pend_color = v.ctx->pend_depth = false;
        if (!r && t.n == 0 && l.n == 0) r = materialize(v.ctx);  // fill kernels: the clear of an empty view
        if (r) set_err(b->err, "view %zu: %s", i, v.ctx->err.c_str());
    }
    for (size_t k = 0; k < ctxs.size() && !r; k++) {
        if (hipEventRecord(b->joins[k], ctxs[k]->stream) != hipSuccess ||
            hipStreamWaitEvent(b->s0, b->joins[k], 0) != hipSuccess)
            r = FGL_E_DEVICE;
    }
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(b->s0, &g);
    for (size_t k = 0; k < ctxs.size(); k++) ctxs[k]->prof = prof[k];
    if (r) {
        if (g) (void)hipGraphDestroy(g);
        if (b->err.empty()) b->err = "recording the batch failed";
        return r;
    }
    if (ee != hipSuccess) {
        set_err(b->err, "hipStreamEndCapture failed: %s", hipGetErrorString(ee));
        return FGL_E_DEVICE;
    }
    if (g_dev_epoch.load() != epoch0) {  // an address recorded above may be stale
        (void)hipGraphDestroy(g);
        return 1;
    }
    b->graph = g;
    BCHK(b, hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    b->epoch = epoch0;
    return FGL_OK;
}

// Prepares every view's draws as a replay launches them (its clears
// pending on its first draw) and builds the groups' tables.  Returns 1 when
// a buffer moved while preparing (prepare again), 2 when the batch has a
// view with an empty mesh (its clear is left to the graph).
int batch_record_multi(fgl_batch *b) {
    struct Item {
        size_t view;
        int kind;
        const TileOps *T;
        DrawArgs a;
        int round;
        uint32_t ndeferred;
    };
    const uint64_t epoch0 = g_dev_epoch.load();
    std::vector<Item> items;
    std::map<fgl_ctx *, int> used;  // draws per context so far
    for (size_t i = 0; i < b->views.size(); i++) {
        fgl_batch::View &v = b->views[i];
        if (v.mesh->ntri == 0 && v.mesh->nline == 0) return 2;
        v.ctx->clear_color = v.clear_color;
        v.ctx->pend_color = true;
        v.ctx->clear_depth = v.clear_depth;
        v.ctx->pend_depth = true;
        for (int kind : {(int)PRIM_TRI, (int)PRIM_LINE}) {
            DevPrims p;
            mesh_prims(v.mesh, kind == PRIM_LINE, p);
            if (p.n == 0) continue;
            Item it;
            it.view = i;
            it.kind = kind;
            int r = draw_begin(v.ctx, &v.st, &v.sh, kind, p, it.T, it.a);
            if (!r) r = draw_buffers(v.ctx, it.a);
            if (r) {
                set_err(b->err, "view %zu: %s", i, v.ctx->err.c_str());
                return r;
            }
            it.round = used[v.ctx]++;
            it.ndeferred = v.ndeferred;
            items.push_back(it);
            v.ctx->pend_color = v.ctx->pend_depth = false;  // consumed by this draw
        }
    }
    if (g_dev_epoch.load() != epoch0) return 1;

    // groups: round, then tile shape and kernel variant, in order of first use
    struct Key {
        int round;
        const TileOps *T;
        int kind, wire, shk, tex;
        bool operator<(const Key &o) const {
            return std::tie(round, T, kind, wire, shk, tex) < std::tie(o.round, o.T, o.kind, o.wire, o.shk, o.tex);
        }
    };
    std::map<Key, size_t> gid;
    std::vector<std::vector<size_t>> members;
    std::vector<Key> keys;
    std::vector<size_t> round_end;  // groups of rounds 0..rd
    int nround = 0;
    for (const Item &it : items) nround = std::max(nround, it.round + 1);
    for (int rd = 0; rd < nround; rd++, round_end.push_back(members.size()))
        for (size_t k = 0; k < items.size(); k++) {
            const Item &it = items[k];
            if (it.round != rd) continue;
            const Key key{rd, it.T, it.kind, it.a.st.wireframe ? 1 : 0, it.a.sh.kind, it.a.sh.has_tex ? 1 : 0};
            auto f = gid.find(key);
            if (f == gid.end()) {
                f = gid.emplace(key, members.size()).first;
                members.emplace_back();
                keys.push_back(key);
            }
            members[f->second].push_back(k);
        }

    // the tables, each 256-byte aligned in one buffer
    std::vector<uint8_t> tab;
    auto put = [&](const void *src, size_t bytes) {
        const size_t off = (tab.size() + 255) & ~(size_t)255;
        tab.resize(off + bytes);
        if (bytes) std::memcpy(tab.data() + off, src, bytes);
        return off;
    };
    std::vector<fgl_batch::MGroup> groups;
    for (size_t g = 0; g < members.size(); g++) {
        const std::vector<size_t> &mem = members[g];
        fgl_batch::MGroup G;
        G.T = keys[g].T;
        G.rep = items[mem[0]].a;
        std::vector<DrawArgs> draws;
        std::vector<uint32_t> ndef;
        int maxpass = 0;
        for (size_t k : mem) {
            draws.push_back(items[k].a);
            ndef.push_back(items[k].ndeferred);
            maxpass = std::max(maxpass, (items[k].a.key_bits + 7) / 8);
        }
        G.draws = put(draws.data(), draws.size() * sizeof(DrawArgs));
        auto add = [&](int stage, int pass) {
            std::vector<uint32_t> first(draws.size() + 1, 0), view_of;
            for (size_t v = 0; v < draws.size(); v++) {
                const DrawArgs &a = draws[v];
                const bool in_pass = (a.key_bits + 7) / 8 > pass;
                const bool radix = stage == ST_RADIX_HIST || stage == ST_RADIX_SCAN || stage == ST_RADIX_SCATTER;
                uint32_t nb = (!radix || in_pass) ? G.T->stage_blocks(stage, a, true) : 0u;
                // the clipping stage's grid-stride loop takes any count; its
                // grid follows the count of the view's direct draw
                if (stage == ST_VS_DEFERRED) nb = std::max(nb, std::min<uint32_t>((ndef[v] + ndef[v] / 4 + 63) / 64, 4096u));
                first[v] = (uint32_t)view_of.size();
                view_of.insert(view_of.end(), nb, (uint32_t)v);
            }
            first[draws.size()] = (uint32_t)view_of.size();
            fgl_batch::MLaunch L;
            L.stage = stage;
            L.pass = pass;
            L.nblocks = (uint32_t)view_of.size();
            L.first = put(first.data(), first.size() * 4);
            L.view_of = put(view_of.data(), view_of.size() * 4);
            G.launches.push_back(L);
        };
        // the stages of draw_pass, in its order
        add(ST_ZERO, 0);
        add(ST_VS, 0);
        if (keys[g].kind == PRIM_TRI && !keys[g].wire) add(ST_VS_DEFERRED, 0);
        add(ST_SCAN_TOP, 0);
        add(ST_EMIT, 0);
        for (int p = 0; p < maxpass; p++) {
            add(ST_RADIX_HIST, p);
            add(ST_RADIX_SCAN, p);
            add(ST_RADIX_SCATTER, p);
        }
        add(ST_SPAN_COUNT, 0);
        add(ST_SPAN_SCAN, 0);
        add(ST_SPANS, 0);
        add(ST_FRAG_SCAN, 0);
        add(ST_SHADE, 0);
        add(ST_ROP, 0);
        groups.push_back(std::move(G));
    }
    // each round's counters, gathered into the per-view slots
    std::vector<fgl_batch::MGather> gathers;
    for (int rd = 0; rd < nround; rd++) {
        std::vector<uint64_t> src;
        std::vector<uint32_t> slot;
        for (const Item &it : items) {
            if (it.round != rd) continue;
            src.push_back((uint64_t)(uintptr_t)it.a.ctr);
            slot.push_back((uint32_t)(2 * it.view + (it.kind == PRIM_LINE ? 1 : 0)));
        }
        fgl_batch::MGather G;
        G.groups_end = round_end[(size_t)rd];
        G.count = (int32_t)src.size();
        G.src = put(src.data(), src.size() * 8);
        G.slot = put(slot.data(), slot.size() * 4);
        gathers.push_back(G);
    }

    BCHK(b, b->mtab.ensure(tab.size()));
    if (g_dev_epoch.load() != epoch0) return 1;  // the table itself may have moved a buffer's address
    BCHK(b, hipMemcpy(b->mtab.p, tab.data(), tab.size(), hipMemcpyHostToDevice));
    b->mgroups = std::move(groups);
    b->mgathers = std::move(gathers);
    b->multi = true;
    b->epoch = g_dev_epoch.load();
    return FGL_OK;
}

// One replay: the groups' launches, each round followed by its counters
// gather, on s0.
int batch_replay_multi(fgl_batch *b) {
    uint8_t *base = b->mtab.as<uint8_t>();
    size_t g = 0;
    for (const fgl_batch::MGather &R : b->mgathers) {
        for (; g < R.groups_end; g++) {
            const fgl_batch::MGroup &G = b->mgroups[g];
            for (const fgl_batch::MLaunch &L : G.launches) {
                MultiArgs m;
                m.draws = reinterpret_cast<const DrawArgs *>(base + G.draws);
                m.view_of = reinterpret_cast<const uint32_t *>(base + L.view_of);
                m.first = reinterpret_cast<const uint32_t *>(base + L.first);
                m.pass = L.pass;
                BCHK(b, G.T->multi(L.stage, m, L.nblocks, G.rep, b->s0));
            }
        }
        BCHK(b, launch_gather_words(reinterpret_cast<const uint32_t *const *>(base + R.src),
                                    reinterpret_cast<const uint32_t *>(base + R.slot), b->ctrs.as<uint32_t>(),
                                    (int32_t)(sizeof(DrawCounters) / 4), R.count, b->s0));
    }
    return FGL_OK;
}

int batch_capture(fgl_batch *b) {
    b->multi = false;
    const char *env = std::getenv("FGL_BATCH_GRAPH");
    const bool graph_only = env && env[0] == '1';
    for (int attempt = 0; attempt < 3 && !graph_only; attempt++) {
        const int r = batch_record_multi(b);
        if (r == 2) break;  // a view with an empty mesh: the graph records its clear
        if (r <= 0) return r;
    }
    for (int attempt = 0; attempt < 3; attempt++) {
        const int r = batch_record(b);
        if (r <= 0) return r;
    }
    b->err = "recording the batch kept reallocating device buffers";
    return FGL_E_DEVICE;
}

// The eager pass: every view drawn directly (sizing every context's
// buffers), then the graph recorded.
int batch_first_run(fgl_batch *b, uint64_t *out) {
    const size_t n = b->views.size();
    for (size_t i = 0; i < n; i++) {
        int r = batch_eager(b, i, out + 2 * i);
        if (r) return r;
    }
    if (b->ctrs.ensure(2 * n * sizeof(DrawCounters)) != hipSuccess) {
        b->err = "fgl_batch_run: device allocation failed";
        return FGL_E_NOMEM;
    }
    for (fgl_ctx *c : b->ctxs) BCHK(b, hipStreamSynchronize(c->stream));
    return batch_capture(b);
}

}  // namespace

extern "C" {

int fgl_batch_create(fgl_batch **out) {
    if (!out) return FGL_E_INVALID;
    *out = new (std::nothrow) fgl_batch;
    return *out ? FGL_OK : FGL_E_NOMEM;
}

void fgl_batch_destroy(fgl_batch *b) {
    if (!b) return;
    if (b->s0) (void)hipStreamSynchronize(b->s0);
    delete b;
}

const char *fgl_batch_error(const fgl_batch *b) { return b ? b->err.c_str() : "null batch"; }

int fgl_batch_add(fgl_batch *b, fgl_ctx *ctx, const uint8_t clear_nrgba[4], double clear_depth, const fgl_state *st,
                  const fgl_shader *sh, const fgl_mesh *mesh) {
    if (!b || check_ctx(ctx) || !clear_nrgba || !st || !sh || !mesh) return FGL_E_INVALID;
    if (b->device >= 0 && ctx->device != b->device) {
        b->err = "fgl_batch_add: every view's context must be on one device";
        return FGL_E_INVALID;
    }
    if (mesh->ctx->device != ctx->device) {
        b->err = "fgl_batch_add: mesh lives on another device";
        return FGL_E_INVALID;
    }
    b->device = ctx->device;
    fgl_batch::View v;
    v.ctx = ctx;
    v.mesh = mesh;
    v.st = *st;
    v.sh = *sh;
    v.clear_color = (uint32_t)clear_nrgba[0] | ((uint32_t)clear_nrgba[1] << 8) | ((uint32_t)clear_nrgba[2] << 16) |
                    ((uint32_t)clear_nrgba[3] << 24);
    v.clear_depth = clear_depth;
    b->views.push_back(v);
    if (std::find(b->ctxs.begin(), b->ctxs.end(), ctx) == b->ctxs.end()) b->ctxs.push_back(ctx);
    if (b->exec) {  // the recorded graph no longer covers every view
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    b->multi = false;
    return FGL_OK;
}

int fgl_batch_run(fgl_batch *b, uint64_t *info) {
    if (!b) return FGL_E_INVALID;
    const size_t n = b->views.size();
    if (n == 0) return FGL_OK;
    BCHK(b, hipSetDevice(b->device));
    if (!b->s0) {
        BCHK(b, hipStreamCreateWithFlags(&b->s0, hipStreamNonBlocking));
        BCHK(b, hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    }
    b->err.clear();
    std::vector<uint64_t> tmp(2 * n);
    uint64_t *out = info ? info : tmp.data();
    if (!(b->multi || b->exec) || b->epoch != g_dev_epoch.load()) return batch_first_run(b, out);
    // the replay follows whatever is already queued on the views' contexts
    // (a resolve or readback of the previous frame, a clear's fill)
    // (a stream with nothing pending, the usual case, needs no event)
    for (fgl_ctx *c : b->ctxs) {
        const hipError_t q = hipStreamQuery(c->stream);
        if (q == hipSuccess) continue;
        if (q != hipErrorNotReady) BCHK(b, q);
        if (!c->xev) BCHK(b, hipEventCreateWithFlags(&c->xev, hipEventDisableTiming));
        BCHK(b, hipEventRecord(c->xev, c->stream));
        BCHK(b, hipStreamWaitEvent(b->s0, c->xev, 0));
    }
    (void)hipGetLastError();  // a not-ready query is no error of a later launch
    if (b->multi) {
        const int r = batch_replay_multi(b);
        if (r) return r;
    } else {
        BCHK(b, hipGraphLaunch(b->exec, b->s0));
    }
    BCHK(b, b->host.ensure(2 * n * sizeof(DrawCounters)));
    BCHK(b, hipMemcpyAsync(b->host.p, b->ctrs.p, 2 * n * sizeof(DrawCounters), hipMemcpyDeviceToHost, b->s0));
    BCHK(b, wait_stream(b->s0));
    const DrawCounters *hc = b->host.as<DrawCounters>();
    // every view cleared its context as it was drawn (its ROP may have left
    // blocks at the clear values): a context's clear values are now its last
    // view's, whatever was cleared on it between runs
    for (fgl_ctx *c : b->ctxs) {
        c->pend_color = c->pend_depth = false;
        c->cblk_any = true;
    }
    for (const fgl_batch::View &v : b->views) {
        v.ctx->clear_color = v.clear_color;
        v.ctx->clear_depth = v.clear_depth;
    }
    // the first view per context whose frame no longer fits the recorded
    // capacities
    std::map<fgl_ctx *, size_t> first_ovf;
    for (size_t i = 0; i < n; i++) {
        const DrawCounters &t = hc[2 * i], &l = hc[2 * i + 1];
        const bool has_t = b->views[i].mesh->ntri > 0, has_l = b->views[i].mesh->nline > 0;
        if ((has_t && t.overflow) || (has_l && l.overflow)) first_ovf.emplace(b->views[i].ctx, i);
    }
    for (size_t i = 0; i < n; i++) {
        fgl_batch::View &v = b->views[i];
        auto it = first_ovf.find(v.ctx);
        if (it != first_ovf.end() && i >= it->second) {
            // redrawn directly (growing its buffers), in view order, so the
            // context ends with the last of its views
            int r = batch_eager(b, i, out + 2 * i);
            if (r) return r;
            continue;
        }
        const DrawCounters &t = hc[2 * i], &l = hc[2 * i + 1];
        const bool has_t = v.mesh->ntri > 0, has_l = v.mesh->nline > 0;
        out[2 * i] = (has_t ? t.total : 0) + (has_l ? l.total : 0);
        out[2 * i + 1] = (has_t ? updated_pixels(t) : 0) + (has_l ? updated_pixels(l) : 0);
        if (has_t || has_l) v.ctx->last = has_l ? l : t;
    }
    if (!first_ovf.empty()) {
        for (fgl_ctx *c : b->ctxs) BCHK(b, hipStreamSynchronize(c->stream));
        return batch_capture(b);
    }
    return FGL_OK;
}

}  // extern "C"
