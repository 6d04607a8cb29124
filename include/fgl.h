/*
 * fgl.h -- C ABI of the MI355X (gfx950) draw path behind fauxgl's Go API.
 *
 * A thin cgo shim in package fauxgl binds these entry points.  Every entry
 * point names the reference function it replaces (file:line in
 * fogleman/fauxgl).
 *
 * Conventions
 *   - Plain C types only.  Host pointers unless the name says "device".
 *   - Every function returns FGL_OK (0) or a negative fgl_status; the
 *     message is available from fgl_last_error(ctx) (or
 *     fgl_last_error(NULL) for failures that have no context yet).
 *     Nothing aborts the process.
 *   - Calls on one context are serialised on that context's HIP stream and
 *     are synchronous unless the name ends in _async.  Different contexts
 *     may be used from different threads.
 *   - The framebuffer lives in device memory: color is NRGBA8
 *     (non-premultiplied, byte order R,G,B,A, row-major, stride 4*W, like
 *     image.NRGBA.Pix) and depth is float64 (W*H, index y*W+x), exactly the
 *     layouts of Context.ColorBuffer / Context.DepthBuffer (context.go:43-44).
 *
 * Semantics follow the single-threaded reference: primitives are drawn in
 * input order (a clipped triangle's fan in fan order, DrawMesh's triangles
 * before its lines), so coverage and depth are bit-exact against fauxgl run
 * with one worker, and RasterizeInfo matches.
 */
#ifndef FGL_H
#define FGL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum fgl_status {
    FGL_OK = 0,
    FGL_E_INVALID = -1,     /* bad argument (null pointer, size, shader / texture / filter kind) */
    FGL_E_NOMEM = -2,       /* device or host allocation failed */
    FGL_E_DEVICE = -3,      /* HIP runtime error or no usable gfx950 device */
    FGL_E_UNSUPPORTED = -4, /* request outside what the device path implements */
    FGL_E_IO = -5           /* a file could not be opened, read or written */
} fgl_status;

/* context.go:11-26: the same values as Go's iota enums (FaceCW = 1,
 * FaceCCW = 2; CullNone = 1, CullFront = 2, CullBack = 3), so a Go caller
 * passes dc.FrontFace and dc.Cull unchanged. */
enum { FGL_FACE_CW = 1, FGL_FACE_CCW = 2 };
enum { FGL_CULL_NONE = 1, FGL_CULL_FRONT = 2, FGL_CULL_BACK = 3 };

/* Built-in shaders (shader.go:11-59). */
enum { FGL_SHADER_SOLID = 1, FGL_SHADER_TEXTURE = 2, FGL_SHADER_PHONG = 3 };

/* Texture texel formats.  NRGBA8: an image.NRGBA (At().RGBA() premultiplies
 * on the fly, color.go:25-29 via image/color); RGBA8: an image.RGBA (already
 * premultiplied); RGBA16: the four 16-bit values At(x,y).RGBA() returns for
 * any Go image type, stored as uint16 R,G,B,A. */
enum { FGL_TEX_NRGBA8 = 1, FGL_TEX_RGBA8 = 2, FGL_TEX_RGBA16 = 3 };

/* SSAA resolve filters (the step after the path: the reference pipeline
 * calls nfnt/resize's Bilinear, README.md:109). */
enum { FGL_RESOLVE_TENT = 1, FGL_RESOLVE_BOX = 2 };

/* Pipeline state = the public Context fields read at draw time
 * (context.go:47-55).  front_face and cull accept any int32, as the Go
 * fields accept any int, and act as drawClippedTriangle's comparisons do
 * (context.go:328-336): only FGL_FACE_CW reverses the winding; FGL_CULL_NONE
 * keeps both faces, FGL_CULL_FRONT culls front faces, and every other value
 * (0 included) culls back faces like FGL_CULL_BACK.  Nothing rejects them. */
typedef struct fgl_state {
    uint8_t read_depth, write_depth, write_color, alpha_blend, wireframe;
    uint8_t pad[3];
    int32_t front_face; /* FGL_FACE_* (any value, see above) */
    int32_t cull;       /* FGL_CULL_* (any value, see above) */
    double line_width;
    double depth_bias;
} fgl_state;

/* A built-in shader (shader.go:11-59).  Fields a shader kind does not have
 * are ignored. */
typedef struct fgl_shader {
    int32_t kind; /* FGL_SHADER_* */
    int32_t pad;
    double matrix[16];       /* Matrix X00..X33, row-major (matrix.go:5-10) */
    double color[4];         /* SolidColorShader.Color */
    double light_dir[3];     /* PhongShader.LightDirection */
    double camera_pos[3];    /* PhongShader.CameraPosition */
    double object_color[4];  /* PhongShader.ObjectColor (all zero = Discard = use vertex color) */
    double ambient[4];       /* PhongShader.AmbientColor */
    double diffuse[4];       /* PhongShader.DiffuseColor */
    double specular[4];      /* PhongShader.SpecularColor */
    double specular_power;   /* PhongShader.SpecularPower */
    int64_t texture;         /* TextureShader.Texture / PhongShader.Texture handle; -1 = nil */
} fgl_shader;

/* A triangle or line soup in SoA form: nv = 3 vertices per triangle, 2 per
 * line, in the order Triangle.V1,V2,V3 / Line.V1,V2 (triangle.go:3-5,
 * line.go:3-5).  position: count*nv*3 doubles (Vertex.Position); normal:
 * count*nv*3 (Vertex.Normal); texture: count*nv*2 (Vertex.Texture.X,Y; the
 * built-in shaders never read Texture.Z); color: count*nv*4 (Vertex.Color).
 * A NULL attribute array stands for all-zero values, like a zero Vertex
 * field. */
typedef struct fgl_prims {
    int64_t count;
    const double *position;
    const double *normal;
    const double *texture;
    const double *color;
} fgl_prims;

typedef struct fgl_ctx fgl_ctx;
typedef struct fgl_mesh fgl_mesh;

/* ---- library ---- */

/* Selects device `device` (0-based) for contexts created afterwards on this
 * thread and checks it is a gfx950 part.  Optional: fgl_ctx_create selects
 * device 0 when this was never called.  No reference counterpart. */
int fgl_init(int32_t device);

/* Message for the last failure on ctx, or the last failure without a
 * context when ctx is NULL.  Never returns NULL. */
const char *fgl_last_error(const fgl_ctx *ctx);

/* ---- context: NewContext, context.go:60-81 ---- */

/* Allocates a W x H color + depth framebuffer on the device and clears the
 * depth buffer to math.MaxFloat64 (context.go:79); the color buffer starts
 * all zero like image.NewNRGBA. */
int fgl_ctx_create(int32_t width, int32_t height, fgl_ctx **out);
void fgl_ctx_destroy(fgl_ctx *ctx);
int fgl_ctx_size(const fgl_ctx *ctx, int32_t *width, int32_t *height);

/* ClearColorBufferWith (context.go:119-131): nrgba is Color.NRGBA() of the
 * clear color, converted by the caller (color.go:56-63).  Deferred: the
 * next draw writes the cleared tiles without reading them first. */
int fgl_clear_color(fgl_ctx *ctx, const uint8_t nrgba[4]);
/* ClearDepthBufferWith (context.go:137-141); ClearDepthBuffer passes
 * math.MaxFloat64.  Deferred like fgl_clear_color. */
int fgl_clear_depth(fgl_ctx *ctx, double value);

/* ---- textures: ImageTexture, texture.go:21-25 ---- */

/* Uploads width*height texels (row-major; 4 uint8 per texel for
 * FGL_TEX_NRGBA8/RGBA8, 4 uint16 per texel for FGL_TEX_RGBA16) and returns
 * a handle usable in fgl_shader.texture on this context. */
int fgl_texture_create(fgl_ctx *ctx, int32_t width, int32_t height, int32_t kind, const void *texels,
                       int64_t *handle);
int fgl_texture_destroy(fgl_ctx *ctx, int64_t handle);

/* ---- drawing ---- */

/* DrawTriangles (context.go:413-433) with one of the built-in shaders:
 * vertex shader, view-volume clipping, culling, rasterization, depth test,
 * shading and blending, in input order.  info[0] = TotalPixels,
 * info[1] = UpdatedPixels (context.go:28-31).  The host arrays are copied
 * to the device during the call. */
int fgl_draw_triangles(fgl_ctx *ctx, const fgl_state *state, const fgl_shader *shader, const fgl_prims *tris,
                       uint64_t info[2]);
/* DrawLines (context.go:391-411): each line becomes a LineWidth-wide quad
 * (context.go:283-294). */
int fgl_draw_lines(fgl_ctx *ctx, const fgl_state *state, const fgl_shader *shader, const fgl_prims *lines,
                   uint64_t info[2]);

/* Device-resident meshes: Mesh{Triangles, Lines} (mesh.go:9-13) uploaded
 * once, for callers that draw the same mesh many times.  Any context on the
 * same device may draw a mesh.  The mesh's own calls (fgl_mesh_read, the
 * preparation calls below) run on the stream of the context it was created
 * on; once that context is destroyed, fgl_mesh_destroy is the only call left
 * (it waits for the device and needs no context). */
int fgl_mesh_create(fgl_ctx *ctx, const fgl_prims *tris, const fgl_prims *lines, fgl_mesh **out);
void fgl_mesh_destroy(fgl_mesh *mesh);
/* DrawMesh (context.go:435-439): triangles, then lines. */
int fgl_draw_mesh(fgl_ctx *ctx, const fgl_state *state, const fgl_shader *shader, const fgl_mesh *mesh,
                  uint64_t info[2]);

/* ---- tile shape and the autotuner (no reference counterpart: the
 * reference has no tiles; for callers that want the fastest shape for their
 * workload) ----
 *
 * The draw kernels are compiled for several tile shapes; every shape gives
 * the same images.  A draw uses the context's shape if one is set, else the
 * tuning table's entry for (width, height, primitive kind, bit length of the
 * primitive count), whose pair / span / fragment capacities it also reserves
 * up front, else 32 x 32.
 *
 * fgl_tile_shapes: the compiled shapes as (w, h) pairs; *n = their count,
 *   at most cap written.
 * fgl_ctx_set_tile: this context's shape (0, 0: the table / default).
 * fgl_ctx_get_tile: the shape a draw of ntri triangles on ctx would use.
 * fgl_autotune_mesh: draws mesh reps + 1 times with each shape (the first
 *   draw per shape sizes the buffers; the fastest of the others counts),
 *   records the fastest shape and the capacities its draw needed in the
 *   table, and returns the shape.  The context's framebuffer, pending clears
 *   and last-draw counters are restored afterwards.
 * fgl_tune_save / fgl_tune_load: the table as a text file (load merges,
 *   rejecting a file with any malformed line or unknown shape);
 *   fgl_tune_clear empties it.  The table is process-wide. */
int fgl_tile_shapes(int32_t *wh, int32_t cap, int32_t *n);
int fgl_ctx_set_tile(fgl_ctx *ctx, int32_t tile_w, int32_t tile_h);
int fgl_ctx_get_tile(fgl_ctx *ctx, int64_t ntri, int32_t out_wh[2]);
int fgl_autotune_mesh(fgl_ctx *ctx, const fgl_state *state, const fgl_shader *shader, const fgl_mesh *mesh,
                      int32_t reps, int32_t out_wh[2]);
int fgl_tune_save(const char *path);
int fgl_tune_load(const char *path);
int fgl_tune_clear(void);

/* ---- batched views (no reference counterpart: callers that render many
 * views of resident meshes per step, hundreds of views of meshes of uneven
 * size at mixed resolutions) ----
 *
 * A batch is a list of views: (context, clear color and depth, state,
 * shader, resident mesh) = ClearColorBufferWith + ClearDepthBufferWith +
 * DrawMesh on that context.  The first fgl_batch_run draws every view one
 * by one, which sizes each context's buffers, and prepares the views' draws;
 * later runs launch each pipeline stage once for all views of one tile shape
 * and kernel variant (a context's views in successive rounds, in order).  A
 * batch with a view whose mesh is empty is recorded into one HIP graph of
 * per-view launches instead.  A view whose frame no longer fits is redrawn
 * directly and the batch prepared again.  info: 2 per view
 * (TotalPixels, UpdatedPixels), may be NULL.  A view's context must not be
 * used by other calls while a run is in flight. */
typedef struct fgl_batch fgl_batch;
int fgl_batch_create(fgl_batch **out);
void fgl_batch_destroy(fgl_batch *batch);
const char *fgl_batch_error(const fgl_batch *batch);
int fgl_batch_add(fgl_batch *batch, fgl_ctx *ctx, const uint8_t clear_nrgba[4], double clear_depth,
                  const fgl_state *state, const fgl_shader *shader, const fgl_mesh *mesh);
int fgl_batch_run(fgl_batch *batch, uint64_t *info);

/* ---- device-side mesh preparation on a resident mesh (in place) ---- */

/* Primitive counts of a resident mesh. */
int fgl_mesh_counts(const fgl_mesh *mesh, int64_t *ntri, int64_t *nline);
/* Copies one attribute out (lines = 0: triangles, 1: lines; attr 0..3 =
 * position, normal, texture, color in the fgl_prims layout; n = the number
 * of doubles).  An attribute the mesh was created without reads as zeros. */
int fgl_mesh_read(const fgl_mesh *mesh, int32_t lines, int32_t attr, double *out, int64_t n);
/* Mesh.BoundingBox (mesh.go:153-165): box = {min x, y, z, max x, y, z},
 * Box.Extend semantics (box.go:48-53: a running box equal to EmptyBox is
 * replaced) and math.Min / Max. */
int fgl_mesh_bounding_box(fgl_mesh *mesh, double box[6]);
/* Mesh.Transform (mesh.go:167-175): positions by MulPosition, normals by
 * MulDirection (matrix.go:209-229); matrix row-major X00..X33. */
int fgl_mesh_transform(fgl_mesh *mesh, const double matrix[16]);
/* Mesh.FitInside (mesh.go:142-151) into box {min, max} at anchor; writes the
 * matrix it applied to matrix_out when not NULL.  BiUnitCube / UnitCube
 * (mesh.go:127-135) are box +-1 / +-0.5 with anchor 0.5. */
int fgl_mesh_fit_inside(fgl_mesh *mesh, const double box[6], const double anchor[3], double matrix_out[16]);
/* Mesh.SmoothNormals (mesh.go:105-125). */
int fgl_mesh_smooth_normals(fgl_mesh *mesh);
/* Mesh.SmoothNormalsThreshold (mesh.go:91-103) given the threshold the
 * reference computes first, math.Cos(radians) (the caller evaluates it, so
 * that the value is Go's bit for bit). */
int fgl_mesh_smooth_normals_threshold(fgl_mesh *mesh, double cos_threshold);

/* ---- mesh ingest into a resident mesh ---- */

/* LoadSTL (stl.go:23-154) from the file's bytes: binary when size is
 * 84 + 50 * (little-endian uint32 at byte 80), decoded on the device
 * (float32 widened, face normals); otherwise ASCII "vertex x y z" lines.
 * Errors where the reference errors or panics (size < 84, an ASCII vertex
 * count that is not a multiple of 3). */
int fgl_mesh_load_stl(fgl_ctx *ctx, const uint8_t *data, int64_t size, fgl_mesh **out);
/* LoadOBJ (obj.go:19-79) from the file's text: v / vt / vn / f, fan
 * triangulation, 1-based and negative indices, FixNormals on the device.
 * Errors where the reference panics (an index out of range, a v / vt / vn
 * line with too few numbers). */
int fgl_mesh_load_obj(fgl_ctx *ctx, const char *text, int64_t size, fgl_mesh **out);

/* ---- encode ---- */

/* SavePNG (util.go:56-63) to memory: the color buffer as the PNG Go's
 * image/png writes for an *image.NRGBA (RGB when every alpha is 255, else
 * RGBA; 8-bit; each row's filter chosen as image/png chooses it), the zlib
 * stream deflated on the device (its bytes are this build's own; it
 * inflates to the same filtered rows).  out must hold
 * fgl_encode_png_bound bytes; *len receives the file size. */
int fgl_encode_png_bound(fgl_ctx *ctx, int64_t *bound);
int fgl_encode_png(fgl_ctx *ctx, uint8_t *out, int64_t cap, int64_t *len);

/* ---- readback and host writes (the public ColorBuffer/DepthBuffer fields) ---- */

/* Image() (context.go:83-85): copies the 4*W*H NRGBA bytes out. */
int fgl_read_color(fgl_ctx *ctx, uint8_t *nrgba);
/* DepthBuffer: copies W*H float64 out. */
int fgl_read_depth(fgl_ctx *ctx, double *depth);
/* For callers that modified ColorBuffer.Pix / DepthBuffer on the host. */
int fgl_write_color(fgl_ctx *ctx, const uint8_t *nrgba);
int fgl_write_depth(fgl_ctx *ctx, const double *depth);
/* DepthImage() (context.go:87-117): W*H Gray16 values. */
int fgl_depth_image(fgl_ctx *ctx, uint16_t *gray16);
/* No counterpart in the reference: fgl_depth_image for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The same kernels and the same W*H Gray16 values,
 * written to gray16_device, a device buffer of at least W*H uint16 on ctx's
 * device.  Nothing is allocated: the min/max partials live in a fixed 8 KiB
 * scratch array per GPU (512 blocks of min and max, enough for any image
 * size), and the call only queues work on ctx's stream (fgl_ctx_sync waits
 * for it).  Every call on one GPU shares that GPU's scratch, so two such
 * calls in flight at once on one GPU (contexts on different streams or
 * threads) race; serialise them. */
int fgl_depth_image_noalloc_async(fgl_ctx *ctx, uint16_t *gray16_device);

/* SSAA resolve: dst (w x h) color = src (w*f x h*f) color downsampled by the
 * integer factor f with `filter`.  FGL_RESOLVE_TENT is a separable 8-bit
 * fixed-point tent filter of radius f (the Bilinear filter widened by the
 * scale factor); FGL_RESOLVE_BOX averages each f x f block.  Both contexts
 * must be on the same device; dst's depth buffer is untouched. */
int fgl_resolve(fgl_ctx *src, fgl_ctx *dst, int32_t filter);

/* ---- measurement (no reference counterpart) ---- */

/* When enabled, every kernel launch on ctx is bracketed by HIP events on the
 * context's stream and its duration accumulated per kernel. */
int fgl_profile_enable(fgl_ctx *ctx, int32_t enable);
/* Restricts the timing to launches of one kernel (its name as reported by
 * fgl_profile_read); NULL or "" times every launch again.  Two events per
 * frame perturb a timed loop far less than two per launch. */
int fgl_profile_only(fgl_ctx *ctx, const char *kernel);
/* Copies up to `cap` records: name (NUL-terminated, 48 bytes), launch count
 * and total milliseconds.  Returns the number of kernels via *n. */
typedef struct fgl_kernel_time {
    char name[48];
    uint64_t launches;
    double total_ms;
} fgl_kernel_time;
int fgl_profile_read(fgl_ctx *ctx, fgl_kernel_time *out, int32_t cap, int32_t *n);
int fgl_profile_reset(fgl_ctx *ctx);
/* Waits for all work queued on ctx. */
int fgl_ctx_sync(fgl_ctx *ctx);

/* Work counts of the last completed draw on ctx (diagnostics; no reference
 * counterpart beyond RasterizeInfo's two totals).  clipped_triangles counts
 * input triangles that crossed a clip plane, side_triangles the raster
 * triangles written by clipping, wireframe and line expansion, pairs the
 * (tile, triangle) bin entries, spans the per-row coverage runs and
 * fragments the covered in-bounds pixels shaded.  overflow_mask: the
 * capacity bits the draw's passes hit (1 side triangles, 2 clipped
 * vertices, 4 pairs, 16 spans, 32 fragments); reruns: how many passes ran
 * again after growing the buffers. */
typedef struct fgl_draw_stat {
    uint64_t total_pixels, updated_pixels;
    uint64_t clipped_triangles, side_triangles;
    uint64_t pairs, spans, fragments;
    uint64_t overflow_mask, reruns;
} fgl_draw_stat;
int fgl_draw_stats(fgl_ctx *ctx, fgl_draw_stat *out);

/* Failure injection (no reference counterpart; for tests of the capacity
 * path).  The next triangle or line draw on ctx starts its first pass with
 * at most caps[0] side triangles, caps[1] clipped vertices, caps[2] pairs,
 * caps[3] spans and caps[4] fragments (0 = as computed).  A kernel that
 * runs out sets its overflow bit and exits normally; the host then grows
 * the buffers and runs the draw again, so the result is unchanged and
 * fgl_draw_stats reports the bits and the reruns. */
int fgl_ctx_force_capacities(fgl_ctx *ctx, const uint32_t caps[5]);

/* ---- learned anti-aliasing (no reference counterpart: it replaces the
 * 16x SSAA render + resolve that the reference pipeline would run) ----
 *
 * A six-layer 3x3 convolutional network (BF16 MFMA, FP32 accumulate) maps
 * a 1x render's color, depth and depth-derived normals to an estimate of
 * the 4x4-supersampled image.  Parameters are FP32, laid out per layer as
 * weights [9][Cout][Cin] then bias [Cout], layers 16->32, 32->32 (x4),
 * 32->16 (input channels 8..15 and output channels 3..15 are padding). */
typedef struct fgl_nn fgl_nn;
/* Allocates the network on ctx's device with seeded He initialisation. */
int fgl_nn_create(fgl_ctx *ctx, uint64_t seed, fgl_nn **out);
void fgl_nn_destroy(fgl_nn *nn);
int fgl_nn_param_count(const fgl_nn *nn, int64_t *n);
int fgl_nn_get_params(fgl_nn *nn, float *params, int64_t n);
/* Also resets the optimiser state. */
int fgl_nn_set_params(fgl_nn *nn, const float *params, int64_t n);
/* dst color = network(src); src and dst have the same size. */
int fgl_nn_infer(fgl_nn *nn, fgl_ctx *src, fgl_ctx *dst);
/* The network's raw output for src: H*W*16 floats, pixel-major; channels
 * 0..2 are the residual added to the input color (in units of 255). */
int fgl_nn_residual(fgl_nn *nn, fgl_ctx *src, float *out, int64_t n);
/* Mean squared error of network(src) against target's color (same size)
 * and, when grad is not NULL, the gradient of every parameter (n floats). */
int fgl_nn_grad(fgl_nn *nn, fgl_ctx *src, fgl_ctx *target, float *loss, float *grad, int64_t n);
/* One Adam step (beta 0.9 / 0.999) on that loss. */
int fgl_nn_train_step(fgl_nn *nn, fgl_ctx *src, fgl_ctx *target, float lr, float *loss);

#ifdef __cplusplus
}
#endif
#endif
