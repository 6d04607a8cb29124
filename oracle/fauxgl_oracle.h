/*
 * fauxgl_oracle.h -- CPU restatement of fauxgl's draw path, used ONLY as the
 * checker (parity tests, golden-fixture generation, smoke check) and as the
 * timed CPU baseline.  It is test infrastructure: the product library never
 * links, loads or calls it.
 *
 * It follows the reference Go sources line by line (citations are
 * file:line into fogleman/fauxgl) with Go's floating-point semantics:
 * IEEE binary64, no multiply-add contraction (built with -ffp-contract=off),
 * left-to-right evaluation, math.Min/Max NaN and signed-zero rules, amd64
 * float->int conversion (NaN / out of range -> INT64_MIN), and Go's
 * math.Pow algorithm for the integer part of the exponent.
 *
 * Two draw modes:
 *   nthreads == 1 : the deterministic oracle ("fauxgl run single-threaded").
 *   nthreads  > 1 : mirror of DrawTriangles/DrawLines (context.go:391-433):
 *                   worker wi takes primitive i when i % wn == wi, framebuffer
 *                   writes under 256 mutexes striped by (x+y)&255, early depth
 *                   read without the lock.  Used only as the CPU baseline.
 *
 * Parity pins (what checks each part; see tests/test_oracle_*.py):
 *   matrices       closed forms of gluLookAt / glFrustum / viewport
 *   clipping       plane invariants, fan count, exact-rational brute force
 *   Barycentric    reconstruction invariant, closed form
 *   coverage       exact-rational brute force on tiny triangles
 *   depth          formula identity recomputed in the test
 *   interpolation  constant-attribute invariance, b.W * sum == 1
 *   Phong          N==L, power 0, N.L<0 special cases
 *   bilinear       constant texture, hand-computed 2x2 textures
 *   blend          hand-computed integer cases
 *   counters       brute force
 *   DepthImage     closed form
 *   go_pow         exact rational power for integer exponents (<= 1 ulp)
 * Parity unpinned: go_pow for non-integer exponents (uses libm exp/log where
 * Go uses its own Exp/Log), Perspective's tan (libm vs Go's math.Tan).
 */
#ifndef FAUXGL_ORACLE_H
#define FAUXGL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_FACE_CW = 1, ORC_FACE_CCW = 2 };
enum { ORC_CULL_NONE = 1, ORC_CULL_FRONT = 2, ORC_CULL_BACK = 3 };
enum { ORC_SHADER_SOLID = 1, ORC_SHADER_TEXTURE = 2, ORC_SHADER_PHONG = 3 };
/* Texture storage kinds.  NRGBA8: image.NRGBA (non-premultiplied, At().RGBA()
 * premultiplies); RGBA8: image.RGBA (already premultiplied); RGBA16: the four
 * 16-bit values At(x,y).RGBA() returns, stored directly (exact for any Go
 * image type). */
enum { ORC_TEX_NRGBA8 = 1, ORC_TEX_RGBA8 = 2, ORC_TEX_RGBA16 = 3 };

typedef struct orc_texture {
    int32_t width, height, kind, pad;
    const void *pix; /* width*height*4 of uint8 (NRGBA8/RGBA8) or uint16 (RGBA16), row-major */
} orc_texture;

/* Flat shader description: the fields of SolidColorShader / TextureShader /
 * PhongShader (shader.go:11-59). */
typedef struct orc_shader {
    int32_t kind, pad;
    double matrix[16];     /* X00..X33 row-major */
    double color[4];       /* SolidColorShader.Color */
    double light_dir[3];   /* PhongShader.LightDirection */
    double camera_pos[3];  /* PhongShader.CameraPosition */
    double object_color[4], ambient[4], diffuse[4], specular[4];
    double specular_power;
    const orc_texture *texture; /* TextureShader.Texture / PhongShader.Texture (NULL = nil) */
} orc_shader;

typedef struct orc_state {
    int32_t read_depth, write_depth, write_color, alpha_blend, wireframe;
    int32_t front_face, cull, pad;
    double line_width, depth_bias;
} orc_state;

/* Instrumentation for the FP64 work model (not part of the reference). */
enum {
    ORC_CNT_INPUT = 0,    /* primitives entering DrawTriangle/DrawLine */
    ORC_CNT_RASTER,       /* rasterize() calls */
    ORC_CNT_ROWS,         /* row iterations */
    ORC_CNT_SCANNED,      /* x iterations */
    ORC_CNT_INSIDE,       /* == TotalPixels */
    ORC_CNT_SHADED,       /* fragment shader invocations */
    ORC_CNT_CLIPPED,      /* primitives that went through ClipTriangle/ClipLine */
    ORC_CNT_SKIPPED,      /* raster triangles rejected as pathological */
    ORC_CNT_N
};

typedef struct orc_ctx orc_ctx;

orc_ctx *orc_ctx_new(int32_t width, int32_t height);
void orc_ctx_free(orc_ctx *c);
void orc_set_state(orc_ctx *c, const orc_state *s);
void orc_set_shader(orc_ctx *c, const orc_shader *s);
void orc_clear_color(orc_ctx *c, const double rgba[4]); /* ClearColorBufferWith */
void orc_clear_depth(orc_ctx *c, double v);             /* ClearDepthBufferWith */
uint8_t *orc_color(orc_ctx *c);                         /* 4*W*H NRGBA */
double *orc_depth(orc_ctx *c);                          /* W*H */
void orc_depth_image(const orc_ctx *c, uint16_t *out);  /* DepthImage Gray16 values */
void orc_counters(const orc_ctx *c, uint64_t out[ORC_CNT_N]);
void orc_reset_counters(orc_ctx *c);

/* Triangle / line soups in SoA form: pos count*nv*3, nrm count*nv*3,
 * uv count*nv*2, col count*nv*4 (nv = 3 for triangles, 2 for lines).
 * NULL attribute arrays mean zeros.  Built once into the reference's
 * 136-byte Vertex structs (vertex.go:3-12), like a Go []*Triangle. */
typedef struct orc_mesh orc_mesh;
orc_mesh *orc_mesh_new(int64_t ntri, const double *tpos, const double *tnrm,
                       const double *tuv, const double *tcol,
                       int64_t nline, const double *lpos, const double *lnrm,
                       const double *luv, const double *lcol);
void orc_mesh_free(orc_mesh *m);

void orc_draw_triangles(orc_ctx *c, const orc_mesh *m, int32_t nthreads, uint64_t info[2]);
void orc_draw_lines(orc_ctx *c, const orc_mesh *m, int32_t nthreads, uint64_t info[2]);
void orc_draw_mesh(orc_ctx *c, const orc_mesh *m, int32_t nthreads, uint64_t info[2]);

/* ---- exported pieces, for the pinning tests ---- */
double orc_go_min(double a, double b);
double orc_go_max(double a, double b);
int64_t orc_go_f2i(double x);
double orc_go_pow(double x, double y);
uint8_t orc_go_u8(double x);    /* uint8(x): int64 conversion then low byte */
void orc_color_nrgba(const double c[4], uint8_t out[4]);               /* Color.NRGBA */
void orc_nrgba_rgba(const uint8_t c[4], uint32_t out[4]);              /* color.NRGBA.RGBA */
void orc_blend(uint8_t dst[4], const double color[4]);                 /* context.go:257-267 */
void orc_barycentric(const double p1[3], const double p2[3], const double p3[3],
                     const double p[3], double out[4]);
int32_t orc_clip_polygon(const double in[3][4], double out[9][4]);    /* sutherlandHodgman */
int32_t orc_clip_line(const double w1[4], const double w2[4], double out[2][4]);
void orc_matrix_mul(const double a[16], const double b[16], double out[16]);
void orc_mul_position_w(const double m[16], const double p[3], double out[4]);
void orc_look_at(const double eye[3], const double center[3], const double up[3], double out[16]);
void orc_frustum(double l, double r, double b, double t, double n, double f, double out[16]);
void orc_perspective(double fovy, double aspect, double near, double far, double out[16]);
void orc_screen(int32_t w, int32_t h, double out[16]);
void orc_bilinear_sample(const orc_texture *t, double u, double v, double out[4]);
/* Fragment of the Phong shader (shader.go:75-96) for one interpolated vertex. */
void orc_phong_fragment(const orc_shader *s, const double pos[3], const double nrm[3],
                        const double uv[2], const double col[4], double out[4]);
/* Rasterize one screen-space triangle with constant attributes through
 * rasterize() (context.go:151-281); returns TotalPixels/UpdatedPixels. */
/* SSAA resolve (the step after the path; the reference pipeline calls
 * nfnt/resize's Bilinear there, README.md:109, which is not vendored, so
 * this is the build's own filter definition, restated independently of
 * the GPU code).  src is sw x sh NRGBA, dst dw x dh NRGBA, sw = f*dw and
 * sh = f*dh.  filter 1 = tent, 2 = box.  Tent: separable; per output
 * coordinate o of an axis of scale s = src/dst, taps t = 0..2*ceil(s)-1 at
 * source index start+t with start = int(s*(o+0.5)-0.5) - ceil(s) + 1,
 * weight int16(max(0, 1 - |(s*(o+0.5)-0.5 - start - t) * min(1/s, 1)|) * 256),
 * indices clamped to [0, n-1]; the horizontal pass premultiplies
 * (c*a/255, integer) and both passes divide the weighted sum by the weight
 * sum with truncation.  Box: per channel round-half-up mean of the f x f
 * block.  orc_tent_weights exposes one axis' table. */
void orc_resolve(const uint8_t *src, int32_t sw, int32_t sh, uint8_t *dst, int32_t dw, int32_t dh, int32_t filter);
int32_t orc_tent_weights(int32_t src, int32_t dst, int16_t *w, int32_t *start);
void orc_rasterize_screen(orc_ctx *c, const double s0[3], const double s1[3],
                          const double s2[3], const double w[3], uint64_t info[2]);

/* Mesh preparation (mesh.go), on SoA vertex arrays (vertex v at [3v, 3v+3);
 * triangles have 3 vertices, lines 2).  Parity pins: tests/test_oracle_mesh.py
 * (a dict-based brute force of mesh.go's map code, the numpy implementation
 * in fauxgl_amd/meshes.py, closed forms and special cases). */
/* Mesh.BoundingBox (mesh.go:153-165; Box.Extend box.go:48-53). */
void orc_mesh_bounding_box(const double *tpos, int64_t ntri, const double *lpos, int64_t nline, double box[6]);
/* Triangle/Line.Transform (triangle.go:66-73, line.go:23-28) of nvert vertices. */
void orc_mesh_transform(double *pos, double *nrm, int64_t nvert, const double m[16]);
/* Mesh.FitInside (mesh.go:142-151): the matrix it builds, then Transform. */
void orc_mesh_fit_inside(double *tpos, double *tnrm, int64_t ntri, double *lpos, double *lnrm, int64_t nline,
                         const double box[6], const double anchor[3], double m_out[16]);
/* Mesh.SmoothNormals (mesh.go:105-125) and SmoothNormalsThreshold
 * (mesh.go:80-103, threshold = math.Cos(radians)) on triangle vertices. */
void orc_mesh_smooth_normals(const double *tpos, double *tnrm, int64_t ntri);
void orc_mesh_smooth_normals_threshold(const double *tpos, double *tnrm, int64_t ntri, double threshold);

/* Mesh loaders.  LoadSTL (stl.go:23-154): binary when the file size is
 * 84 + 50 * (uint32 at byte 80), else ASCII "vertex x y z" lines; face
 * normals (Triangle.Normal, triangle.go:33-37).  LoadOBJ (obj.go:19-79):
 * v / vt / vn / f with fan triangulation, 1-based and negative indices
 * (strconv.ParseInt base 0), FixNormals.  Positions / normals are ntri*9
 * doubles, uv ntri*6 (Texture.X, Y).  Call with NULL outputs to get ntri.
 * Returns 0, or -1 where the reference returns an error or panics (short
 * header, a vertex count not a multiple of 3, an index out of range, too
 * few numbers on a v / vt / vn line).  Floats follow strconv.ParseFloat
 * (decimal, inf / infinity / nan; a malformed token reads as 0). */
int orc_load_stl(const uint8_t *data, int64_t size, int64_t *ntri, double *pos, double *nrm);
int orc_load_obj(const char *text, int64_t size, int64_t *ntri, double *pos, double *nrm, double *uv);

/* SavePNG (util.go:56-63) goes through Go's image/png encoder (standard
 * library, not in the reference tree): for *image.NRGBA it writes RGB when
 * every alpha is 255, else RGBA, and filters each row with the filter that
 * minimises the sum of |byte| read as int8, trying Up, Paeth, None, Sub,
 * Average in that order and keeping the first minimum.  This writes the
 * filtered stream (per row: filter type byte, then the row) for nrgba
 * (w*h*4 bytes) into out (h * (w*bpp + 1) bytes) and returns bpp (3 or 4);
 * out NULL: only bpp. */
int32_t orc_png_filter(const uint8_t *nrgba, int32_t w, int32_t h, uint8_t *out);

#ifdef __cplusplus
}
#endif
#endif
