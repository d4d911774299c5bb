/*
 * skity_hip.h -- C ABI of the MI355X (gfx950) DrawPath backend.
 *
 * This is the drop-in boundary for skity's software rasteriser.  A skity
 * canvas draw (skity::Canvas::DrawPath, include/skity/render/canvas.hpp:365-424
 * and src/render/canvas.cc:237-239, dispatched to SWCanvas::OnDrawPath in
 * src/render/sw/sw_canvas.cc:357-411) becomes one path record in a
 * hipr_scene; hipr_surface_draw_scene rasterises a whole ordered list of such
 * draws onto a device-resident RGBA8 surface with hand-written HIP kernels.
 *
 * Conventions follow skity's own C wrapper (module/capi):
 *   - opaque handles `typedef struct X_s* X` (module/capi/include/skity_c/
 *     skity_base.h:34), destroy calls are safe on NULL;
 *   - result codes equal skity_result (skity_base.h:47-56), extended with
 *     device-side failures below -5;
 *   - no exceptions cross the boundary, symbols are exported with
 *     visibility("default") from a library built -fvisibility=hidden
 *     (module/capi/CMakeLists.txt:28-37);
 *   - the caller owns every host array until the call returns; the library
 *     owns every device buffer it allocates.
 *
 * Enumerations use the numeric values of the reference's C++ enums so that a
 * skity Path/Paint can be packed without translation tables.
 */
#ifndef SKITY_HIP_H_
#define SKITY_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HIPR_API __attribute__((visibility("default")))

/* skity_result values (module/capi/include/skity_c/skity_base.h:47-56). */
typedef enum hipr_result {
  HIPR_SUCCESS = 0,
  HIPR_ERROR_INVALID_HANDLE = -1,
  HIPR_ERROR_INVALID_ARGUMENT = -2,
  HIPR_ERROR_INITIALIZATION_FAILED = -3,
  HIPR_ERROR_OUT_OF_HOST_MEMORY = -4,
  HIPR_ERROR_NOT_SUPPORTED = -5,
  /* Extensions for the device backend. */
  HIPR_ERROR_DEVICE = -6,               /* a HIP runtime call failed        */
  HIPR_ERROR_OUT_OF_DEVICE_MEMORY = -7, /* device allocation failed         */
  HIPR_ERROR_CAPACITY = -8              /* a fixed device limit was exceeded */
} hipr_result;

/* Path::Verb (include/skity/graphic/path.hpp:45-60). */
enum { HIPR_VERB_MOVE = 0, HIPR_VERB_LINE = 1, HIPR_VERB_QUAD = 2,
       HIPR_VERB_CONIC = 3, HIPR_VERB_CUBIC = 4, HIPR_VERB_CLOSE = 5 };
/* Path::PathFillType (path.hpp:62-67). */
enum { HIPR_FILL_WINDING = 0, HIPR_FILL_EVEN_ODD = 1 };
/* Paint::Style / Cap / Join (include/skity/graphic/paint.hpp:46-105). */
enum { HIPR_STYLE_FILL = 0, HIPR_STYLE_STROKE = 1, HIPR_STYLE_STROKE_AND_FILL = 2,
       HIPR_STYLE_STROKE_THEN_FILL = 3 };
enum { HIPR_CAP_BUTT = 0, HIPR_CAP_ROUND = 1, HIPR_CAP_SQUARE = 2 };
enum { HIPR_JOIN_MITER = 0, HIPR_JOIN_ROUND = 1, HIPR_JOIN_BEVEL = 2 };
/* Shader::GradientType (include/skity/effect/shader.hpp:40-47). */
enum {
  HIPR_GRADIENT_LINEAR = 2,
  HIPR_GRADIENT_RADIAL = 3,
  HIPR_GRADIENT_SWEEP = 4,
  HIPR_GRADIENT_CONICAL = 5 /* two-point conical: points = start, end centres;
                               radius = start, end radii */
};
/* TileMode (include/skity/graphic/tile_mode.hpp:13-30). */
enum { HIPR_TILE_CLAMP = 0, HIPR_TILE_REPEAT = 1, HIPR_TILE_MIRROR = 2, HIPR_TILE_DECAL = 3 };
/* BlurStyle (include/skity/effect/mask_filter.hpp:15-20). */
enum { HIPR_BLUR_NORMAL = 1, HIPR_BLUR_SOLID = 2, HIPR_BLUR_OUTER = 3, HIPR_BLUR_INNER = 4 };
/* BlendMode (include/skity/graphic/blend_mode.hpp:45-80).  The software
 * backend blends kClear .. kScreen and kSoftLight; every other mode falls back
 * to kSrcOver in PorterDuffBlend (src/graphic/blend_mode.cc:125-134), and so
 * does the device. */
enum {
  HIPR_BLEND_CLEAR = 0, HIPR_BLEND_SRC = 1, HIPR_BLEND_DST = 2, HIPR_BLEND_SRC_OVER = 3,
  HIPR_BLEND_DST_OVER = 4, HIPR_BLEND_SRC_IN = 5, HIPR_BLEND_DST_IN = 6,
  HIPR_BLEND_SRC_OUT = 7, HIPR_BLEND_DST_OUT = 8, HIPR_BLEND_SRC_ATOP = 9,
  HIPR_BLEND_DST_ATOP = 10, HIPR_BLEND_XOR = 11, HIPR_BLEND_PLUS = 12,
  HIPR_BLEND_MODULATE = 13, HIPR_BLEND_SCREEN = 14, HIPR_BLEND_SOFT_LIGHT = 21,
  HIPR_BLEND_LAST = 28 /* kLuminosity */
};
/* AlphaType (include/skity/graphic/alpha_type.hpp:33-39) and ColorType
 * (include/skity/graphic/color_type.hpp:16-23) of a surface.  Targets that
 * are not premultiplied are converted on every blend
 * (SWRenderTarget::BlendPixel, src/render/sw/sw_render_target.cc:12-34). */
enum { HIPR_ALPHA_OPAQUE = 1, HIPR_ALPHA_PREMUL = 2, HIPR_ALPHA_UNPREMUL = 3 };
enum { HIPR_COLOR_RGBA = 1, HIPR_COLOR_BGRA = 2 };
/* What one entry of a hipr_scene draws: Canvas::DrawPath, or
 * Canvas::DrawPaint (SWCanvas::OnDrawPaint, src/render/sw/sw_canvas.cc:
 * 412-438: the paint over the whole target, whose entry has no verbs). */
enum { HIPR_DRAW_PATH = 0, HIPR_DRAW_PAINT = 1 };

#define HIPR_MAX_GRADIENT_STOPS 16

/* One Paint, as SWCanvas reads it (src/render/sw/sw_canvas.cc:357-411,727-795).
 * `color` and `stroke_color` are the Color4f fill and stroke colours the
 * reference Paint stores (Paint::SetColor sets both to Color4fFromColor(c),
 * src/graphic/paint.cc:48-84). */
typedef struct hipr_paint {
  float color[4];       /* fill colour r, g, b, a in [0,1] (unpremultiplied) */
  float stroke_color[4];/* stroke colour                                     */
  int32_t style;        /* HIPR_STYLE_*                                     */
  float stroke_width;   /* Paint::GetStrokeWidth                            */
  float miter_limit;    /* Paint::GetStrokeMiter (default 4)                */
  int32_t cap;          /* HIPR_CAP_*                                       */
  int32_t join;         /* HIPR_JOIN_*                                      */
  int32_t gradient;     /* index into hipr_scene.gradients, or -1           */
  float blur_radius;    /* MaskFilter::GetBlurRadius, 0 = no mask filter    */
  int32_t blur_style;   /* HIPR_BLUR_*                                      */
  int32_t blend;        /* HIPR_BLEND_*                                     */
} hipr_paint;

/* A gradient shader as Shader::AsGradient reports it
 * (include/skity/effect/shader.hpp:49-65). For sweep gradients radius[0] is
 * the bias and radius[1] the scale computed by Shader::MakeSweep
 * (src/effect/shader.cc:88-105). local_matrix is the affine shader local
 * matrix {scale_x, skew_x, trans_x, skew_y, scale_y, trans_y}. */
typedef struct hipr_gradient {
  int32_t type;        /* HIPR_GRADIENT_*                                    */
  int32_t tile_mode;   /* HIPR_TILE_*                                        */
  int32_t color_count; /* 2 .. HIPR_MAX_GRADIENT_STOPS                       */
  int32_t has_offsets; /* 0: evenly spaced colours, 1: `offsets` is used     */
  float points[4];     /* point[0].x, point[0].y, point[1].x, point[1].y     */
  float radius[2];
  float local_matrix[6];
  float colors[HIPR_MAX_GRADIENT_STOPS][4]; /* r, g, b, a                     */
  float offsets[HIPR_MAX_GRADIENT_STOPS];
} hipr_gradient;

/* An ordered list of DrawPath calls in structure-of-arrays form.  Path p owns
 * verbs [verb_offsets[p], verb_offsets[p+1]), points (x,y pairs)
 * [point_offsets[p], point_offsets[p+1]) and conic weights
 * [weight_offsets[p], weight_offsets[p+1]).  The verb/point streams must be
 * what skity's Path builder produces (every contour starts with a move).
 * `matrices` holds the canvas matrix of each draw (Canvas::GetTotalMatrix,
 * src/render/canvas.cc:126-128), applied as SWRaster::RastePath does after
 * the quad conversion (Path::CopyWithMatrix, src/graphic/path.cc:1274-1295).
 * The fields after n_gradients may be left zero (identity matrices, all
 * draws DrawPath). */
/* A paint's ColorFilter (include/skity/effect/color_filter.hpp:26-37), applied
 * to every source pixel after its cover, before the blend
 * (SWSpanBrush::BrushH, src/render/sw/sw_span_brush.cc:106-137). */
enum {
  HIPR_COLOR_FILTER_NONE = 0,
  HIPR_COLOR_FILTER_BLEND = 1,           /* ColorFilters::Blend(color, blend)   */
  HIPR_COLOR_FILTER_MATRIX = 2,          /* ColorFilters::Matrix(matrix)        */
  HIPR_COLOR_FILTER_LINEAR_TO_SRGB = 3,  /* ColorFilters::LinearToSRGBGamma()   */
  HIPR_COLOR_FILTER_SRGB_TO_LINEAR = 4,  /* ColorFilters::SRGBToLinearGamma()   */
  HIPR_COLOR_FILTER_COMPOSE = 5          /* ColorFilters::Compose(outer, inner):
                                            the software backend passes colours
                                            through it unchanged
                                            (color_filter.cc:182-185)          */
};
typedef struct hipr_color_filter {
  int32_t type;       /* HIPR_COLOR_FILTER_*                                  */
  uint32_t color;     /* BLEND: the Color, 0xAARRGGBB, unpremultiplied        */
  int32_t blend;      /* BLEND: HIPR_BLEND_*                                  */
  float matrix[20];   /* MATRIX: row-major 4x5, each entry with |255 m| < 2^15 */
} hipr_color_filter;

/* An Image's pixels (Image::MakeImage of a Pixmap, kRGBA_8888): rows of
 * width * 4 bytes, R, G, B, A; premultiplied (HIPR_ALPHA_PREMUL or
 * HIPR_ALPHA_OPAQUE) or not (HIPR_ALPHA_UNPREMUL). */
typedef struct hipr_image {
  int32_t width, height;
  int32_t alpha_type;     /* HIPR_ALPHA_*                                      */
  const uint8_t* pixels;
} hipr_image;

enum { HIPR_FILTER_NEAREST = 0, HIPR_FILTER_LINEAR = 1 };

/* A pixmap shader (Shader::MakeShader(image, sampling, tile_x, tile_y,
 * local_matrix), sampled by PixmapBrush / BitmapSampler,
 * src/render/sw/sw_span_brush.cc:555-579, src/graphic/bitmap_sampler.cc:
 * 26-108).  Cubic sampling is drawn as linear, as the software backend
 * does. */
typedef struct hipr_image_shader {
  int32_t image;          /* index into hipr_scene.images; -1 = no shader      */
  int32_t filter;         /* HIPR_FILTER_*                                     */
  int32_t tile_x, tile_y; /* HIPR_TILE_*                                       */
  float local_matrix[6];  /* {scale_x, skew_x, trans_x, skew_y, scale_y, trans_y} */
} hipr_image_shader;

typedef struct hipr_scene {
  int32_t n_paths;
  const uint8_t* verbs;
  const int32_t* verb_offsets;   /* n_paths + 1 */
  const float* points;           /* 2 floats per point */
  const int32_t* point_offsets;  /* n_paths + 1, in points */
  const float* weights;
  const int32_t* weight_offsets; /* n_paths + 1 */
  const uint8_t* fill_types;     /* n_paths, HIPR_FILL_* */
  const hipr_paint* paints;      /* n_paths */
  const hipr_gradient* gradients;
  int32_t n_gradients;
  /* n_paths x 6 floats {scale_x, skew_x, trans_x, skew_y, scale_y, trans_y}:
   * x' = (scale_x * x + skew_x * y) + trans_x,
   * y' = (skew_y * x + scale_y * y) + trans_y; NULL = identity. */
  const float* matrices;
  const uint8_t* draw_kinds;     /* n_paths, HIPR_DRAW_*; NULL = all paths */
  const hipr_color_filter* color_filters; /* n_paths; NULL = none            */
  const hipr_image* images;
  int32_t n_images;
  /* n_paths; NULL = none.  A draw with an image shader has no gradient. */
  const hipr_image_shader* image_shaders;
  /* n_paths x 4 floats {left, top, right, bottom}: the canvas's global clip
   * bounds at each draw in device space (Canvas::ClipRect intersections,
   * canvas.cc:130-135,463-476), the scan clip of the draw's path, stroke and
   * blurred-layer rasters (SWRaster::RastePath, sw_raster.cc:731-786;
   * DrawPaint rasters without it, sw_canvas.cc:413-439).  NULL = none; a
   * draw without a clip holds kMaxCullRect {-1e9, -1e9, 1e9, 1e9}. */
  const float* clip_rects;
} hipr_scene;

typedef struct hipr_context_s* hipr_context;
typedef struct hipr_surface_s* hipr_surface;
typedef struct hipr_prepared_s* hipr_prepared; /* see hipr_scene_prepare */

/* Per-draw statistics of the last hipr_surface_draw_scene call. */
typedef struct hipr_stats {
  int64_t paths;          /* draws processed                                */
  int64_t raster_paths;   /* fill/stroke rasterisations (strokes count once)*/
  int64_t edges;          /* line + quad edges built (after culling)        */
  int64_t tiles;          /* 16x16 coverage tiles written                   */
  int64_t tile_planes;    /* coverage planes written (>= tiles)             */
  int64_t fallbacks;      /* paths re-run by the exact two-pass method      */
  int64_t max_active;     /* largest active edge list seen                  */
  double ms_total;        /* device time of the whole draw                  */
  int64_t blends;         /* pixel blend operations on the target (F)       */
  int64_t pixels;         /* distinct target pixels touched (U)             */
  int64_t underestimates; /* estimate modes: paths whose tile planes
                             exceeded their estimate (served from the shared
                             overflow pool, or re-run when it ran out)      */
} hipr_stats;

/* Execution options.  mode 0 = plain exact two-pass (count, scan, write);
 * mode 1 = pre-sized single pass from the caller-supplied per-path estimates
 * (estimates != NULL), falling back to the exact two-pass method for every
 * path whose estimate was too small; mode 2 = pre-sized from the context's
 * learned pre-sizer (hipr_context_set_presizer), with the same fallback. */
enum { HIPR_MODE_EXACT = 0, HIPR_MODE_ESTIMATES = 1, HIPR_MODE_LEARNED = 2 };
/* StackBlur implementations; both are exact (identical output). */
enum { HIPR_BLUR_IMPL_MFMA = 0,  /* banded-matrix products on MFMA (fp16, fp32 sums) */
       HIPR_BLUR_IMPL_LINES = 1  /* one thread per line, integer running sums    */ };
typedef struct hipr_options {
  int32_t mode;
  const int32_t* edge_estimates;  /* n_paths or NULL */
  const int32_t* tile_estimates;  /* n_paths or NULL */
  int32_t blur_impl;              /* HIPR_BLUR_IMPL_*; 0 = MFMA */
  int32_t band_rows;              /* rows per walk band: 0 = default (128
                                     exact, 64 estimate modes), < 0 = one
                                     band per path */
  int32_t band_min_edges;         /* estimate modes: band only the paths whose
                                     active-edge bound reaches this; 0 = default
                                     (257: the 8- and 16-slot walks), < 0 =
                                     every path, a huge value = none */
  int32_t flags;                  /* HIPR_FLAG_* */
  int32_t chunk_points;           /* a draw of more source points than this
                                     (or of more bounding-box tiles than four
                                     times it) runs as consecutive sub-plans
                                     of whole draws, in order: 0 = 33,554,432.
                                     Draws with mask filters run whole. */
} hipr_options;
enum {
  HIPR_FLAG_NO_GRAPH = 1,   /* launch the walk passes directly, not as hipGraphs */
  HIPR_FLAG_PERSISTENT = 2  /* persistent walk kernels on an atomic work queue   */
};

/* --- context / surface ---------------------------------------------------
 * Replaces Canvas::MakeSoftwareCanvas(Bitmap*) (src/render/sw/sw_canvas.cc:
 * 146-156) and the GPUSurface/LockCanvas pattern
 * (include/skity/gpu/gpu_surface.hpp:95-113). */
HIPR_API hipr_result hipr_context_create(int32_t device, hipr_context* out);
/* Destroy calls may come in any order: a context's device resources are
 * released once its handle and every surface and prepared scene made from
 * it have been destroyed. */
HIPR_API void hipr_context_destroy(hipr_context ctx);
/* A premultiplied RGBA8 target (bytes R,G,B,A), cleared to transparent.
 * width or height 0 -> HIPR_ERROR_INVALID_ARGUMENT, mirroring
 * MakeSoftwareCanvas returning null for an empty bitmap. */
HIPR_API hipr_result hipr_surface_create(hipr_context ctx, int32_t width,
                                         int32_t height, hipr_surface* out);
/* A target of the given HIPR_ALPHA_* and HIPR_COLOR_* types, the two
 * properties of skity's Bitmap(width, height, alpha_type, color_type)
 * (include/skity/graphic/bitmap.hpp:27-31, whose default is kUnpremul /
 * kRGBA).  Pixels are 4 bytes, in the colour type's byte order; read and
 * write calls move them as stored. */
HIPR_API hipr_result hipr_surface_create_ex(hipr_context ctx, int32_t width, int32_t height,
                                            int32_t alpha_type, int32_t color_type,
                                            hipr_surface* out);
HIPR_API void hipr_surface_destroy(hipr_surface surface);
HIPR_API hipr_result hipr_surface_clear(hipr_surface surface, uint32_t rgba);
/* Copy the surface (width*height*4 bytes) to host memory.
 * Replaces GPUSurface::ReadPixels (gpu_surface.hpp:110). */
HIPR_API hipr_result hipr_surface_read_pixels(hipr_surface surface,
                                              uint8_t* dst, size_t dst_bytes);
HIPR_API hipr_result hipr_surface_write_pixels(hipr_surface surface,
                                               const uint8_t* src,
                                               size_t src_bytes);

/* --- drawing --------------------------------------------------------------
 * Draw every path of `scene` in order, exactly as consecutive
 * Canvas::DrawPath calls on a software canvas would (sw_canvas.cc:357-411).
 * Asynchronous on the surface's stream; hipr_surface_finish waits. */
HIPR_API hipr_result hipr_surface_draw_scene(hipr_surface surface,
                                             const hipr_scene* scene,
                                             const hipr_options* options);
HIPR_API hipr_result hipr_surface_finish(hipr_surface surface);
HIPR_API hipr_result hipr_surface_get_stats(hipr_surface surface,
                                            hipr_stats* out);

/* --- learned pre-sizer ---------------------------------------------------
 * No reference counterpart: the software backend sizes its edge and span
 * vectors by growing them.  A 16-64-64-2 MLP (FP16 weights, FP32
 * accumulation on MFMA) predicts per raster job log1p(edges) and
 * log1p(tile planes) from 16 path features; its forward and backward passes
 * and its Adam training loop run on the device.  Parameters are
 * HIPR_PRESIZER_PARAMS floats: W1[16][64] b1[64] W2[64][64] b2[64] W3[64][2]
 * b3[2] (trained), then feature mean[16] and scale[16], target mean[2] and
 * scale[2] (fixed normalisation: x' = (x - mean) * scale,
 * y' = (y - mean) / scale). */
#define HIPR_PRESIZER_FEATURES 16
#define HIPR_PRESIZER_TRAINED 5378
#define HIPR_PRESIZER_PARAMS 5414
/* Raster jobs of a prepared scene (fills, strokes and layer composites). */
HIPR_API hipr_result hipr_prepared_job_count(hipr_prepared prepared, int32_t* n_jobs);
/* Per-job features (n_jobs x 16 floats, zero rows for layer composites) for
 * a draw onto a width x height target. */
HIPR_API hipr_result hipr_prepared_features(hipr_prepared prepared, int32_t width,
                                            int32_t height, float* features,
                                            int32_t n_jobs);
/* Exact per-job counts of the last draw on the surface (-1 for layer
 * composites): the labels of the pre-sizer. */
HIPR_API hipr_result hipr_surface_get_job_counts(hipr_surface surface, int32_t* edges,
                                                 int32_t* planes, int32_t n_jobs);
/* Walk profile of the last draw per job: sub-steps taken, the walk's wall
 * time and (phases, 5 per job) its split into roles + trapezoids, quad
 * updates, re-sort, row flushes and edge insertion, in 10 ns units, and the
 * largest active edge list (last write pass).  Any pointer may be NULL.
 * No reference counterpart. */
HIPR_API hipr_result hipr_surface_get_job_profile(hipr_surface surface, int32_t* steps,
                                                  int32_t* time_10ns, int32_t* phases,
                                                  int32_t* max_active, int32_t n_jobs);
/* Raster jobs of the last draw on the surface, in draw order: the DrawPath
 * call each comes from, its kind (0 fill, 1 stroke outline, 2 composite of a
 * blurred mask-filter layer) and its canvas (0 the surface, >= 1 the layers).
 * *n_jobs receives the count; at most `cap` entries are written to each
 * non-NULL array.  No reference counterpart: SWCanvas::OnDrawPath
 * (src/render/sw/sw_canvas.cc:357-411) makes the same rasterisations one
 * after another without naming them. */
HIPR_API hipr_result hipr_surface_get_jobs(hipr_surface surface, int32_t* path,
                                          int32_t* kind, int32_t* canvas, int32_t cap,
                                          int32_t* n_jobs);
/* Edge records of one job of the last draw, in emission order, 16 int32 each
 * (the hipr_debug_build_edges layout): the edges SWEdgeBuilder::BuildEdges
 * (src/render/sw/sw_edge.cc:299-336) builds for the fill or the stroke outline.
 * *n_edges receives the count; edges may be NULL to query it.  Valid until
 * the next draw on the surface's context (then HIPR_ERROR_INVALID_ARGUMENT). */
HIPR_API hipr_result hipr_surface_get_job_edges(hipr_surface surface, int32_t job,
                                               int32_t* edges, int64_t cap,
                                               int64_t* n_edges);
/* Tile bins of the last draw: for every job, the 16x16 tiles of its canvas
 * that hold coverage, as codes ((tile_y * tiles_x + tile_x) << 1 | plane),
 * ascending; offsets[j]..offsets[j+1] (n_jobs + 1 entries) delimit job j.
 * Plane 0 holds each pixel's first non-zero cover, plane 1 its accumulated
 * cover where a directly emitted span came first (the two span kinds of
 * RealSpanBuilder / SpanBuilder, src/render/sw/sw_raster.cc:19-136).  The
 * software backend has spans, not tiles; the bins are this integer function
 * of its spans.  bins may be NULL to query the sizes.  Valid until the next
 * draw on the surface's context. */
HIPR_API hipr_result hipr_surface_get_tile_bins(hipr_surface surface, int64_t* offsets,
                                               int32_t* bins, int64_t cap, int32_t n_jobs);
/* Predictions (n x 2, log1p units) for host-side features. */
HIPR_API hipr_result hipr_presizer_predict(hipr_context ctx, const float* params,
                                           const float* features, int64_t n, float* out);
/* Gradient of the mean squared error (normalised target space) over n
 * samples (HIPR_PRESIZER_TRAINED floats) and the mean loss. */
HIPR_API hipr_result hipr_presizer_gradients(hipr_context ctx, const float* params,
                                             const float* features, const float* targets,
                                             int64_t n, float* grads, float* loss);
/* Adam training on the device: `steps` mini-batches of `batch` samples drawn
 * in a seeded shuffled order; params is updated in place; loss_curve (steps
 * floats, may be NULL) receives each batch's mean loss. */
HIPR_API hipr_result hipr_presizer_train(hipr_context ctx, float* params,
                                         const float* features, const float* targets,
                                         int64_t n, int32_t steps, int32_t batch, float lr,
                                         uint64_t seed, float* loss_curve);
/* Installs the parameters used by HIPR_MODE_LEARNED.  Capacities are
 * ceil(expm1(prediction + margin)) per job and count. */
HIPR_API hipr_result hipr_context_set_presizer(hipr_context ctx, const float* params,
                                               float margin_edges, float margin_planes);

/* Device time of each pipeline stage of the last draw, measured with HIP
 * events recorded on the surface's stream around the stage's kernels
 * (ms summed over the stage's launches; launches counts them).  No
 * reference counterpart: the software backend has no stages to time. */
typedef enum hipr_stage {
  HIPR_STAGE_STROKE = 0,        /* outline count + write                  */
  HIPR_STAGE_FLATTEN_COUNT = 1, /* edge count pass                        */
  HIPR_STAGE_FLATTEN = 2,       /* edge write pass                        */
  HIPR_STAGE_SORT = 3,          /* per-path edge sort + gather            */
  HIPR_STAGE_WALK_COUNT = 4,    /* coverage count pass                    */
  HIPR_STAGE_WALK = 5,          /* coverage write pass (tile planes)      */
  HIPR_STAGE_LAYERS = 6,        /* mask-filter layers composite           */
  HIPR_STAGE_BLUR = 7,          /* StackBlur of the layers                */
  HIPR_STAGE_COMPOSITE = 8,     /* tile composite onto the target         */
  HIPR_STAGE_PRESIZE = 9,       /* features + MLP inference               */
  HIPR_STAGE_WALK_SCHEDULE = 10, /* step schedule pass (band snapshots)    */
  HIPR_STAGE_COUNT = 11
} hipr_stage;
HIPR_API hipr_result hipr_surface_get_stage_times(hipr_surface surface,
                                                  double* ms,
                                                  int32_t* launches,
                                                  int32_t n_stages);

/* A validated scene whose geometry, paints and draw plan are resident in
 * device memory, so that it can be drawn repeatedly without host-to-device
 * copies (the recorded DisplayList of include/skity/recorder/
 * display_list.hpp:56-77 played onto a HIP surface).  The scene arrays may
 * be released after hipr_scene_prepare returns. */
HIPR_API hipr_result hipr_scene_prepare(hipr_context ctx,
                                        const hipr_scene* scene,
                                        hipr_prepared* out);
HIPR_API void hipr_prepared_destroy(hipr_prepared prepared);
HIPR_API hipr_result hipr_surface_draw_prepared(hipr_surface surface,
                                                hipr_prepared prepared,
                                                const hipr_options* options);

/* Batched form: many canvases, each with its own recorded scene, drawn by
 * one pipeline run (one launch per stage for all of them).  Replaces a loop
 * of per-canvas Canvas::Flush calls (canvas.hpp:365-424) over independent
 * canvases.  A surface appears at most once; the surfaces and scenes added
 * must stay alive until the batch is cleared or destroyed.  Path indices of
 * the HIPR_MODE_ESTIMATES arrays run over the pairs in order. */
typedef struct hipr_batch_s* hipr_batch;
HIPR_API hipr_result hipr_batch_create(hipr_context ctx, hipr_batch* out);
HIPR_API void hipr_batch_destroy(hipr_batch batch);
HIPR_API hipr_result hipr_batch_add(hipr_batch batch, hipr_surface surface,
                                    hipr_prepared scene);
HIPR_API hipr_result hipr_batch_clear(hipr_batch batch);
HIPR_API hipr_result hipr_batch_draw(hipr_batch batch, const hipr_options* options);
/* The work split used when options leave band_rows / band_min_edges /
 * flags at 0. */
HIPR_API hipr_result hipr_batch_set_split(hipr_batch batch, int32_t band_rows,
                                          int32_t band_min_edges, int32_t flags);
/* Autotuner: draws the batch with each candidate work split (rows per walk
 * band, banded paths, persistent or dispatched walk kernels) in `mode`
 * (EXACT or LEARNED), keeps the fastest by device time and returns it.  The
 * targets' contents are left undefined (clear them before drawing). */
HIPR_API hipr_result hipr_batch_autotune(hipr_batch batch, int32_t mode, int32_t repeats,
                                         int32_t* out_band_rows,
                                         int32_t* out_band_min_edges, int32_t* out_flags);
HIPR_API hipr_result hipr_batch_get_stats(hipr_batch batch, hipr_stats* out);
HIPR_API hipr_result hipr_batch_get_stage_times(hipr_batch batch, double* ms,
                                                int32_t* launches, int32_t n_stages);

/* --- stage entry points used by the parity tests -------------------------
 * Stage (a): flatten + edge build (Stroke::QuadPath, src/geometry/
 * stroke.cc:914-962 and SWEdgeBuilder::BuildEdges, src/render/sw/
 * sw_edge.cc:299-336) for every path of `scene`, returning the edge records
 * in emission order and per-path edge offsets (n_paths + 1).  An edge is 16
 * int32: x, snapped_x, dx, dy, upper_x, upper_y, lower_y (16.16 fixed
 * point), packed (winding | curve count << 8 | curve shift << 16 | quad <<
 * 24), then the quad state qx, qy, qdx, qdy, qddx, qddy, q_last_x, q_last_y. */
HIPR_API hipr_result hipr_debug_build_edges(hipr_context ctx,
                                            const hipr_scene* scene,
                                            int32_t* edge_offsets,
                                            int32_t* edges, int64_t edge_cap);

/* StackBlur on a device-resident premultiplied RGBA8 image
 * (SWStackBlur::Blur, src/render/sw/sw_stack_blur.cc:18-284): radius is
 * round(MaskFilter radius), clamped to 254; radius <= 1 copies. */
/* Self-check of the wavefront primitives the kernels use (DPP scans and
 * shifts against their shuffle forms): for every wave of 64 inputs writes
 * 6 x 64 outputs (DPP sum scan, shuffle sum scan, DPP max scan, shuffle max
 * scan, shift up with fill -7, shift down with fill -9).  No reference
 * counterpart. */
HIPR_API hipr_result hipr_debug_wave_ops(hipr_context ctx, const int32_t* in, int32_t* out,
                                         int32_t n_waves);

/* StackBlur with a chosen implementation (HIPR_BLUR_IMPL_*). */
HIPR_API hipr_result hipr_stackblur_ex(hipr_context ctx, const uint8_t* src_host,
                                       uint8_t* dst_host, int32_t width, int32_t height,
                                       int32_t radius, int32_t impl);
HIPR_API hipr_result hipr_stackblur(hipr_context ctx, const uint8_t* src_host,
                                    uint8_t* dst_host, int32_t width,
                                    int32_t height, int32_t radius);
/* hipr_stackblur_ex with HIPR_BLUR_IMPL_MFMA, on device-resident images, that
 * allocates nothing.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  src and dst are device pointers
 * to width*height*4 bytes (premultiplied RGBA8) that do not overlap; the
 * intermediate of the two passes lives in one static device buffer per GPU,
 * so width*height is at most 4,194,304 (else HIPR_ERROR_CAPACITY), and calls
 * that overlap on one GPU race.  Returns when the blur is done. */
HIPR_API hipr_result hipr_stackblur_device_noalloc(hipr_context ctx, const uint8_t* src,
                                                   uint8_t* dst, int32_t width,
                                                   int32_t height, int32_t radius);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif  /* SKITY_HIP_H_ */
