/* oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * C entry points of the CPU restatement of skity's software DrawPath path.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, as the checker.  The product library never links it.
 */
#ifndef ORACLE_H_
#define ORACLE_H_

#include <stdint.h>

#include "../include/skity_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_API __attribute__((visibility("default")))

/* Draw every path of the scene, in order, onto a premultiplied RGBA8 image
 * (the caller initialises it), as consecutive SWCanvas::DrawPath calls. */
ORC_API int orc_render_scene(const hipr_scene* scene, uint8_t* pixels,
                             int32_t width, int32_t height);

/* orc_render_scene with the draws rasterised on `threads` threads and their
 * spans applied in draw order per row band: the same image. */
ORC_API int orc_render_scene_mt(const hipr_scene* scene, uint8_t* pixels,
                                int32_t width, int32_t height, int32_t threads);

/* orc_render_scene onto a target of the given HIPR_ALPHA_* / HIPR_COLOR_*
 * types (bytes in the colour type's order), on `threads` threads. */
ORC_API int orc_render_scene_ex(const hipr_scene* scene, uint8_t* pixels, int32_t width,
                                int32_t height, int32_t alpha_type, int32_t color_type,
                                int32_t threads);

/* Edge records of path p after Stroke::QuadPath + SWEdgeBuilder::BuildEdges,
 * 16 int32 per edge in the hipr_debug_build_edges layout.  stroke != 0 builds
 * the edges of the stroke outline instead of the fill.  Returns the number of
 * edges (even when it exceeds cap; only cap are written). */
ORC_API int64_t orc_path_edges(const hipr_scene* scene, int32_t p,
                               int32_t stroke, int32_t* out, int64_t cap);

/* Spans {x, y, len, cover} of path p in emission order; bounds receives the
 * floored/ceiled path bounds (l, t, r, b).  Returns the span count. */
ORC_API int64_t orc_path_spans(const hipr_scene* scene, int32_t p,
                               int32_t stroke, int32_t* out, int64_t cap,
                               float* bounds);

/* Tile bins of one raster (fill, or stroke outline if stroke != 0) of path p
 * on a width x height canvas: 16x16 tiles, plane 0 = each pixel's first
 * non-zero cover, plane 1 = its second (the accumulated cover after a direct
 * one); codes ((tile_y * tiles_x + tile_x) << 1 | plane), ascending.
 * *stacked receives the pixels with two direct covers (expected 0).
 * Returns the bin count (only cap are written). */
ORC_API int64_t orc_path_tile_bins(const hipr_scene* scene, int32_t p, int32_t stroke,
                                   int32_t width, int32_t height, int32_t* out,
                                   int64_t cap, int64_t* stacked);

/* Number of quads Cubic::ToQuads produces for one cubic. */
ORC_API int32_t orc_cubic_quad_count(const float* pts8);

/* Quad-converted path of p (Stroke::QuadPath): verbs and points. */
ORC_API int64_t orc_quad_path(const hipr_scene* scene, int32_t p,
                              uint8_t* verbs, int64_t verb_cap, float* pts,
                              int64_t pt_cap, int64_t* n_pts);

/* Stroke outline of path p (Stroke::StrokePath of its quad path with the
 * draw's paint): verbs, points, conic weights; counts = {verbs, points,
 * weights}.  Returns the verb count. */
ORC_API int64_t orc_stroke_outline(const hipr_scene* scene, int32_t p,
                                   uint8_t* verbs, int64_t verb_cap, float* pts,
                                   int64_t pt_cap, float* weights, int64_t w_cap,
                                   int64_t* counts);

/* StackBlur of a premultiplied RGBA8 image (radius = round(r), clamp 254). */
ORC_API void orc_stackblur(const uint8_t* src, uint8_t* dst, int32_t w,
                           int32_t h, int32_t radius);
ORC_API void orc_stackblur_tables(int32_t r, uint32_t* mul, uint32_t* shr);

/* Premultiplied colour of gradient g at device pixel (x, y), ARGB. */
ORC_API uint32_t orc_gradient_color(const hipr_gradient* g, int32_t x,
                                    int32_t y);

/* The blurred premultiplied layer of a mask-filter draw and its filter
 * bounds (l, t, r, b).  Returns W*H*4 (bytes needed). */
ORC_API int64_t orc_blur_layer(const hipr_scene* scene, int32_t p,
                               uint8_t* out, int64_t cap, int32_t* bounds);

#ifdef __cplusplus
}
#endif

#endif /* ORACLE_H_ */
