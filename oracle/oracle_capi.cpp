// oracle_capi.cpp -- TEST INFRASTRUCTURE ONLY: C wrappers of the oracle.
#include <algorithm>
#include <cstring>
#include <vector>

#include "oracle.h"
#include "oracle_raster.hpp"

namespace orc {
void RenderScene(const hipr_scene* s, uint8_t* px, int w, int h);
void RenderSceneParallel(const hipr_scene* s, uint8_t* px, int w, int h, int threads);
void RenderSceneEx(const hipr_scene* s, uint8_t* px, int w, int h, int alpha_type,
                   int color_type, int threads);
void PathSpans(const hipr_scene* s, int p, int stroke, std::vector<Span>* spans,
               float bounds[4], std::vector<Edge>* edges);
void StackBlur(const uint8_t* src, uint8_t* dst, int w, int h, int radius);
void StackBlurTablesPublic(int r, uint32_t* mul, uint32_t* shr);
uint32_t GradientColorAt(const hipr_gradient& g, int x, int y);
void BlurredLayer(const hipr_scene* s, int p, std::vector<uint8_t>* blurred,
                  int fb[4]);
int CubicQuadCountPublic(const Pt p[4]);
PathData StrokeOutline(const PathData& quad_path, const hipr_paint& paint);
}  // namespace orc

using namespace orc;

// 16-int edge layout shared with hipr_debug_build_edges:
//  0 x, 1 snapped_x, 2 dx, 3 dy, 4 upper_x, 5 upper_y, 6 lower_y,
//  7 winding&0xFF | curve_count<<8 | curve_shift<<16 | is_quad<<24,
//  8 qx, 9 qy, 10 qdx, 11 qdy, 12 qddx, 13 qddy, 14 q_last_x, 15 q_last_y.
static void PackEdge(const Edge& e, int32_t* o) {
  o[0] = e.x;
  o[1] = e.is_quad ? e.snapped_x : 0;
  o[2] = e.dx;
  o[3] = e.dy;
  o[4] = e.upper_x;
  o[5] = e.upper_y;
  o[6] = e.lower_y;
  o[7] = (int32_t)((uint32_t)(uint8_t)e.winding |
                   ((uint32_t)(uint8_t)e.curve_count << 8) |
                   ((uint32_t)e.curve_shift << 16) |
                   ((uint32_t)(e.is_quad ? 1 : 0) << 24));
  if (e.is_quad) {
    o[8] = e.qx;
    o[9] = e.qy;
    o[10] = e.qdx;
    o[11] = e.qdy;
    o[12] = e.qddx;
    o[13] = e.qddy;
    o[14] = e.q_last_x;
    o[15] = e.q_last_y;
  } else {
    for (int i = 8; i < 16; i++) o[i] = 0;
  }
}

extern "C" {

int orc_render_scene(const hipr_scene* scene, uint8_t* pixels, int32_t width,
                     int32_t height) {
  if (!scene || !pixels || width <= 0 || height <= 0) return -2;
  RenderScene(scene, pixels, width, height);
  return 0;
}

int orc_render_scene_mt(const hipr_scene* scene, uint8_t* pixels, int32_t width,
                        int32_t height, int32_t threads) {
  if (!scene || !pixels || width <= 0 || height <= 0) return -2;
  RenderSceneParallel(scene, pixels, width, height, threads);
  return 0;
}

int orc_render_scene_ex(const hipr_scene* scene, uint8_t* pixels, int32_t width,
                        int32_t height, int32_t alpha_type, int32_t color_type,
                        int32_t threads) {
  if (!scene || !pixels || width <= 0 || height <= 0) return -2;
  if (alpha_type < HIPR_ALPHA_OPAQUE || alpha_type > HIPR_ALPHA_UNPREMUL) return -2;
  if (color_type != HIPR_COLOR_RGBA && color_type != HIPR_COLOR_BGRA) return -2;
  RenderSceneEx(scene, pixels, width, height, alpha_type, color_type, threads);
  return 0;
}

int64_t orc_path_edges(const hipr_scene* scene, int32_t p, int32_t stroke,
                       int32_t* out, int64_t cap) {
  std::vector<Span> spans;
  std::vector<Edge> edges;
  float b[4];
  PathSpans(scene, p, stroke, &spans, b, &edges);
  for (int64_t i = 0; i < (int64_t)edges.size() && i < cap; i++)
    PackEdge(edges[i], out + 16 * i);
  return (int64_t)edges.size();
}

int64_t orc_path_spans(const hipr_scene* scene, int32_t p, int32_t stroke,
                       int32_t* out, int64_t cap, float* bounds) {
  std::vector<Span> spans;
  float b[4] = {0, 0, 0, 0};
  PathSpans(scene, p, stroke, &spans, b, nullptr);
  if (bounds) std::memcpy(bounds, b, sizeof(b));
  for (int64_t i = 0; i < (int64_t)spans.size() && i < cap; i++) {
    out[4 * i + 0] = spans[i].x;
    out[4 * i + 1] = spans[i].y;
    out[4 * i + 2] = spans[i].len;
    out[4 * i + 3] = spans[i].cover;
  }
  return (int64_t)spans.size();
}

// Tile bins of one raster of path p on a width x height canvas (the
// definition the device's 16x16 coverage tiles follow).  Per visible pixel,
// the non-zero covers the reference's span builders emit for it, in
// emission order (RealSpanBuilder spans first, then the row accumulator's
// flush, sw_raster.cc:19-136); spans of cover 0 blend nothing and are left
// out.  Plane 0 of a tile holds the first cover of each pixel, plane 1 the
// second.  A tile's bin (plane k) exists iff some pixel of the tile has a
// k-th cover.  Codes: ((tile_y * tiles_x + tile_x) << 1 | plane), ascending.
// *stacked receives the number of pixels with more than one direct cover
// (the device stores one direct cover per pixel; the reference never emits
// two for one pixel, since full-row spans of different intervals are
// disjoint and too-close edges are routed to the accumulator).
int64_t orc_path_tile_bins(const hipr_scene* scene, int32_t p, int32_t stroke,
                           int32_t width, int32_t height, int32_t* out, int64_t cap,
                           int64_t* stacked) {
  std::vector<Span> spans;
  float b[4];
  PathSpans(scene, p, stroke, &spans, b, nullptr);
  const int tiles_x = (width + 15) / 16;
  std::vector<uint8_t> direct_count((size_t)width * height, 0);
  std::vector<uint8_t> has_acc((size_t)width * height, 0);
  int64_t n_stacked = 0;
  for (const Span& s : spans) {
    if (s.cover == 0 || s.y < 0 || s.y >= height) continue;
    int x0 = std::max(s.x, 0), x1 = std::min(s.x + s.len, width);
    for (int x = x0; x < x1; x++) {
      size_t i = (size_t)s.y * width + x;
      if (s.accumulated) {
        has_acc[i] = 1;
      } else {
        if (direct_count[i] == 1) n_stacked++;
        if (direct_count[i] < 2) direct_count[i]++;
      }
    }
  }
  std::vector<int32_t> codes;
  for (int y = 0; y < height; y++) {
    for (int x = 0; x < width; x++) {
      size_t i = (size_t)y * width + x;
      int covers = (direct_count[i] ? 1 : 0) + (has_acc[i] ? 1 : 0);
      int tile = (y >> 4) * tiles_x + (x >> 4);
      for (int k = 0; k < covers; k++) codes.push_back(tile << 1 | k);
    }
  }
  std::sort(codes.begin(), codes.end());
  codes.erase(std::unique(codes.begin(), codes.end()), codes.end());
  if (stacked) *stacked = n_stacked;
  for (int64_t i = 0; i < (int64_t)codes.size() && i < cap; i++) out[i] = codes[i];
  return (int64_t)codes.size();
}

int32_t orc_cubic_quad_count(const float* pts8) {
  Pt p[4];
  for (int i = 0; i < 4; i++) p[i] = Pt{pts8[2 * i], pts8[2 * i + 1]};
  return CubicQuadCountPublic(p);
}

int64_t orc_quad_path(const hipr_scene* scene, int32_t p, uint8_t* verbs,
                      int64_t verb_cap, float* pts, int64_t pt_cap,
                      int64_t* n_pts) {
  PathData q = QuadPath(ScenePath(scene, p));
  for (int64_t i = 0; i < (int64_t)q.verbs.size() && i < verb_cap; i++)
    verbs[i] = q.verbs[i];
  for (int64_t i = 0; i < (int64_t)q.pts.size() && i < pt_cap; i++) {
    pts[2 * i] = q.pts[i].x;
    pts[2 * i + 1] = q.pts[i].y;
  }
  if (n_pts) *n_pts = (int64_t)q.pts.size();
  return (int64_t)q.verbs.size();
}

int64_t orc_stroke_outline(const hipr_scene* scene, int32_t p, uint8_t* verbs,
                           int64_t verb_cap, float* pts, int64_t pt_cap, float* weights,
                           int64_t w_cap, int64_t* counts) {
  PathData o = StrokeOutline(QuadPath(ScenePath(scene, p)), scene->paints[p]);
  for (int64_t i = 0; i < (int64_t)o.verbs.size() && i < verb_cap; i++) verbs[i] = o.verbs[i];
  for (int64_t i = 0; i < (int64_t)o.pts.size() && i < pt_cap; i++) {
    pts[2 * i] = o.pts[i].x;
    pts[2 * i + 1] = o.pts[i].y;
  }
  for (int64_t i = 0; i < (int64_t)o.weights.size() && i < w_cap; i++) weights[i] = o.weights[i];
  if (counts) {
    counts[0] = (int64_t)o.verbs.size();
    counts[1] = (int64_t)o.pts.size();
    counts[2] = (int64_t)o.weights.size();
  }
  return (int64_t)o.verbs.size();
}

void orc_stackblur(const uint8_t* src, uint8_t* dst, int32_t w, int32_t h,
                   int32_t radius) {
  StackBlur(src, dst, w, h, radius);
}

void orc_stackblur_tables(int32_t r, uint32_t* mul, uint32_t* shr) {
  StackBlurTablesPublic(r, mul, shr);
}

uint32_t orc_gradient_color(const hipr_gradient* g, int32_t x, int32_t y) {
  return GradientColorAt(*g, x, y);
}

int64_t orc_blur_layer(const hipr_scene* scene, int32_t p, uint8_t* out,
                       int64_t cap, int32_t* bounds) {
  std::vector<uint8_t> img;
  int fb[4];
  BlurredLayer(scene, p, &img, fb);
  for (int i = 0; i < 4; i++) bounds[i] = fb[i];
  if ((int64_t)img.size() <= cap && out) std::memcpy(out, img.data(), img.size());
  return (int64_t)img.size();
}

}  // extern "C"
