// skity_hip.hip -- host runtime and C ABI of the MI355X DrawPath backend.
//
// hipr_surface_draw_scene turns an ordered list of DrawPath calls into
// raster jobs (SWCanvas::OnDrawPath, src/render/sw/sw_canvas.cc:357-411:
// fill and/or stroke in paint order; a mask filter draws into a layer that is
// blurred and composited back, sw_canvas.cc:797-826) and runs the device
// pipeline:
//   flatten (count) -> scan -> flatten (write) -> sort -> walk (count) ->
//   scan -> walk (write) -> composite layers -> blur -> composite canvas
// or, with pre-sized capacities (mode 1), flatten (write) -> sort ->
// walk (write) -> overflow check -> exact two-pass for the overflowed jobs
// only -> composite.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <memory>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <numeric>
#include <random>
#include <vector>

#include "../../include/skity_hip.h"
#include "hipr_blur.hpp"
#include "hipr_blur_mfma.hpp"
#include "hipr_composite.hpp"
#include "hipr_flatten.hpp"
#include "hipr_mlp.hpp"
#include "hipr_noalloc.hpp"
#include "hipr_sort.hpp"
#include "hipr_stroke.hpp"
#include "hipr_walk.hpp"

using namespace hipr;

namespace {

#define HIP_TRY(expr)                                   \
  do {                                                  \
    hipError_t err__ = (expr);                          \
    if (err__ != hipSuccess) {                          \
      std::fprintf(stderr, "skity_hip: %s failed: %s\n", #expr, \
                   hipGetErrorString(err__));           \
      return HIPR_ERROR_DEVICE;                         \
    }                                                   \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  hipError_t Ensure(size_t bytes) {
    if (bytes <= cap && p) return hipSuccess;
    if (p) hipFree(p);
    p = nullptr;
    cap = 0;
    size_t want = std::max<size_t>(bytes, 256);
    want = want + want / 4;
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) cap = want;
    return e;
  }
  // Grows like Ensure but keeps the first `keep` bytes of the old contents.
  hipError_t EnsureKeep(size_t bytes, size_t keep, hipStream_t s) {
    if (bytes <= cap && p) return hipSuccess;
    size_t want = std::max<size_t>(bytes, 256);
    want = want + want / 4;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, want);
    if (e != hipSuccess) return e;
    if (p) {
      keep = std::min(keep, cap);
      if (keep) e = hipMemcpyAsync(q, p, keep, hipMemcpyDeviceToDevice, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      hipFree(p);
    }
    p = q;
    cap = want;
    return e;
  }
  template <class T>
  T* as() const {
    return (T*)p;
  }
  void Free() {
    if (p) hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

// ---- host-side affine helpers (the same arithmetic as the reference's 4x4
// Matrix restricted to the plane; see GenerateBrush, sw_canvas.cc:727-795) --
struct Aff {
  float sx = 1, kx = 0, tx = 0, ky = 0, sy = 1, ty = 0;
  bool Identity() const {
    return sx == 1 && kx == 0 && tx == 0 && ky == 0 && sy == 1 && ty == 0;
  }
};
Aff Mul(const Aff& a, const Aff& b) {
  if (a.Identity()) return b;
  if (b.Identity()) return a;
  Aff r;
  r.sx = a.sx * b.sx + a.kx * b.ky;
  r.kx = a.sx * b.kx + a.kx * b.sy;
  r.tx = (a.sx * b.tx + a.kx * b.ty) + a.tx;
  r.ky = a.ky * b.sx + a.sy * b.ky;
  r.sy = a.ky * b.kx + a.sy * b.sy;
  r.ty = (a.ky * b.tx + a.sy * b.ty) + a.ty;
  return r;
}
Aff Translate(float x, float y) {
  Aff m;
  m.tx = x;
  m.ty = y;
  return m;
}
Aff Scale(float x, float y) {
  Aff m;
  m.sx = x;
  m.sy = y;
  return m;
}
// Matrix::Invert (src/geometry/matrix.cc:134-142,242-279); the general case
// uses a double-precision adjugate (glm::inverse in the reference).
bool Invert(const Aff& m, Aff* out) {
  if (m.Identity()) {
    *out = Aff();
    return true;
  }
  if (m.kx == 0 && m.ky == 0) {
    Aff r;
    if (m.sx != 1.f || m.sy != 1.f) {
      float ix = 1.f / m.sx, iy = 1.f / m.sy;
      if (std::isinf(ix) || std::isinf(iy)) return false;
      r.sx = ix;
      r.sy = iy;
      r.tx = -m.tx * ix;
      r.ty = -m.ty * iy;
    } else {
      r.tx = -m.tx;
      r.ty = -m.ty;
    }
    *out = r;
    return true;
  }
  double det = (double)m.sx * m.sy - (double)m.kx * m.ky;
  if (std::fabs((float)det) <= 1.f / 4096) return false;
  double id = 1.0 / det;
  Aff r;
  r.sx = (float)(m.sy * id);
  r.kx = (float)(-m.kx * id);
  r.ky = (float)(-m.ky * id);
  r.sy = (float)(m.sx * id);
  r.tx = (float)(((double)m.kx * m.ty - (double)m.sy * m.tx) * id);
  r.ty = (float)(((double)m.ky * m.tx - (double)m.sx * m.ty) * id);
  *out = r;
  return true;
}

uint8_t UnitToByte(float v) {
  float s = v * 255;
  s = std::min(std::max(s, 0.f), 255.f);
  return (uint8_t)s;
}
uint32_t MulDiv255(uint32_t a, uint32_t b) {
  uint32_t prod = a * b + 128;
  return (prod + (prod >> 8)) >> 8;
}
// Color4fToColor + ColorToPMColor, RGBA byte order.
uint32_t PremulRGBAHost(const float c[4]) {
  uint32_t a = UnitToByte(c[3]), r = UnitToByte(c[0]), g = UnitToByte(c[1]),
           b = UnitToByte(c[2]);
  if (a != 255) {
    r = MulDiv255(r, a);
    g = MulDiv255(g, a);
    b = MulDiv255(b, a);
  }
  return r | (g << 8) | (b << 16) | (a << 24);
}

// PointsToUnit (sw_span_brush.cc:164-198) * inverse(local) * inverse(CTM).
// PointsToUnit(p0, p1) (sw_span_brush.cc:200-219): rotation . scale .
// translation taking p0 to the origin and p1 to (1, 0).
Aff TwoPointsToUnit(float x0, float y0, float x1, float y1) {
  float ssx = x1 - x0, ssy = y1 - y0;
  float length = std::sqrt(ssx * ssx + ssy * ssy);
  float scale = length > 0 ? 1.0f / length : 0;
  float ux = ssx * scale, uy = ssy * scale;
  Aff rot;
  rot.sx = ux;
  rot.kx = uy;  // -sine with sine = -uy
  rot.ky = -uy;
  rot.sy = ux;
  return Mul(Mul(rot, Scale(scale, scale)), Translate(-x0, -y0));
}

// ConicalGradientColorBrush::OnPreBrush (sw_span_brush.cc:407-440), in the
// same float operations.  Points are (x, y, 0, 1), so the focal point's w is
// (1 - f) + f and enters |c1 - cf|.
ConicalDev ConicalSetup(const hipr_gradient& g) {
  ConicalDev c{};
  const float kNear = 1.f / 4096;
  float c0x = g.points[0], c0y = g.points[1], c1x = g.points[2], c1y = g.points[3];
  float r0 = g.radius[0], r1 = g.radius[1];
  float dcx = c1x - c0x, dcy = c1y - c0y;
  float delta_center = std::sqrt(dcx * dcx + dcy * dcy);
  float delta_radius = std::fabs(r1 - r0);
  c.kind = kConicalNone;
  c.sign = 1.f;
  if (r0 < 0 || r1 < 0) return c;
  bool radial = delta_center < kNear, strip = delta_radius < kNear;
  Aff u;
  if (radial) {
    if (strip) return c;
    c.kind = kConicalRadial;
    c.scale = (float)(1.0 / delta_radius);
    c.sign = delta_radius < 0 ? -1.f : 1.f;
    c.bias = r0 / delta_radius;
    c.c0x = c0x;
    c.c0y = c0y;
    return c;
  }
  if (strip) {
    c.kind = kConicalStrip;
    float r = r0 / delta_center;
    c.r2 = r * r;
    u = TwoPointsToUnit(c0x, c0y, c1x, c1y);
  } else {
    c.kind = kConicalFocal;
    c.swap = r1 < kNear;
    if (c.swap) {
      std::swap(c0x, c1x);
      std::swap(c0y, c1y);
      std::swap(r0, r1);
    }
    c.f = r0 / (r0 - r1);
    float cfx = c0x * (1.f - c.f) + c1x * c.f;
    float cfy = c0y * (1.f - c.f) + c1y * c.f;
    float cfw = 1.f * (1.f - c.f) + 1.f * c.f;
    float ex = c1x - cfx, ey = c1y - cfy, ew = 1.f - cfw;
    float len = std::sqrt(ex * ex + ey * ey + 0.f * 0.f + ew * ew);
    c.r1 = r1 / len;
    c.r1sq = c.r1 * c.r1;
    u = TwoPointsToUnit(cfx, cfy, c1x, c1y);
  }
  c.ux = u.sx;
  c.ukx = u.kx;
  c.utx = u.tx;
  c.uky = u.ky;
  c.uy = u.sy;
  c.uty = u.ty;
  return c;
}

GradDev MakeGradDev(const hipr_gradient& g, const Aff& ctm) {
  GradDev d{};
  d.type = g.type;
  d.tile = g.tile_mode;
  d.count = g.color_count;
  d.has_offsets = g.has_offsets;
  Aff ptu;
  if (g.type == HIPR_GRADIENT_LINEAR) {
    ptu = TwoPointsToUnit(g.points[0], g.points[1], g.points[2], g.points[3]);
  } else if (g.type == HIPR_GRADIENT_CONICAL) {
    d.con = ConicalSetup(g);  // the brush maps device -> local itself
  } else if (g.type == HIPR_GRADIENT_RADIAL) {
    float r = g.radius[0];
    float scale = r > 0 ? 1.0f / r : 0;
    ptu = Mul(Scale(scale, scale), Translate(-g.points[0], -g.points[1]));
  } else {
    ptu = Translate(-g.points[0], -g.points[1]);
    d.bias = g.radius[0];
    d.scale = g.radius[1];
  }
  Aff local;
  local.sx = g.local_matrix[0];
  local.kx = g.local_matrix[1];
  local.tx = g.local_matrix[2];
  local.ky = g.local_matrix[3];
  local.sy = g.local_matrix[4];
  local.ty = g.local_matrix[5];
  Aff d2l, l2l;
  if (!Invert(local, &d2l)) d2l = Aff();
  if (!Invert(ctm, &l2l)) l2l = Aff();
  Aff m = Mul(ptu, Mul(d2l, l2l));
  d.sx = m.sx;
  d.kx = m.kx;
  d.tx = m.tx;
  d.ky = m.ky;
  d.sy = m.sy;
  d.ty = m.ty;
  for (int i = 0; i < g.color_count && i < 16; i++) {
    d.offsets[i] = g.offsets[i];
    for (int c = 0; c < 4; c++) d.colors[i][c] = g.colors[i][c];
  }
  if (g.has_offsets) {  // EnsureColorOffsetsNonDescending (gradient_shader.cc:14-21)
    for (int i = 0; i < g.color_count - 1; i++)
      if (d.offsets[i] > d.offsets[i + 1]) d.offsets[i + 1] = d.offsets[i];
  }
  return d;
}

}  // namespace

struct hipr_context_s {
  int refs = 1;  // the handle itself + live surfaces and prepared scenes
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;  // walk tiers run here and on side2, beside the main stream
  hipStream_t side2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_join2 = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // per-draw device buffers (grow-only)
  DevBuf canvases, layer_canvases, geo, cov, edge_base, edge_cap, edges, wedges, gkeys;
  DevBuf job_tiles;  // per-job tile ranges for the composite
  DevBuf plane_base, plane_cap, planes, map_base, map0, map1, sub_list, big_list, bound,
      gather_list;
  DevBuf wcov, qstate, q_base, snaps, snap_base, band_count;  // write-pass coverage, row bands
  DevBuf queues;                          // persistent walk: item counters
  DevBuf srgb;                            // sRGB colour-filter tables (512 bytes)
  DevBuf flatten_redo;                    // jobs whose edges outgrew their estimate
  DevBuf stroke_redo, stroke_redo_out;    // stroke jobs written again (estimate forms)
  DevBuf pool_ctr;                        // overflow plane pool counter
  int n_cus = 0;                          // compute units of the device
  int max_lds = 0;                        // LDS bytes one workgroup may take
  std::vector<hipGraph_t> graphs;         // graphs of the walk passes in flight
  std::vector<hipGraphExec_t> graph_execs;
  DevBuf stroke_out, o_verbs, o_pts, o_w, s_verbs, s_pts, s_w;
  DevBuf layers, blur_rows, blur_cols, blur_rb, blur_cb, blur_tmp, blur_jobs, blur_lines;
  DevBuf blur_raw, blur_styles;  // unblurred layers and jobs of the style pass
  // debug/utility uploads
  DevBuf d_verbs, d_pts, d_w, d_jobs, run_jobs, counters;
  // learned pre-sizer
  bool has_presizer = false;
  bool mlp_dev_current = false;  // mlp_prm/mlp_ops hold presizer_host
  std::vector<float> presizer_host;
  float margin_edges = 0.f, margin_planes = 0.f;
  DevBuf mlp_prm, mlp_ops, mlp_feats, mlp_preds, mlp_targets, mlp_grads, mlp_loss, mlp_m,
      mlp_v, mlp_gsum, mlp_curve;
  // stage timing: events recorded on `stream` around each kernel launch
  struct StageMark {
    int32_t stage;
    hipEvent_t begin, end;
  };
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<StageMark> marks;
  // pinned staging for read-backs
  JobGeo* h_geo = nullptr;
  JobCov* h_cov = nullptr;
  JobCov* h_wcov = nullptr;
  StrokeOut* h_so = nullptr;
  size_t h_cap = 0;
  uint64_t draw_serial = 0;  // pipeline runs so far (see hipr_surface_s)
};

struct hipr_surface_s {
  hipr_context ctx = nullptr;
  int32_t width = 0, height = 0;
  int32_t format = 0;  // kFormat* bits
  uint8_t* pixels = nullptr;
  hipr_stats stats{};
  std::vector<int32_t> job_edges, job_planes;  // exact counts of the last draw
  std::vector<int32_t> job_steps, job_time;    // walk profile of the last draw
  std::vector<int32_t> job_phase;              // 5 per job
  std::vector<int32_t> job_active;             // largest active edge list
  // Where the last draw left each job's edges (emission order, context
  // buffer `edges`) and tile maps (context buffers `map0`/`map1`); valid while
  // draw_serial equals the context's.
  uint64_t draw_serial = 0;
  std::vector<int32_t> job_path, job_kind, job_canvas;
  std::vector<int64_t> job_ebase, job_mbase;
  std::vector<int32_t> job_tiles;  // per job: tc0, tr0, ntc, ntr, canvas tiles_x
  double stage_ms[HIPR_STAGE_COUNT] = {};
  int32_t stage_launches[HIPR_STAGE_COUNT] = {};
};

// A scene validated, planned and resident in device memory.
struct hipr_prepared_s {
  hipr_context ctx = nullptr;
  int32_t n_paths = 0;
  std::vector<Job> jobs;
  std::vector<Canvas> canvases;  // [0] is the target, then mask-filter layers
  std::vector<PaintDev> paints;
  std::vector<GradDev> grads;
  std::vector<BlurJob> blurs;
  std::vector<int32_t> layer_of_blur;
  std::vector<int32_t> blur_style;   // per blur: HIPR_BLUR_*
  std::vector<uint8_t> x_verbs;      // geometry appended after the scene's:
  std::vector<float> x_pts;          // the rectangles of walked layer composites
  std::vector<int32_t> walk_jobs;    // job indices that are rasterised
  std::vector<int32_t> stroke_jobs;  // job indices that need an outline
  std::vector<int32_t> stroke_quads; // per stroke job: segments of its quad path
  int64_t layer_bytes = 0;
  std::vector<ImageDev> images;      // pixmap shader images (offsets into image_pixels)
  int64_t image_bytes = 0;
  int32_t n_targets = 1;             // canvases [0, n_targets) are draw targets
  int64_t n_verbs = 0, n_pts = 0, n_weights = 0;
  DevBuf verbs, pts, weights, scene_paints, d_jobs, d_paints, d_grads, d_walk, d_stroke;
  DevBuf d_images, image_pixels;
  // Per job, the cost that bounds a sub-plan: source points (edges grow
  // with them) and the tiles of its bounding box (coverage planes).
  std::vector<int64_t> job_points, job_tiles;
  // A plan too large for one pipeline run is drawn as these sub-plans of
  // consecutive jobs, which borrow its geometry and paints (`borrowed`).
  std::vector<std::unique_ptr<hipr_prepared_s>> chunks;
  bool chunks_built = false;
  int64_t chunk_budget = 0;
  bool borrowed = false;
  int32_t job_offset = 0;  // a sub-plan: index of its first job in the plan
  ~hipr_prepared_s() {
    DevBuf* own[] = {&d_jobs, &d_walk, &d_stroke};
    for (DevBuf* x : own) x->Free();
    if (!borrowed) {
      DevBuf* b[] = {&verbs, &pts, &weights, &scene_paints, &d_paints, &d_grads,
                     &d_images, &image_pixels};
      for (DevBuf* x : b) x->Free();
    }
  }
};

// Many (surface, prepared scene) pairs drawn by one pipeline run: the plans
// are merged into one (`merged`), whose canvases [0, pairs) are the pairs'
// surfaces and whose jobs are the pairs' jobs in pair order.
struct hipr_batch_s {
  hipr_context ctx = nullptr;
  std::vector<std::pair<hipr_surface, hipr_prepared>> pairs;
  hipr_prepared_s merged;
  bool dirty = true;        // pairs changed since the last merge
  hipr_surface_s sink;      // statistics of the last draw (no pixels)
  int32_t tuned_band_rows = 0, tuned_band_min_edges = 0;  // 0: library defaults
  int32_t tuned_flags = 0;
};

namespace {

template <class T>
hipr_result Upload(DevBuf* b, const T* src, size_t n, hipStream_t s) {
  if (b->Ensure(std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess)
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  if (n) HIP_TRY(hipMemcpyAsync(b->p, src, n * sizeof(T), hipMemcpyHostToDevice, s));
  return HIPR_SUCCESS;
}

hipEvent_t NextEvent(hipr_context ctx) {
  if (ctx->ev_used == ctx->ev_pool.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    ctx->ev_pool.push_back(e);
  }
  return ctx->ev_pool[ctx->ev_used++];
}
hipEvent_t StageBegin(hipr_context ctx, hipStream_t st = nullptr) {
  hipEvent_t e = NextEvent(ctx);
  if (e) (void)hipEventRecord(e, st ? st : ctx->stream);
  return e;
}
void StageEnd(hipr_context ctx, int stage, hipEvent_t begin, hipStream_t st = nullptr) {
  hipEvent_t e = NextEvent(ctx);
  if (!e || !begin) return;
  (void)hipEventRecord(e, st ? st : ctx->stream);
  ctx->marks.push_back({stage, begin, e});
}
// Runs a kernel launch statement between two stage events on `st`.
#define STAGE_LAUNCH_ON(ctx, st, stage, ...)        \
  do {                                              \
    hipEvent_t stage_begin__ = StageBegin(ctx, st); \
    __VA_ARGS__;                                    \
    StageEnd(ctx, stage, stage_begin__, st);        \
  } while (0)
#define STAGE_LAUNCH(ctx, stage, ...) STAGE_LAUNCH_ON(ctx, ctx->stream, stage, __VA_ARGS__)

// Both StackBlur passes of every job with the MFMA banded-product kernels:
// rows src -> tmp, then columns tmp -> dst (jobs hold src/dst of the row
// pass; `tmp` per job is in `tmps`).
hipr_result LaunchBlurMfma(hipr_context ctx, const std::vector<BlurJob>& rows,
                           const std::vector<BlurJob>& cols, DevBuf* d_rows, DevBuf* d_cols,
                           hipStream_t st, bool timed);

hipr_result HostStage(hipr_context ctx, size_t n) {
  if (n > ctx->h_cap) {
    if (ctx->h_geo) (void)hipHostFree(ctx->h_geo);
    if (ctx->h_cov) (void)hipHostFree(ctx->h_cov);
    if (ctx->h_wcov) (void)hipHostFree(ctx->h_wcov);
    if (ctx->h_so) (void)hipHostFree(ctx->h_so);
    ctx->h_geo = nullptr;
    ctx->h_cov = nullptr;
    ctx->h_wcov = nullptr;
    ctx->h_so = nullptr;
    size_t want = n + n / 2 + 64;
    HIP_TRY(hipHostMalloc((void**)&ctx->h_geo, want * sizeof(JobGeo)));
    HIP_TRY(hipHostMalloc((void**)&ctx->h_cov, want * sizeof(JobCov)));
    HIP_TRY(hipHostMalloc((void**)&ctx->h_wcov, want * sizeof(JobCov)));
    HIP_TRY(hipHostMalloc((void**)&ctx->h_so, want * sizeof(StrokeOut)));
    ctx->h_cap = want;
  }
  return HIPR_SUCCESS;
}

int VerbPointCount(int verb) {
  return verb == HIPR_VERB_MOVE || verb == HIPR_VERB_LINE ? 1
         : verb == HIPR_VERB_QUAD || verb == HIPR_VERB_CONIC ? 2
         : verb == HIPR_VERB_CUBIC ? 3 : 0;
}

hipr_result ValidateScene(const hipr_scene* s) {
  if (!s) return HIPR_ERROR_INVALID_ARGUMENT;
  if (s->n_paths < 0) return HIPR_ERROR_INVALID_ARGUMENT;
  if (s->n_paths == 0) return HIPR_SUCCESS;
  if (!s->verb_offsets || !s->point_offsets || !s->weight_offsets || !s->paints)
    return HIPR_ERROR_INVALID_ARGUMENT;
  if (s->verb_offsets[0] != 0 || s->point_offsets[0] != 0 || s->weight_offsets[0] != 0)
    return HIPR_ERROR_INVALID_ARGUMENT;
  // the arrays may be NULL only when empty (a scene of DrawPaint calls)
  if ((!s->verbs && s->verb_offsets[s->n_paths] != 0) ||
      (!s->points && s->point_offsets[s->n_paths] != 0))
    return HIPR_ERROR_INVALID_ARGUMENT;
  if (s->n_images < 0 || (s->n_images > 0 && !s->images)) return HIPR_ERROR_INVALID_ARGUMENT;
  for (int i = 0; s->images && i < s->n_images; i++) {
    const hipr_image& im = s->images[i];
    if (im.width <= 0 || im.height <= 0 || !im.pixels ||
        (int64_t)im.width * im.height > (int64_t(1) << 28) ||
        im.alpha_type < HIPR_ALPHA_OPAQUE || im.alpha_type > HIPR_ALPHA_UNPREMUL)
      return HIPR_ERROR_INVALID_ARGUMENT;
  }
  for (int p = 0; p < s->n_paths; p++) {
    int v0 = s->verb_offsets[p], v1 = s->verb_offsets[p + 1];
    if (v1 < v0) return HIPR_ERROR_INVALID_ARGUMENT;
    int need_pts = 0, need_w = 0;
    bool open = false;
    for (int v = v0; v < v1; v++) {
      int verb = s->verbs[v];
      if (verb < 0 || verb > HIPR_VERB_CLOSE) return HIPR_ERROR_INVALID_ARGUMENT;
      if (verb == HIPR_VERB_MOVE) {
        if (v > v0 && s->verbs[v - 1] == HIPR_VERB_MOVE)
          return HIPR_ERROR_INVALID_ARGUMENT;  // the Path builder merges moves
        open = true;
      } else if (verb == HIPR_VERB_CLOSE) {
        open = false;
      } else if (!open) {
        return HIPR_ERROR_INVALID_ARGUMENT;  // a segment without a move
      }
      need_pts += VerbPointCount(verb);
      need_w += verb == HIPR_VERB_CONIC ? 1 : 0;
    }
    if (s->point_offsets[p + 1] - s->point_offsets[p] != need_pts)
      return HIPR_ERROR_INVALID_ARGUMENT;
    if (s->weight_offsets[p + 1] - s->weight_offsets[p] != need_w)
      return HIPR_ERROR_INVALID_ARGUMENT;
    if (need_w && !s->weights) return HIPR_ERROR_INVALID_ARGUMENT;
    for (int i = s->point_offsets[p]; i < s->point_offsets[p + 1]; i++) {
      // non-finite points are rejected (Rect::SetBoundsCheck, src/geometry/
      // rect.cc:12-49, only tests isinf of 0 * x products, which a NaN passes)
      if (!std::isfinite(s->points[2 * i]) || !std::isfinite(s->points[2 * i + 1]))
        return HIPR_ERROR_INVALID_ARGUMENT;
      // far beyond any device coordinate the walk can take (checked again in
      // device space after the flatten)
      if (std::fabs(s->points[2 * i]) > kSourceCoordLimit ||
          std::fabs(s->points[2 * i + 1]) > kSourceCoordLimit)
        return HIPR_ERROR_NOT_SUPPORTED;
    }
    const int kind = s->draw_kinds ? s->draw_kinds[p] : HIPR_DRAW_PATH;
    if (kind != HIPR_DRAW_PATH && kind != HIPR_DRAW_PAINT) return HIPR_ERROR_INVALID_ARGUMENT;
    if (kind == HIPR_DRAW_PAINT && v1 != v0) return HIPR_ERROR_INVALID_ARGUMENT;
    if (s->matrices) {
      for (int k = 0; k < 6; k++)
        if (!std::isfinite(s->matrices[6 * (size_t)p + k])) return HIPR_ERROR_INVALID_ARGUMENT;
    }
    for (int i = s->weight_offsets[p]; i < s->weight_offsets[p + 1]; i++) {
      float w = s->weights[i];
      if (!std::isfinite(w) || !(w > 0)) return HIPR_ERROR_INVALID_ARGUMENT;
    }
    const hipr_paint& pt = s->paints[p];
    if (pt.blend < HIPR_BLEND_CLEAR || pt.blend > HIPR_BLEND_LAST)
      return HIPR_ERROR_INVALID_ARGUMENT;
    if (pt.style < 0 || pt.style > 3) return HIPR_ERROR_INVALID_ARGUMENT;
    // Width 0 is stroked like any width: SWCanvas::OnDrawPath runs the
    // PathStroker with radius 0 (sw_canvas.cc:388-405, stroke.cc:865-866),
    // which leaves near-empty outlines.  A negative width is refused.
    if (kind == HIPR_DRAW_PATH && pt.style != HIPR_STYLE_FILL &&
        (!std::isfinite(pt.stroke_width) || pt.stroke_width < 0.f))
      return HIPR_ERROR_NOT_SUPPORTED;
    if (pt.cap < 0 || pt.cap > 2 || pt.join < 0 || pt.join > 2)
      return HIPR_ERROR_INVALID_ARGUMENT;
    if (pt.gradient >= 0) {
      if (!s->gradients || pt.gradient >= s->n_gradients) return HIPR_ERROR_INVALID_ARGUMENT;
      const hipr_gradient& g = s->gradients[pt.gradient];
      if (g.color_count < 2 || g.color_count > HIPR_MAX_GRADIENT_STOPS)
        return HIPR_ERROR_INVALID_ARGUMENT;
      if (g.type < HIPR_GRADIENT_LINEAR || g.type > HIPR_GRADIENT_CONICAL)
        return HIPR_ERROR_NOT_SUPPORTED;
    }
    if (kind == HIPR_DRAW_PATH && pt.blur_radius > 0 &&
        (pt.blur_style < HIPR_BLUR_NORMAL || pt.blur_style > HIPR_BLUR_INNER))
      return HIPR_ERROR_INVALID_ARGUMENT;
    if (s->image_shaders && s->image_shaders[p].image >= 0) {
      const hipr_image_shader& sh = s->image_shaders[p];
      if (!s->images || sh.image >= s->n_images || pt.gradient >= 0)
        return HIPR_ERROR_INVALID_ARGUMENT;
      if (sh.filter < HIPR_FILTER_NEAREST || sh.filter > HIPR_FILTER_LINEAR ||
          sh.tile_x < HIPR_TILE_CLAMP || sh.tile_x > HIPR_TILE_DECAL ||
          sh.tile_y < HIPR_TILE_CLAMP || sh.tile_y > HIPR_TILE_DECAL)
        return HIPR_ERROR_INVALID_ARGUMENT;
      for (int k = 0; k < 6; k++)
        if (!std::isfinite(sh.local_matrix[k])) return HIPR_ERROR_INVALID_ARGUMENT;
    }
    if (s->clip_rects) {
      for (int k = 0; k < 4; k++)
        if (std::isnan(s->clip_rects[4 * (size_t)p + k])) return HIPR_ERROR_INVALID_ARGUMENT;
    }
    if (s->color_filters) {
      const hipr_color_filter& f = s->color_filters[p];
      if (f.type < HIPR_COLOR_FILTER_NONE || f.type > HIPR_COLOR_FILTER_COMPOSE)
        return HIPR_ERROR_INVALID_ARGUMENT;
      if (f.type == HIPR_COLOR_FILTER_BLEND &&
          (f.blend < HIPR_BLEND_CLEAR || f.blend > HIPR_BLEND_LAST))
        return HIPR_ERROR_INVALID_ARGUMENT;
      if (f.type == HIPR_COLOR_FILTER_MATRIX)
        for (int k = 0; k < 20; k++)  // the reference keeps (int16_t)(m * 255)
          if (!std::isfinite(f.matrix[k]) || !(std::fabs(f.matrix[k] * 255) < 32767.f))
            return HIPR_ERROR_NOT_SUPPORTED;
    }
  }
  return HIPR_SUCCESS;
}

// Path::GetBounds of the source path (Path::ComputePtBounds): all points but
// a trailing move.
void SourceBounds(const hipr_scene* s, int p, float b[4]) {
  int p0 = s->point_offsets[p], p1 = s->point_offsets[p + 1];
  int v0 = s->verb_offsets[p], v1 = s->verb_offsets[p + 1];
  if (v1 - v0 > 1 && s->verbs[v1 - 1] == HIPR_VERB_MOVE) p1--;
  if (p1 <= p0) {
    b[0] = b[1] = b[2] = b[3] = 0;
    return;
  }
  b[0] = b[2] = s->points[2 * p0];
  b[1] = b[3] = s->points[2 * p0 + 1];
  for (int i = p0 + 1; i < p1; i++) {
    b[0] = std::min(b[0], s->points[2 * i]);
    b[2] = std::max(b[2], s->points[2 * i]);
    b[1] = std::min(b[1], s->points[2 * i + 1]);
    b[3] = std::max(b[3], s->points[2 * i + 1]);
  }
}

PaintDev SolidOrGradient(const hipr_scene* s, const hipr_paint& pt, const Aff& ctm,
                         bool stroke, std::vector<GradDev>* grads) {
  PaintDev pd{};
  pd.ga = 255;
  pd.blend = pt.blend;
  pd.layer = -1;
  if (pt.gradient >= 0) {
    pd.kind = kPaintGradient;
    pd.grad = (int32_t)grads->size();
    grads->push_back(MakeGradDev(s->gradients[pt.gradient], ctm));
  } else {
    pd.kind = kPaintSolid;
    pd.color = PremulRGBAHost(stroke ? pt.stroke_color : pt.color);
  }
  return pd;
}

// The paint's colour filter on a device paint record (every brush of the
// draw carries it: fills, strokes, DrawPaint, the blurred layer's draw and
// its composite).
void ApplyColorFilter(const hipr_scene* s, int p, PaintDev* pd) {
  pd->cf_type = HIPR_COLOR_FILTER_NONE;
  if (!s->color_filters) return;
  const hipr_color_filter& f = s->color_filters[p];
  pd->cf_type = f.type;
  pd->cf_mode = f.blend;
  // ColorToPMColor of the filter colour (0xAARRGGBB), RGBA byte order
  uint32_t a = f.color >> 24, r = (f.color >> 16) & 255, g = (f.color >> 8) & 255,
           b = f.color & 255;
  if (a != 255) {
    r = MulDiv255(r, a);
    g = MulDiv255(g, a);
    b = MulDiv255(b, a);
  }
  pd->cf_pm = r | (g << 8) | (b << 16) | (a << 24);
  for (int i = 0; i < 20; i++) pd->cf_m[i] = (int16_t)(f.matrix[i] * 255);
}

// The draw's pixmap shader on a device paint record (GenerateBrush,
// sw_canvas.cc:755-788): points to unit = Scale(1/w, 1/h) . inverse(local)
// . inverse(CTM); global alpha Paint::GetAlphaF (the stroke colour's alpha
// for kStroke_Style, else the fill colour's, paint.cc:96-98).
void ApplyImageShader(const hipr_scene* s, int p, const Aff& ctm, PaintDev* pd) {
  if (!s->image_shaders || s->image_shaders[p].image < 0) return;
  const hipr_image_shader& sh = s->image_shaders[p];
  const hipr_image& im = s->images[sh.image];
  const hipr_paint& pt = s->paints[p];
  pd->kind = kPaintPixmap;
  pd->pix_img = sh.image;
  pd->pix_filter = sh.filter;
  pd->pix_tx = sh.tile_x;
  pd->pix_ty = sh.tile_y;
  pd->pix_unpremul = im.alpha_type == HIPR_ALPHA_UNPREMUL;
  Aff local;
  local.sx = sh.local_matrix[0];
  local.kx = sh.local_matrix[1];
  local.tx = sh.local_matrix[2];
  local.ky = sh.local_matrix[3];
  local.sy = sh.local_matrix[4];
  local.ty = sh.local_matrix[5];
  Aff inv, l2l;
  if (!Invert(local, &inv)) inv = Aff();
  if (!Invert(ctm, &l2l)) l2l = Aff();
  Aff m = Mul(Mul(Scale(1.f / (uint32_t)im.width, 1.f / (uint32_t)im.height), inv), l2l);
  pd->im[0] = m.sx;
  pd->im[1] = m.kx;
  pd->im[2] = m.tx;
  pd->im[3] = m.ky;
  pd->im[4] = m.sy;
  pd->im[5] = m.ty;
  const float alpha = pt.style == HIPR_STYLE_STROKE ? pt.stroke_color[3] : pt.color[3];
  pd->ga = (int32_t)(uint8_t)(255 * alpha);
}

// A filter that can give a transparent source a colour makes the draw's
// zero-cover spans visible (blend and matrix filters; the sRGB tables map 0
// to 0 and compose passes colours through).
bool FilterColoursTransparent(const PaintDev& pd) {
  return pd.cf_type == HIPR_COLOR_FILTER_BLEND || pd.cf_type == HIPR_COLOR_FILTER_MATRIX;
}

Affine ToAffine(const Aff& a) { return Affine{a.sx, a.kx, a.tx, a.ky, a.sy, a.ty}; }

// Whether SWRenderTarget::BlendPixel changes a pixel for a transparent
// source (sw_render_target.cc:12-141, blend_mode.cc:125-191): FastBlend
// returns early for kSrcOver, kDst and kDstOut; kDstOver, kSrcATop, kXor,
// kPlus, kScreen and the modes falling back to kSrcOver are identities on
// premultiplied targets but round-trip through premultiplication on the
// others; the rest clear or rescale the pixel.
bool TransparentSourceMatters(int mode, bool premul_target) {
  switch (mode) {
    case HIPR_BLEND_SRC_OVER:
    case HIPR_BLEND_DST:
    case HIPR_BLEND_DST_OUT:
      return false;
    case HIPR_BLEND_DST_OVER:
    case HIPR_BLEND_SRC_ATOP:
    case HIPR_BLEND_XOR:
    case HIPR_BLEND_PLUS:
    case HIPR_BLEND_SCREEN:
      return !premul_target;
    default:
      if (mode > HIPR_BLEND_SCREEN && mode != HIPR_BLEND_SOFT_LIGHT) return !premul_target;
      return true;
  }
}

// The canvas matrix of draw p (identity when the scene has none).
Aff DrawMatrix(const hipr_scene* s, int p) {
  Aff m;
  if (s->matrices) {
    const float* f = s->matrices + 6 * (size_t)p;
    m.sx = f[0];
    m.kx = f[1];
    m.tx = f[2];
    m.ky = f[3];
    m.sy = f[4];
    m.ty = f[5];
  }
  return m;
}

// Fill and/or stroke jobs in paint order (DrawFillStrokeInPaintOrder,
// src/render/paint_order.hpp:12-31), rasterised through `ctm`.
void AddShapeJobs(const hipr_scene* s, int p, int canvas, const Aff& ctm,
                  hipr_prepared_s* pl, std::vector<Job>* out, const float* clip = nullptr) {
  const hipr_paint& pt = s->paints[p];
  bool need_fill = pt.style != HIPR_STYLE_STROKE;
  bool need_stroke = pt.style != HIPR_STYLE_FILL;
  auto add = [&](bool stroke) {
    Job jb{};
    jb.path = p;
    jb.canvas = canvas;
    jb.kind = stroke ? kJobStroke : kJobFill;
    jb.fill = stroke ? HIPR_FILL_WINDING
                     : (s->fill_types ? s->fill_types[p] : HIPR_FILL_WINDING);
    jb.geom = stroke ? 1 : 0;
    jb.verb_begin = s->verb_offsets[p];
    jb.verb_end = s->verb_offsets[p + 1];
    jb.pt_begin = s->point_offsets[p];
    jb.w_begin = s->weight_offsets[p];
    jb.m = ToAffine(ctm);
    jb.image = -1;
    if (clip) std::copy(clip, clip + 4, jb.clip);
    jb.paint = (int32_t)pl->paints.size();
    pl->paints.push_back(SolidOrGradient(s, pt, ctm, stroke, &pl->grads));
    ApplyColorFilter(s, p, &pl->paints.back());
    ApplyImageShader(s, p, ctm, &pl->paints.back());
    out->push_back(jb);
  };
  if (pt.style == HIPR_STYLE_STROKE_THEN_FILL) {
    if (need_stroke) add(true);
    if (need_fill) add(false);
  } else {
    if (need_fill) add(false);
    if (need_stroke) add(true);
  }
}

// Segments of a path's quad path (Stroke::QuadPath, stroke.cc:914-962):
// lines and quads one each, conics two, cubics Cubic::ToQuads's count, plus
// a closing line per contour.  Sizes the single-pass stroke outlines.
int32_t QuadPathSegments(const hipr_scene* s, int p) {
  int64_t n = 0;
  int pi = s->point_offsets[p];
  for (int v = s->verb_offsets[p]; v < s->verb_offsets[p + 1]; v++) {
    int verb = s->verbs[v];
    if (verb == HIPR_VERB_LINE || verb == HIPR_VERB_QUAD) {
      n += 1;
    } else if (verb == HIPR_VERB_CONIC) {
      n += 2;
    } else if (verb == HIPR_VERB_CUBIC) {
      P2 c[4] = {{s->points[2 * (pi - 1)], s->points[2 * (pi - 1) + 1]},
                 {s->points[2 * pi], s->points[2 * pi + 1]},
                 {s->points[2 * (pi + 1)], s->points[2 * (pi + 1) + 1]},
                 {s->points[2 * (pi + 2)], s->points[2 * (pi + 2) + 1]}};
      CubicQ q;
      CubicSetup(c, &q);
      n += q.n;
    } else if (verb == HIPR_VERB_CLOSE) {
      n += 1;
    }
    pi += VerbPointCount(verb);
  }
  return (int32_t)std::min<int64_t>(n + 1, 1 << 24);
}

// Stroke jobs in decreasing order of quad-path length: the stroke kernel's
// workgroups start in launch order, so the longest outlines start first.
void LongestStrokesFirst(hipr_prepared_s* pl) {
  const size_t n = pl->stroke_jobs.size();
  if (pl->stroke_quads.size() != n) return;
  std::vector<size_t> order(n);
  for (size_t i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    return pl->stroke_quads[a] > pl->stroke_quads[b];
  });
  std::vector<int32_t> jobs(n), quads(n);
  for (size_t i = 0; i < n; i++) {
    jobs[i] = pl->stroke_jobs[order[i]];
    quads[i] = pl->stroke_quads[order[i]];
  }
  pl->stroke_jobs.swap(jobs);
  pl->stroke_quads.swap(quads);
}

// Sub-plan budget: a run of at most this many source points (the edges of
// a run grow with them: about 0.5 KB of edge records, sorted copies and quad
// state per point) and four times as many bounding-box tiles (8 B of tile
// map each, and at most three 256-byte planes per touched tile): about 16 GB
// of edges and, for bounding boxes that are mostly covered, tens of GB of
// planes, a fraction of the 288 GB of HBM.  A 2048^2 frame of 4,096 large
// paths (about 70M bounding-box tiles) is one run.
constexpr int64_t kChunkPointsDefault = int64_t(32) << 20;
constexpr int64_t kChunkTilesPerPoint = 4;

// Per job: source points (an outline has several per source point) and the
// tiles of its device-space bounding box (corners of the source bounds
// through the job's matrix, widened by the stroke width).
void JobCosts(const hipr_scene* s, hipr_prepared_s* pl) {
  const size_t nj = pl->jobs.size();
  pl->job_points.assign(nj, 0);
  pl->job_tiles.assign(nj, 0);
  for (size_t i = 0; i < nj; i++) {
    const Job& jb = pl->jobs[i];
    if (jb.kind == kJobRect) continue;
    const int p = jb.path;
    const bool src_path = jb.pt_begin == s->point_offsets[p];
    int64_t pts = src_path ? s->point_offsets[p + 1] - s->point_offsets[p] : 4;
    if (jb.kind == kJobStroke) pts *= 8;
    pl->job_points[i] = pts;
    float b[4] = {0, 0, 0, 0};
    if (src_path) SourceBounds(s, p, b);
    float w = jb.kind == kJobStroke ? s->paints[p].stroke_width : 0.f;
    float xs[4] = {b[0] - w, b[2] + w, b[0] - w, b[2] + w};
    float ys[4] = {b[1] - w, b[1] - w, b[3] + w, b[3] + w};
    float l = INFINITY, t = INFINITY, r = -INFINITY, bt = -INFINITY;
    for (int k = 0; k < 4; k++) {
      P2 q = jb.m.Map(P2{xs[k], ys[k]});
      l = std::min(l, q.x);
      r = std::max(r, q.x);
      t = std::min(t, q.y);
      bt = std::max(bt, q.y);
    }
    const double side = kMaxCanvasSide / kTile;
    double tw = std::min(std::max((r - l) / kTile + 2.0, 1.0), side);
    double th = std::min(std::max((bt - t) / kTile + 2.0, 1.0), side);
    pl->job_tiles[i] = (int64_t)(tw * th);
  }
}

// A job of full coverage over an integer rectangle (see kJobRect).
Job RectJob(int p, int canvas, int paint) {
  Job jb{};
  jb.path = p;
  jb.canvas = canvas;
  jb.kind = kJobRect;
  jb.fill = HIPR_FILL_WINDING;
  jb.m = ToAffine(Aff());
  jb.image = -1;
  jb.paint = paint;
  return jb;
}

void BuildPlan(const hipr_scene* s, hipr_prepared_s* pl) {
  std::vector<std::vector<Job>> canvas_jobs(1);
  pl->canvases.push_back(Canvas{});
  const int32_t scene_verbs = s->verb_offsets[s->n_paths];
  const int32_t scene_pts = s->point_offsets[s->n_paths];
  for (int p = 0; p < s->n_paths; p++) {
    const hipr_paint& pt = s->paints[p];
    const Aff ctm = DrawMatrix(s, p);
    const int kind = s->draw_kinds ? s->draw_kinds[p] : HIPR_DRAW_PATH;
    // the canvas's global clip bounds: the scan clip of the path, stroke and
    // blurred-layer rasters of this draw (DrawPaint rasters without it)
    const float* clip = s->clip_rects && !ClipIsCull(s->clip_rects + 4 * (size_t)p)
                            ? s->clip_rects + 4 * (size_t)p
                            : nullptr;
    if (kind == HIPR_DRAW_PAINT) {
      // OnDrawPaint (sw_canvas.cc:412-438): the target rectangle under the
      // identity, the paint's fill brush (its shader through the canvas
      // matrix); no mask filter, style ignored.
      PaintDev pd = SolidOrGradient(s, pt, ctm, false, &pl->grads);
      pd.L = 0;
      pd.T = 0;
      pd.R = INT32_MAX;
      pd.B = INT32_MAX;
      ApplyColorFilter(s, p, &pd);
      ApplyImageShader(s, p, ctm, &pd);
      canvas_jobs[0].push_back(RectJob(p, 0, (int32_t)pl->paints.size()));
      pl->paints.push_back(pd);
      continue;
    }
    if (pt.blur_radius > 0) {
      // HandleFilter (sw_canvas.cc:797-826): local bounds, stroke-expanded
      // (floor on all sides, sw_canvas.cc:135-144), +- radius, floored /
      // ceiled (image_filter.cc:33-41).
      float b[4];
      SourceBounds(s, p, b);
      if (pt.style != HIPR_STYLE_FILL) {
        float w = pt.stroke_width;
        b[0] = std::floor(b[0] - w);
        b[1] = std::floor(b[1] - w);
        b[2] = std::floor(b[2] + w);
        b[3] = std::floor(b[3] + w);
      }
      float r = pt.blur_radius;
      float L = std::floor(b[0] - r), T = std::floor(b[1] - r);
      float R = std::ceil(b[2] + r), B = std::ceil(b[3] + r);
      int W = (int)(R - L), H = (int)(B - T);
      if (W <= 0 || H <= 0) continue;
      int layer = (int)pl->canvases.size();
      Canvas lc{};
      lc.width = W;
      lc.height = H;
      lc.pixels = (uint8_t*)(intptr_t)pl->layer_bytes;  // offset into the layer pool
      pl->layer_bytes += ((int64_t)W * H * 4 + 255) & ~(int64_t)255;
      pl->canvases.push_back(lc);
      canvas_jobs.emplace_back();
      // the layer is a fresh canvas translated by -LT: the draw's canvas
      // matrix does not reach it
      AddShapeJobs(s, p, layer, Translate(-L, -T), pl, &canvas_jobs[layer]);
      BlurJob bj{};
      bj.width = W;
      bj.height = H;
      bj.radius = std::min((int)std::round(r), 254);
      pl->blurs.push_back(bj);
      pl->layer_of_blur.push_back(layer);
      pl->blur_style.push_back(pt.blur_style);
      // The composite draw: OnDrawImageRect of the layer over the filter
      // bounds onto this canvas (sw_canvas.cc:641-677): the rectangle through
      // the canvas matrix, a decal nearest pixmap shader mapping device
      // points to the layer's unit square (GenerateBrush, sw_canvas.cc:
      // 765-789: Scale(1/w, 1/h) * inverse(local) * inverse(CTM)), and the
      // paint with its style set to fill, so that its alpha (Paint::GetAlphaF,
      // paint.cc:96-98) is the fill colour's.
      PaintDev ip{};
      ApplyColorFilter(s, p, &ip);
      ip.kind = kPaintImage;
      ip.layer = layer;
      ip.blend = pt.blend;
      ip.ga = (int32_t)(uint8_t)(255 * pt.color[3]);
      Aff inv, l2l;
      Invert(Translate(L, T), &inv);
      if (!Invert(ctm, &l2l)) l2l = Aff();
      Aff m = Mul(Mul(Scale(1.f / W, 1.f / H), inv), l2l);
      ip.im[0] = m.sx;
      ip.im[1] = m.kx;
      ip.im[2] = m.tx;
      ip.im[3] = m.ky;
      ip.im[4] = m.sy;
      ip.im[5] = m.ty;
      const int32_t paint = (int32_t)pl->paints.size();
      if (ctm.Identity()) {
        // an integer rectangle under the identity: full coverage over its
        // scan bounds (the rectangle intersected with the clip bounds,
        // floored / ceiled), no walk
        float sl = L, st = T, sr = R, sb = B;
        if (clip) {
          sl = std::max(L, clip[0]);
          st = std::max(T, clip[1]);
          sr = std::min(R, clip[2]);
          sb = std::min(B, clip[3]);
          if (!(sl < sr && st < sb)) sl = st = sr = sb = 0.f;
          sl = std::floor(sl);
          st = std::floor(st);
          sr = std::ceil(sr);
          sb = std::ceil(sb);
        }
        ip.L = (int)sl;
        ip.T = (int)st;
        ip.R = (int)sr;
        ip.B = (int)sb;
        Job jb = RectJob(p, 0, paint);
        jb.image = layer;
        canvas_jobs[0].push_back(jb);
      } else {
        // the rectangle path (Path::AddRect, path.cc:815-835) walked through
        // the canvas matrix
        Job jb{};
        jb.path = p;
        jb.canvas = 0;
        jb.kind = kJobFill;
        jb.fill = HIPR_FILL_WINDING;
        jb.geom = 0;
        jb.verb_begin = scene_verbs + (int32_t)pl->x_verbs.size();
        jb.pt_begin = scene_pts + (int32_t)(pl->x_pts.size() / 2);
        jb.w_begin = s->weight_offsets[s->n_paths];
        const uint8_t rv[5] = {HIPR_VERB_MOVE, HIPR_VERB_LINE, HIPR_VERB_LINE, HIPR_VERB_LINE,
                               HIPR_VERB_CLOSE};
        const float rp[8] = {L, T, R, T, R, B, L, B};
        pl->x_verbs.insert(pl->x_verbs.end(), rv, rv + 5);
        pl->x_pts.insert(pl->x_pts.end(), rp, rp + 8);
        jb.verb_end = jb.verb_begin + 5;
        jb.m = ToAffine(ctm);
        jb.image = layer;
        jb.paint = paint;
        if (clip) std::copy(clip, clip + 4, jb.clip);
        canvas_jobs[0].push_back(jb);
      }
      pl->paints.push_back(ip);
    } else {
      AddShapeJobs(s, p, 0, ctm, pl, &canvas_jobs[0], clip);
    }
  }
  // layers first (composited and blurred before the canvas), canvas last
  for (size_t c = 1; c <= canvas_jobs.size(); c++) {
    size_t ci = c % canvas_jobs.size();
    Canvas& cv = pl->canvases[ci];
    cv.job_begin = (int)pl->jobs.size();
    for (auto& jb : canvas_jobs[ci]) pl->jobs.push_back(jb);
    cv.job_end = (int)pl->jobs.size();
  }
  for (size_t i = 0; i < pl->jobs.size(); i++) {
    if (pl->jobs[i].kind != kJobRect) pl->walk_jobs.push_back((int32_t)i);
    if (pl->jobs[i].kind == kJobStroke) pl->stroke_jobs.push_back((int32_t)i);
  }
  JobCosts(s, pl);
  // pixmap shader images, each at a 256-byte aligned offset
  pl->images.clear();
  pl->image_bytes = 0;
  for (int i = 0; s->images && i < s->n_images; i++) {
    ImageDev d{};
    d.width = s->images[i].width;
    d.height = s->images[i].height;
    d.offset = pl->image_bytes;
    pl->image_bytes += ((int64_t)d.width * d.height * 4 + 255) & ~(int64_t)255;
    pl->images.push_back(d);
  }
}

void SetTiles(std::vector<Canvas>* cvs) {
  int total = 0;
  for (auto& cv : *cvs) {
    cv.tiles_x = (cv.width + kTile - 1) / kTile;
    cv.tiles_y = (cv.height + kTile - 1) / kTile;
    cv.tile_begin = total;
    total += cv.tiles_x * cv.tiles_y;
  }
}

// Nominal rows per walk band (the write pass runs the bands of a path
// concurrently).  Exact form: every path is banded by its count pass;
// estimate forms: only the widest paths, after a schedule pass, whose
// banded write then ends the walk stage, so shorter bands pay there
// (curves + strokes at 2048^2: learned 155.0 -> 152.6 ms with 64 rows;
// exact 305 -> 308 ms).
constexpr int kBandRowsExact = 128;
constexpr int kBandRowsEstimates = 64;
// Estimate modes: paths whose active-edge bound reaches this are banded:
// the paths of the 8- and 16-slot walks, the longest sequential walks of a
// frame, whose schedule pass and banded write run beside the 4-slot writes.
// (Banding every path costs more schedule work than it saves; the
// autotuner tries other thresholds.)
constexpr int kBandMinEdges = 4 * kWave + 1;
// ... and so are paths whose scan bounds span this many rows: a large
// outline with a short active list (the tiger at 4096^2: 2,000-4,000 steps
// per path) is as long a sequential walk, and its schedule pass costs a
// fraction of the write it splits (tiger.skp at 4096^2, learned: walk
// 34.5 -> 10 ms when every path is banded at 64 rows).  In a frame of many
// paths they are not the critical path, and banding all of them costs more
// than it saves (curves + strokes at 2048^2, learned: 175 -> 231 ms), hence
// the frame-relative threshold below.
constexpr int kBandMinRows = 512;
constexpr int kBandSlots = 512;

// LDS of a walk wave: row deltas, three per-tile-column rows (two slots,
// written rows), five scratch arrays of K * 64 entries.
size_t WalkLdsBytes(int K, int lds_width) {
  return (size_t)((lds_width + 1) + 3 * (lds_width / kTile + 2) + 5 * K * kWave) *
         sizeof(int32_t);
}

hipr_result WalkLdsCheck(hipr_context ctx, int K, size_t lds) {
  if (ctx->max_lds > 0 && lds > (size_t)ctx->max_lds) {
    std::fprintf(stderr, "skity_hip: a %d-slot walk needs %zu bytes of LDS (device limit %d)\n",
                 K, lds, ctx->max_lds);
    return HIPR_ERROR_CAPACITY;
  }
  return HIPR_SUCCESS;
}

template <int K, int kMode>
hipr_result LaunchWalk(hipr_context ctx, WalkParams P, hipStream_t st, bool timed,
                       int32_t* queue, double share) {
  if (P.n_walk == 0) return HIPR_SUCCESS;
  size_t lds = WalkLdsBytes(K, P.lds_width);
  if (hipr_result r = WalkLdsCheck(ctx, K, lds)) return r;
  unsigned grid = (unsigned)P.n_walk;
  P.queue = queue;
  if (queue) {
    // persistent form: as many waves as fit at once (this tier's share)
    int per_cu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, WalkKernel<K, kMode>, kWave,
                                                         lds));
    int64_t cap = (int64_t)std::max(per_cu, 1) * std::max(ctx->n_cus, 1);
    cap = std::max<int64_t>(1, (int64_t)((double)cap * share));
    grid = (unsigned)std::min<int64_t>(P.n_walk, cap);
  }
  const int stage = kMode == kWalkCount   ? HIPR_STAGE_WALK_COUNT
                    : kMode == kWalkWrite ? HIPR_STAGE_WALK
                                          : HIPR_STAGE_WALK_SCHEDULE;
  hipEvent_t b = timed ? StageBegin(ctx, st) : nullptr;
  hipLaunchKernelGGL((WalkKernel<K, kMode>), dim3(grid), dim3(kWave), lds, st, P);
  HIP_TRY(hipGetLastError());
  if (timed) StageEnd(ctx, stage, b, st);
  return HIPR_SUCCESS;
}

// The 32- and 64-slot walks (hipr_walk_wide.hip): one wave per job, never
// persistent, count and write modes only.
hipr_result LaunchWalkWide(hipr_context ctx, int K, int kmode, WalkParams P, hipStream_t st,
                           bool timed) {
  if (P.n_walk == 0) return HIPR_SUCCESS;
  if (kmode == kWalkSchedule) return HIPR_ERROR_INVALID_ARGUMENT;  // wide jobs are unbanded
  size_t lds = WalkLdsBytes(K, P.lds_width);
  if (hipr_result r = WalkLdsCheck(ctx, K, lds)) return r;
  P.queue = nullptr;
  const int stage = kmode == kWalkCount ? HIPR_STAGE_WALK_COUNT : HIPR_STAGE_WALK;
  hipEvent_t b = timed ? StageBegin(ctx, st) : nullptr;
  HIP_TRY(LaunchWideWalk(K, kmode, (unsigned)P.n_walk, lds, st, P));
  if (timed) StageEnd(ctx, stage, b, st);
  return HIPR_SUCCESS;
}

// Graphs launched by the walk passes; every pass is followed by a stream
// synchronisation, so they are released at the start of the next pass.
void ReleaseGraphs(hipr_context ctx) {
  for (hipGraphExec_t e : ctx->graph_execs) (void)hipGraphExecDestroy(e);
  for (hipGraph_t g : ctx->graphs) (void)hipGraphDestroy(g);
  ctx->graph_execs.clear();
  ctx->graphs.clear();
}

// ---- the device pipeline -----------------------------------------------------
struct Frame {
  hipr_context ctx;
  hipr_prepared_s* pl;
  hipStream_t st;
  int nj;
  Job* d_jobs = nullptr;  // the jobs as rasterised (stroke jobs point at outlines)
  std::vector<Canvas> canvases;
  std::vector<Job> jobs;  // stroke jobs get their outline ranges here
  hipr_stats stats{};
};

// Stroke outlines of the frame's stroke jobs.  Exact form: a count pass,
// then a write pass into exact ranges.  Estimate forms (`single_pass`): one
// write pass into ranges sized from each job's quad-path segment count
// (an outline has at most ~12 verbs, ~19 points and ~0.7 conic weights per
// segment on the workloads measured); a job whose outline outgrew its range
// reports its exact sizes and is written again into an exact range.
hipr_result RunStrokes(Frame& f, bool single_pass) {
  hipr_context ctx = f.ctx;
  hipr_prepared_s* pl = f.pl;
  int ns = (int)pl->stroke_jobs.size();
  if (ns == 0) return HIPR_SUCCESS;
  hipStream_t st = f.st;
  if (ctx->stroke_out.Ensure(sizeof(StrokeOut) * ns) != hipSuccess)
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  GeomSrc gs{pl->verbs.as<uint8_t>(), pl->pts.as<float>(), pl->weights.as<float>()};
  // One stroke job per wavefront (its lanes stroke quad segments ahead of
  // the stroker): 64 different paths in one wave would serialise their
  // branches.
  int blocks = ns;
  int64_t tv = 0, tp = 0, tw = 0, ts = 0;
  auto place = [&](StrokeOut& so) {
    so.verb_begin = (int32_t)tv;
    so.pt_begin = (int32_t)tp;
    so.w_begin = (int32_t)tw;
    so.scratch_begin = (int32_t)ts;
    tv += so.nv;
    tp += so.np;
    tw += so.nw;
    ts += std::max(so.inner_nv, std::max(so.inner_np, so.inner_nw));
  };
  auto ensure_outlines = [&](bool keep) -> hipr_result {
    if (tv > INT32_MAX || tp > INT32_MAX || ts > INT32_MAX) return HIPR_ERROR_CAPACITY;
    DevBuf* bufs[6] = {&ctx->o_verbs, &ctx->o_pts, &ctx->o_w, &ctx->s_verbs, &ctx->s_pts,
                       &ctx->s_w};
    const int64_t need[6] = {tv, 8 * tp, 4 * tw, ts, 8 * ts, 4 * ts};
    for (int i = 0; i < 6; i++) {
      size_t bytes = (size_t)std::max<int64_t>(need[i], 1);
      hipError_t e = keep ? bufs[i]->EnsureKeep(bytes, bufs[i]->cap, st) : bufs[i]->Ensure(bytes);
      if (e != hipSuccess) return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
    }
    return HIPR_SUCCESS;
  };
  auto launch = [&](auto mode_tag, const int32_t* list, int n, StrokeOut* outs) -> hipr_result {
    constexpr int kM = decltype(mode_tag)::value;
    STAGE_LAUNCH(ctx, HIPR_STAGE_STROKE,
      hipLaunchKernelGGL((StrokeKernel<kM>), dim3(n), dim3(64), 0, st, pl->d_jobs.as<Job>(), list,
                         n, gs, pl->scene_paints.as<hipr_paint>(), outs,
                         kM == 0 ? nullptr : ctx->o_verbs.as<uint8_t>(),
                         kM == 0 ? nullptr : ctx->o_pts.as<float>(),
                         kM == 0 ? nullptr : ctx->o_w.as<float>(),
                         kM == 0 ? nullptr : ctx->s_verbs.as<uint8_t>(),
                         kM == 0 ? nullptr : ctx->s_pts.as<float>(),
                         kM == 0 ? nullptr : ctx->s_w.as<float>()));
    HIP_TRY(hipGetLastError());
    return HIPR_SUCCESS;
  };
  auto set_job = [&](int i, const StrokeOut& so) {
    Job& jb = f.jobs[pl->stroke_jobs[i]];
    jb.verb_begin = so.verb_begin;
    jb.verb_end = so.verb_begin + so.nv;
    jb.pt_begin = so.pt_begin;
    jb.w_begin = so.w_begin;
  };
  hipr_result r;
  if (!single_pass) {
    if ((r = launch(std::integral_constant<int, 0>{}, pl->d_stroke.as<int32_t>(), blocks,
                    ctx->stroke_out.as<StrokeOut>())) != HIPR_SUCCESS)
      return r;
    HIP_TRY(hipMemcpyAsync(ctx->h_so, ctx->stroke_out.p, sizeof(StrokeOut) * ns,
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (int i = 0; i < ns; i++) {
      place(ctx->h_so[i]);
      set_job(i, ctx->h_so[i]);
    }
    if ((r = ensure_outlines(false)) != HIPR_SUCCESS) return r;
    HIP_TRY(hipMemcpyAsync(ctx->stroke_out.p, ctx->h_so, sizeof(StrokeOut) * ns,
                           hipMemcpyHostToDevice, st));
    return launch(std::integral_constant<int, 1>{}, pl->d_stroke.as<int32_t>(), blocks,
                  ctx->stroke_out.as<StrokeOut>());
  }
  std::vector<StrokeOut> cap(ns);
  for (int i = 0; i < ns; i++) {
    const int64_t q = pl->stroke_quads.empty() ? 64 : pl->stroke_quads[i];
    StrokeOut so{};
    so.nv = (int32_t)std::min<int64_t>(12 * q + 64, INT32_MAX / 4);
    so.np = (int32_t)std::min<int64_t>(20 * q + 64, INT32_MAX / 4);
    so.nw = (int32_t)std::min<int64_t>(q + 16, INT32_MAX / 4);
    so.inner_nv = so.nv;
    so.inner_np = so.np;
    so.inner_nw = so.nw;
    place(so);
    cap[i] = so;
  }
  if ((r = ensure_outlines(false)) != HIPR_SUCCESS) return r;
  HIP_TRY(hipMemcpyAsync(ctx->stroke_out.p, cap.data(), sizeof(StrokeOut) * ns,
                         hipMemcpyHostToDevice, st));
  if ((r = launch(std::integral_constant<int, 2>{}, pl->d_stroke.as<int32_t>(), blocks,
                  ctx->stroke_out.as<StrokeOut>())) != HIPR_SUCCESS)
    return r;
  HIP_TRY(hipMemcpyAsync(ctx->h_so, ctx->stroke_out.p, sizeof(StrokeOut) * ns,
                         hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  std::vector<int32_t> redo_list;
  std::vector<StrokeOut> redo_out;
  for (int i = 0; i < ns; i++) {
    StrokeOut so = ctx->h_so[i];  // ranges as placed, sizes as found
    const StrokeOut& c = cap[i];
    bool fits = so.nv <= c.nv && so.np <= c.np && so.nw <= c.nw && so.inner_nv <= c.inner_nv &&
                so.inner_np <= c.inner_np && so.inner_nw <= c.inner_nw;
    if (fits) {
      set_job(i, so);
      continue;
    }
    place(so);  // an exact range after the others
    set_job(i, so);
    redo_list.push_back(pl->stroke_jobs[i]);
    redo_out.push_back(so);
  }
  if (redo_list.empty()) return HIPR_SUCCESS;
  f.stats.fallbacks += (int64_t)redo_list.size();
  if ((r = ensure_outlines(true)) != HIPR_SUCCESS) return r;
  if ((r = Upload(&ctx->stroke_redo, redo_list.data(), redo_list.size(), st)) != HIPR_SUCCESS ||
      (r = Upload(&ctx->stroke_redo_out, redo_out.data(), redo_out.size(), st)) != HIPR_SUCCESS)
    return r;
  return launch(std::integral_constant<int, 1>{}, ctx->stroke_redo.as<int32_t>(),
                (int)redo_list.size(), ctx->stroke_redo_out.as<StrokeOut>());
}

// Flatten + sort for the jobs in `list` (indices into f.jobs).
hipr_result Flatten(Frame& f, int mode, const int32_t* edge_est, std::vector<int64_t>* ebase,
                    std::vector<int32_t>* ecap) {
  hipr_context ctx = f.ctx;
  hipr_prepared_s* pl = f.pl;
  hipStream_t st = f.st;
  int nj = f.nj;
  GeomSrc gs{pl->verbs.as<uint8_t>(), pl->pts.as<float>(), pl->weights.as<float>()};
  GeomSrc go{ctx->o_verbs.as<uint8_t>(), ctx->o_pts.as<float>(), ctx->o_w.as<float>()};
  ebase->assign(nj + 1, 0);
  ecap->assign(nj, 0);
  if (mode == 0) {
    STAGE_LAUNCH(ctx, HIPR_STAGE_FLATTEN_COUNT,
    hipLaunchKernelGGL((FlattenKernel<0>), dim3(nj), dim3(64), 0, st, f.d_jobs, nj,
                       gs, go, ctx->edge_base.as<int64_t>(), ctx->edge_cap.as<int32_t>(),
                       (EdgeRec*)nullptr, ctx->geo.as<JobGeo>(), nullptr));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(ctx->h_geo, ctx->geo.p, sizeof(JobGeo) * nj, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (int j = 0; j < nj; j++) (*ecap)[j] = ctx->h_geo[j].edges;
  } else {
    for (int j = 0; j < nj; j++) (*ecap)[j] = f.jobs[j].kind == kJobRect ? 0 : edge_est[j];
  }
  for (int j = 0; j < nj; j++) (*ebase)[j + 1] = (*ebase)[j] + (*ecap)[j];
  int64_t total = (*ebase)[nj];
  if (ctx->edges.Ensure(sizeof(EdgeRec) * std::max<int64_t>(total, 1)) != hipSuccess)
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  HIP_TRY(hipMemcpyAsync(ctx->edge_base.p, ebase->data(), sizeof(int64_t) * nj,
                         hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(ctx->edge_cap.p, ecap->data(), sizeof(int32_t) * nj,
                         hipMemcpyHostToDevice, st));
  STAGE_LAUNCH(ctx, HIPR_STAGE_FLATTEN,
    hipLaunchKernelGGL((FlattenKernel<1>), dim3(nj), dim3(64), 0, st, f.d_jobs, nj, gs,
                     go, ctx->edge_base.as<int64_t>(), ctx->edge_cap.as<int32_t>(),
                     ctx->edges.as<EdgeRec>(), ctx->geo.as<JobGeo>(), nullptr));
  HIP_TRY(hipGetLastError());
  if (mode == 1) {
    // one read-back: exact counts, bounds and overflow flags
    HIP_TRY(hipMemcpyAsync(ctx->h_geo, ctx->geo.p, sizeof(JobGeo) * nj, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    std::vector<int32_t> redo;
    for (int j = 0; j < nj; j++)
      if (ctx->h_geo[j].overflow) redo.push_back(j);
    if (!redo.empty()) {
      // exact fallback: the write pass already counted; append exact space
      f.stats.fallbacks += (int64_t)redo.size();
      for (int j : redo) {
        (*ebase)[j] = total;
        (*ecap)[j] = ctx->h_geo[j].edges;
        total += (*ecap)[j];
      }
      // the other jobs' edges stay where they are (kept if the buffer grows)
      const size_t written = sizeof(EdgeRec) * (size_t)(*ebase)[nj];
      if (ctx->edges.EnsureKeep(sizeof(EdgeRec) * std::max<int64_t>(total, 1), written, st) !=
          hipSuccess)
        return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
      HIP_TRY(hipMemcpyAsync(ctx->edge_base.p, ebase->data(), sizeof(int64_t) * nj,
                             hipMemcpyHostToDevice, st));
      HIP_TRY(hipMemcpyAsync(ctx->edge_cap.p, ecap->data(), sizeof(int32_t) * nj,
                             hipMemcpyHostToDevice, st));
      hipr_result rr = Upload(&ctx->flatten_redo, redo.data(), redo.size(), st);
      if (rr != HIPR_SUCCESS) return rr;
      STAGE_LAUNCH(ctx, HIPR_STAGE_FLATTEN,
    hipLaunchKernelGGL((FlattenKernel<1>), dim3((unsigned)redo.size()), dim3(64), 0, st, f.d_jobs,
                         (int)redo.size(), gs, go, ctx->edge_base.as<int64_t>(),
                         ctx->edge_cap.as<int32_t>(), ctx->edges.as<EdgeRec>(),
                         ctx->geo.as<JobGeo>(), ctx->flatten_redo.as<int32_t>()));
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMemcpyAsync(ctx->h_geo, ctx->geo.p, sizeof(JobGeo) * nj, hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
    }
  }
  f.stats.edges = 0;
  for (int j = 0; j < nj; j++) f.stats.edges += ctx->h_geo[j].edges;
  if (ctx->wedges.Ensure(sizeof(EdgeRec) * std::max<int64_t>(total, 1)) != hipSuccess ||
      ctx->gkeys.Ensure(sizeof(uint64_t) * 2 * std::max<int64_t>(total, 1)) != hipSuccess)
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  return HIPR_SUCCESS;
}

hipr_result SortGather(Frame& f, const int32_t* d_list, int n, int32_t* d_bound = nullptr) {
  if (n == 0) return HIPR_SUCCESS;
  hipr_context ctx = f.ctx;
  STAGE_LAUNCH(ctx, HIPR_STAGE_SORT,
    hipLaunchKernelGGL(SortGatherKernel, dim3(n), dim3(kSortThreads), 0, f.st, d_list, n,
                     ctx->geo.as<JobGeo>(), ctx->edge_base.as<int64_t>(),
                     ctx->edges.as<EdgeRec>(), ctx->wedges.as<EdgeRec>(),
                     ctx->gkeys.as<uint64_t>(), d_bound));
  HIP_TRY(hipGetLastError());
  return HIPR_SUCCESS;
}

// Features of every job of a prepared scene into ctx->mlp_feats (device).
hipr_result LaunchFeatures(hipr_context ctx, hipr_prepared_s* pl, const Canvas* d_canvases) {
  int nj = (int)pl->jobs.size();
  if (ctx->mlp_feats.Ensure(sizeof(float) * kMlpIn * std::max(nj, 1)) != hipSuccess)
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  if (nj == 0) return HIPR_SUCCESS;
  FeatureParams fp{};
  fp.jobs = pl->d_jobs.as<Job>();
  fp.n_jobs = nj;
  fp.scene = GeomSrc{pl->verbs.as<uint8_t>(), pl->pts.as<float>(), pl->weights.as<float>()};
  fp.paints = pl->scene_paints.as<hipr_paint>();
  fp.canvases = d_canvases;
  fp.out = ctx->mlp_feats.as<float>();
  hipLaunchKernelGGL(FeaturesKernel, dim3(nj), dim3(64), 0, ctx->stream, fp);
  HIP_TRY(hipGetLastError());
  return HIPR_SUCCESS;
}

hipr_result UploadPresizer(hipr_context ctx, const float* params) {
  hipr_result r;
  if ((r = Upload(&ctx->mlp_prm, params, kMlpTotal, ctx->stream)) != HIPR_SUCCESS) return r;
  if (ctx->mlp_ops.Ensure(sizeof(_Float16) * kOpBlocks * 256) != hipSuccess)
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  hipLaunchKernelGGL(MlpPackKernel, dim3(kOpBlocks), dim3(256), 0, ctx->stream,
                     ctx->mlp_prm.as<float>(), ctx->mlp_ops.as<_Float16>());
  HIP_TRY(hipGetLastError());
  return HIPR_SUCCESS;
}

// Learned capacities for every job: features -> MLP -> ceil(expm1(p + m)).
hipr_result RunPresize(Frame& f, std::vector<int32_t>* eest, std::vector<int32_t>* test) {
  hipr_context ctx = f.ctx;
  int nj = f.nj;
  hipr_result r;
  if (ctx->mlp_preds.Ensure(sizeof(float) * kMlpOut * std::max(nj, 1)) != hipSuccess)
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  if (!ctx->mlp_dev_current) {
    if ((r = UploadPresizer(ctx, ctx->presizer_host.data())) != HIPR_SUCCESS) return r;
    ctx->mlp_dev_current = true;
  }
  hipEvent_t b = StageBegin(ctx);
  if ((r = LaunchFeatures(ctx, f.pl, ctx->canvases.as<Canvas>())) != HIPR_SUCCESS) return r;
  int blocks = std::min((nj + 15) / 16, 8192);
  hipLaunchKernelGGL(MlpInferKernel, dim3(blocks), dim3(64), 0, f.st, ctx->mlp_ops.as<_Float16>(),
                     ctx->mlp_prm.as<float>(), ctx->mlp_feats.as<float>(), (int64_t)nj,
                     ctx->mlp_preds.as<float>());
  HIP_TRY(hipGetLastError());
  StageEnd(ctx, HIPR_STAGE_PRESIZE, b);
  std::vector<float> pred((size_t)kMlpOut * nj);
  HIP_TRY(hipMemcpyAsync(pred.data(), ctx->mlp_preds.p, sizeof(float) * pred.size(),
                         hipMemcpyDeviceToHost, f.st));
  HIP_TRY(hipStreamSynchronize(f.st));
  eest->resize(nj);
  test->resize(nj);
  auto cap = [](float p, float m) {
    double v = std::ceil(std::expm1((double)p + m));
    return (int32_t)std::min(std::max(v, 1.0), 1.0e9);
  };
  for (int j = 0; j < nj; j++) {
    // A prediction far outside the training distribution (glyph-sized
    // outlines, say) can be huge; capacities beyond 64 edges per verb only
    // waste memory (a job that needs more falls back to its exact count).
    const int64_t verbs = std::max<int64_t>(f.jobs[j].verb_end - f.jobs[j].verb_begin, 0);
    (*eest)[j] = (int32_t)std::min<int64_t>(cap(pred[2 * j], ctx->margin_edges), 64 * verbs + 64);
    (*test)[j] = cap(pred[2 * j + 1], ctx->margin_planes);
  }
  return HIPR_SUCCESS;
}

// Draws a planned scene whose canvases [0, targets.size()) are the given
// surfaces (the others are mask-filter layers); statistics, stage times and
// per-job profiles go to `surf`.
hipr_result RunPipeline(const std::vector<hipr_surface>& targets, hipr_surface surf,
                        hipr_prepared_s* pl, const hipr_options* opt) {
  hipr_context ctx = surf->ctx;
  const int n_targets = (int)targets.size();
  if (n_targets != pl->n_targets) return HIPR_ERROR_INVALID_ARGUMENT;
  Frame f;
  f.ctx = ctx;
  f.pl = pl;
  f.st = ctx->stream;
  f.nj = (int)pl->jobs.size();
  f.jobs = pl->jobs;
  f.canvases = pl->canvases;
  f.stats.paths = pl->n_paths;
  f.stats.raster_paths = (int64_t)pl->walk_jobs.size();
  hipStream_t st = f.st;
  int nj = f.nj;
  int mode = opt ? opt->mode : HIPR_MODE_EXACT;
  const int blur_impl = opt ? opt->blur_impl : HIPR_BLUR_IMPL_MFMA;
  if (blur_impl != HIPR_BLUR_IMPL_MFMA && blur_impl != HIPR_BLUR_IMPL_LINES)
    return HIPR_ERROR_INVALID_ARGUMENT;
  if (mode < HIPR_MODE_EXACT || mode > HIPR_MODE_LEARNED) return HIPR_ERROR_INVALID_ARGUMENT;
  if (mode == HIPR_MODE_ESTIMATES && (!opt->edge_estimates || !opt->tile_estimates))
    return HIPR_ERROR_INVALID_ARGUMENT;
  if (mode == HIPR_MODE_LEARNED && !ctx->has_presizer) return HIPR_ERROR_INVALID_ARGUMENT;
  if (nj >= (1 << (31 - kBandBits))) return HIPR_ERROR_CAPACITY;  // walk items: job, band
  ctx->ev_used = 0;
  ctx->marks.clear();
  for (int k = 0; k < HIPR_STAGE_COUNT; k++) {
    surf->stage_ms[k] = 0;
    surf->stage_launches[k] = 0;
  }
  HIP_TRY(hipEventRecord(ctx->ev0, st));
  if (nj == 0) {
    HIP_TRY(hipEventRecord(ctx->ev1, st));
    surf->stats = f.stats;
    return HIPR_SUCCESS;
  }
  // canvases: the target and the mask-filter layers
  for (int t = 0; t < n_targets; t++) {
    f.canvases[t].pixels = targets[t]->pixels;
    f.canvases[t].width = targets[t]->width;
    f.canvases[t].height = targets[t]->height;
    f.canvases[t].format = targets[t]->format;
  }
  // Walked draws whose blend changes pixels under a transparent source keep
  // the reference's zero-cover spans on this target.
  bool any_keep_zero = false;
  for (Job& jb : f.jobs) {
    jb.flags = 0;
    if (jb.kind == kJobRect) continue;
    const bool premul = (f.canvases[jb.canvas].format & kFormatUnpremul) == 0;
    const PaintDev& jp = pl->paints[jb.paint];
    if (TransparentSourceMatters(jp.blend, premul) || FilterColoursTransparent(jp)) {
      jb.flags |= kJobKeepZero;
      any_keep_zero = true;
    }
  }
  if (pl->layer_bytes > 0) {
    if (ctx->layers.Ensure(pl->layer_bytes) != hipSuccess) return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
    for (size_t c = n_targets; c < f.canvases.size(); c++)
      f.canvases[c].pixels = ctx->layers.as<uint8_t>() + (intptr_t)pl->canvases[c].pixels;
    HIP_TRY(hipMemsetAsync(ctx->layers.p, 0, pl->layer_bytes, st));
  }
  SetTiles(&f.canvases);
  hipr_result r;
  if ((r = Upload(&ctx->canvases, f.canvases.data(), f.canvases.size(), st)) != HIPR_SUCCESS) return r;
  if ((r = HostStage(ctx, std::max<size_t>(nj, pl->stroke_jobs.size()))) != HIPR_SUCCESS) return r;
  if (ctx->geo.Ensure(sizeof(JobGeo) * nj) != hipSuccess ||
      ctx->cov.Ensure(sizeof(JobCov) * nj) != hipSuccess ||
      ctx->edge_base.Ensure(sizeof(int64_t) * nj) != hipSuccess ||
      ctx->edge_cap.Ensure(sizeof(int32_t) * nj) != hipSuccess ||
      ctx->plane_base.Ensure(sizeof(int64_t) * nj) != hipSuccess ||
      ctx->plane_cap.Ensure(sizeof(int32_t) * nj) != hipSuccess ||
      ctx->map_base.Ensure(sizeof(int64_t) * nj) != hipSuccess)
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  // ---- stage (a): stroke outlines, flatten, sort ----
  if ((r = RunStrokes(f, mode != HIPR_MODE_EXACT)) != HIPR_SUCCESS) return r;
  // the prepared job table keeps the source ranges; stroke jobs are redirected
  // to their outlines in a per-draw copy
  f.d_jobs = pl->d_jobs.as<Job>();
  if (!pl->stroke_jobs.empty() || any_keep_zero) {
    if ((r = Upload(&ctx->run_jobs, f.jobs.data(), f.jobs.size(), st)) != HIPR_SUCCESS) return r;
    f.d_jobs = ctx->run_jobs.as<Job>();
  }
  // per-job estimates are indexed by draw (path); spread them to jobs
  std::vector<int32_t> eest, test;
  if (mode == HIPR_MODE_LEARNED) {
    if ((r = RunPresize(f, &eest, &test)) != HIPR_SUCCESS) return r;
    mode = 1;  // from here on exactly the estimates path
  } else if (mode == 1) {
    eest.resize(nj);
    test.resize(nj);
    for (int j = 0; j < nj; j++) {
      eest[j] = std::max(opt->edge_estimates[f.jobs[j].path], 1);
      test[j] = std::max(opt->tile_estimates[f.jobs[j].path], 1);
    }
  }
  std::vector<int64_t> ebase;
  std::vector<int32_t> ecap;
  if ((r = Flatten(f, mode, eest.data(), &ebase, &ecap)) != HIPR_SUCCESS) return r;
  if (mode == 0) {
    // (the count pass read the bounds back already)
  }
  int lds_width = 256;
  std::vector<int64_t> mbase(nj + 1, 0);
  for (int j = 0; j < nj; j++) {
    int64_t msz = 0;
    if (f.jobs[j].kind != kJobRect) {
      const JobGeo& g = ctx->h_geo[j];
      if (g.l < g.r && g.t < g.b &&
          (std::fabs(g.l) > kCoordLimit || std::fabs(g.r) > kCoordLimit ||
           std::fabs(g.t) > kCoordLimit || std::fabs(g.b) > kCoordLimit)) {
        // beyond 2^23 px the reference's float -> 26.6 conversion
        // ((int)(4x * 64), sw_edge.cc:24-29) overflows int32
        std::fprintf(stderr, "skity_hip: draw %d has device coordinates beyond 2^23 px\n",
                     f.jobs[j].path);
        return HIPR_ERROR_NOT_SUPPORTED;
      }
      PixBounds pb = JobScanBounds(g);
      const Canvas& cv = f.canvases[f.jobs[j].canvas];
      TileRange tr = JobTiles(pb, cv.width, cv.height);
      // the walk's row holds the job's visible columns only
      if (tr.ntc > 0)
        lds_width = std::max(lds_width, std::min(pb.r, cv.width) - std::max(pb.l, 0));
      msz = (int64_t)tr.ntc * tr.ntr;
    }
    mbase[j + 1] = mbase[j] + msz;
  }
  lds_width = (lds_width + 255) & ~255;
  if (lds_width > kMaxCanvasSide) return HIPR_ERROR_CAPACITY;  // a layer wider than a canvas
  int64_t total_map = mbase[nj];
  if (ctx->map0.Ensure(sizeof(int32_t) * std::max<int64_t>(total_map, 1)) != hipSuccess ||
      ctx->map1.Ensure(sizeof(int32_t) * std::max<int64_t>(total_map, 1)) != hipSuccess)
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  HIP_TRY(hipMemcpyAsync(ctx->map_base.p, mbase.data(), sizeof(int64_t) * nj,
                         hipMemcpyHostToDevice, st));
  int nwalk = (int)pl->walk_jobs.size();
  if (ctx->bound.Ensure(sizeof(int32_t) * std::max(nj, 1)) != hipSuccess)
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  if ((r = SortGather(f, pl->d_walk.as<int32_t>(), nwalk, ctx->bound.as<int32_t>())) !=
      HIPR_SUCCESS)
    return r;
  WalkParams wp{};
  wp.jobs = f.d_jobs;
  wp.walk_list = pl->d_walk.as<int32_t>();
  wp.n_walk = nwalk;
  wp.geo = ctx->geo.as<JobGeo>();
  wp.edge_base = ctx->edge_base.as<int64_t>();
  wp.wedges = ctx->wedges.as<EdgeRec>();
  wp.canvases = ctx->canvases.as<Canvas>();
  wp.cov = ctx->cov.as<JobCov>();
  wp.lds_width = lds_width;
  wp.map_base = ctx->map_base.as<int64_t>();
  // ---- stage (b): walk ----
  // The walk keeps the active edge list in registers.  The sort bounded each
  // job's list by its largest row overlap: jobs within 256 edges take the
  // 4-slot walk on the main stream, the others the 8-slot (512 edges) or
  // 16-slot (1024 edges) walk on a side stream, at the same time, and the
  // rare longer lists the unbanded 32- or 64-slot walk (2048 and 4096
  // edges).  A longer list is refused (HIPR_ERROR_CAPACITY).
  std::vector<int32_t> bound(nj, 0);
  HIP_TRY(hipMemcpyAsync(bound.data(), ctx->bound.p, sizeof(int32_t) * nj, hipMemcpyDeviceToHost,
                         st));
  HIP_TRY(hipStreamSynchronize(st));
  const std::vector<int32_t>& all = pl->walk_jobs;
  std::vector<int32_t> small, wide;
  std::vector<int8_t> tier(nj, 4);  // slots of each job's walk
  for (int32_t j : all) {
    if (bound[j] > kMaxSlots * kWave) {
      std::fprintf(stderr, "skity_hip: draw %d keeps up to %d edges active (limit %d)\n",
                   f.jobs[j].path, bound[j], kMaxSlots * kWave);
      return HIPR_ERROR_CAPACITY;
    }
    int k = 4;
    while (bound[j] > k * kWave) k *= 2;
    tier[j] = (int8_t)k;
    (tier[j] == 4 ? small : wide).push_back(j);
  }
  // Row bands: each job's rows are split into bands walked concurrently by
  // the write pass; the count or schedule pass records where they start.
  const int band_rows = opt && opt->band_rows != 0 ? std::max(opt->band_rows, 0)
                        : mode == 0                    ? kBandRowsExact
                                                       : kBandRowsEstimates;
  // In the estimate modes banding costs a schedule pass, so only the paths
  // with the widest active lists (the longest walks) are banded.
  const int band_min =
      mode == 0 ? 0
      : (opt && opt->band_min_edges != 0) ? std::max(opt->band_min_edges, 0)
                                          : kBandMinEdges;
  // (estimate modes, default split only) paths this tall are banded too:
  // at least kBandMinRows, and taller than the frame's rows spread over
  // kBandSlots concurrent walks, so that only paths that would outlast the
  // rest of the walk are split (a frame of many paths keeps its unbanded
  // walks; one of a few large outlines bands them)
  int band_min_rows = 0;
  if (mode != 0 && !(opt && opt->band_min_edges != 0)) {
    int64_t total_rows = 0;
    for (int32_t j : all) {
      const PixBounds b = JobScanBounds(ctx->h_geo[j]);
      if (!b.empty) total_rows += b.b - b.t;
    }
    band_min_rows = (int)std::min<int64_t>(
        std::max<int64_t>(kBandMinRows, total_rows / kBandSlots), INT32_MAX);
  }
  std::vector<int32_t> nbands(nj, 1);
  std::vector<int64_t> qb(nj, 0), sb(nj, 0);
  int64_t q_total = 0, snap_total = 0;
  for (int32_t j : all) {
    // the 32- and 64-slot walks are unbanded
    const PixBounds sb_j = JobScanBounds(ctx->h_geo[j]);
    // a tall path is a long sequential walk too, whatever its list width
    const bool tall = band_min_rows > 0 && !sb_j.empty && sb_j.b - sb_j.t >= band_min_rows;
    nbands[j] = (bound[j] >= band_min || tall) && tier[j] <= 16 ? BandCount(sb_j, band_rows) : 1;
    qb[j] = q_total;
    q_total += (int64_t)nbands[j] * ctx->h_geo[j].edges;
  }
  auto assign_snaps = [&](const std::vector<int32_t>& js) {
    for (int32_t j : js) {
      sb[j] = snap_total;
      snap_total += (int64_t)(nbands[j] - 1) * SnapStride(tier[j] * kWave);
    }
  };
  assign_snaps(all);
  if (ctx->qstate.Ensure(sizeof(EdgeRec) * std::max<int64_t>(q_total, 1)) != hipSuccess ||
      ctx->q_base.Ensure(sizeof(int64_t) * nj) != hipSuccess ||
      ctx->snap_base.Ensure(sizeof(int64_t) * nj) != hipSuccess ||
      ctx->band_count.Ensure(sizeof(int32_t) * nj) != hipSuccess ||
      ctx->snaps.Ensure(std::max<int64_t>(snap_total, 1)) != hipSuccess ||
      ctx->wcov.Ensure(sizeof(JobCov) * nj) != hipSuccess)
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  HIP_TRY(hipMemcpyAsync(ctx->q_base.p, qb.data(), sizeof(int64_t) * nj, hipMemcpyHostToDevice,
                         st));
  HIP_TRY(hipMemcpyAsync(ctx->snap_base.p, sb.data(), sizeof(int64_t) * nj,
                         hipMemcpyHostToDevice, st));
  wp.qstate = ctx->qstate.as<EdgeRec>();
  wp.q_base = ctx->q_base.as<int64_t>();
  wp.snaps = ctx->snaps.as<uint8_t>();
  wp.snap_base = ctx->snap_base.as<int64_t>();
  HIP_TRY(hipMemcpyAsync(ctx->band_count.p, nbands.data(), sizeof(int32_t) * nj,
                         hipMemcpyHostToDevice, st));
  wp.band_count = ctx->band_count.as<int32_t>();
  std::vector<int64_t> pbase(nj, 0);
  std::vector<int32_t> pcap(nj, 0);
  int64_t plane_total = 0;
  auto upload_planes = [&]() -> hipr_result {
    if (plane_total >= kNoPlane) return HIPR_ERROR_CAPACITY;  // plane numbers are int32
    if (ctx->planes.EnsureKeep((size_t)kTilePlaneBytes * std::max<int64_t>(plane_total, 1),
                               ctx->planes.cap, st) != hipSuccess)
      return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
    HIP_TRY(hipMemcpyAsync(ctx->plane_base.p, pbase.data(), sizeof(int64_t) * nj,
                           hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(ctx->plane_cap.p, pcap.data(), sizeof(int32_t) * nj,
                           hipMemcpyHostToDevice, st));
    return HIPR_SUCCESS;
  };
  auto assign_planes = [&](const std::vector<int32_t>& js) {
    for (int32_t j : js) {
      pbase[j] = plane_total;
      plane_total += pcap[j];
    }
  };
  auto read_cov = [&](DevBuf* src, JobCov* dst) -> hipr_result {
    HIP_TRY(hipMemcpyAsync(dst, src->p, sizeof(JobCov) * nj, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return HIPR_SUCCESS;
  };
  auto limit_error = [&](const JobCov* c, const std::vector<int32_t>& js,
                         int mask) -> hipr_result {
    for (int32_t j : js) {
      if (c[j].flags & mask) {
        std::fprintf(stderr, "skity_hip: draw %d exceeds a device limit (flags %d)\n",
                     f.jobs[j].path, c[j].flags);
        return HIPR_ERROR_CAPACITY;
      }
    }
    return HIPR_SUCCESS;
  };
  // Work items of a pass: one per job (count, schedule) or one per row band
  // (write), the bands of the longest walks first so they start first.
  auto make_items = [&](const std::vector<int32_t>& js, int kmode) {
    std::vector<int32_t> it;
    if (kmode != kWalkWrite) {
      for (int32_t j : js) it.push_back(j << kBandBits);
      return it;
    }
    std::vector<std::pair<int64_t, int32_t>> order;
    for (int32_t j : js) {
      PixBounds pb = JobScanBounds(ctx->h_geo[j]);
      int64_t work = (int64_t)std::max(bound[j], 1) * std::max(pb.b - pb.t, 1) / nbands[j];
      order.push_back({-work, j});
    }
    std::stable_sort(order.begin(), order.end());
    for (auto& o : order)
      for (int b = 0; b < nbands[o.second]; b++) it.push_back((o.second << kBandBits) | b);
    return it;
  };
  // One walk pass: the 4-slot jobs on the main stream, the 8- and 16-slot
  // jobs on the side stream; the main stream waits for both.
  // One walk pass: the 4-slot jobs on the main stream, the 8- and 16-slot
  // jobs on the side stream; the main stream waits for both.  The item lists
  // are uploaded first; the launches (queue reset, fork, tier kernels, join)
  // are captured into a hipGraph and launched as one unless the caller asks
  // for direct launches.
  const int run_flags = opt ? opt->flags : 0;
  const bool use_graph = (run_flags & HIPR_FLAG_NO_GRAPH) == 0;
  const bool persistent = (run_flags & HIPR_FLAG_PERSISTENT) != 0;
  // A launch plan: walk kernels, each on one of three streams (0 = main);
  // launches on one stream run in order, the streams concurrently.  The item
  // lists are uploaded first; the launches (queue reset, fork, kernels, join)
  // are captured into a hipGraph and launched as one unless the caller asks
  // for direct launches.
  struct WalkLaunch {
    int kmode, K, stream;
    std::vector<int32_t> jobs;
  };
  auto run_plan = [&](const std::vector<WalkLaunch>& plan) -> hipr_result {
    hipr_result rr;
    ReleaseGraphs(ctx);
    std::vector<int32_t> all_items;
    std::vector<size_t> off, cnt;
    std::vector<double> work;
    double work_all = 0;
    for (const WalkLaunch& l : plan) {
      std::vector<int32_t> it = make_items(l.jobs, l.kmode);
      off.push_back(all_items.size());
      cnt.push_back(it.size());
      all_items.insert(all_items.end(), it.begin(), it.end());
      double w = 0;
      if (persistent) {
        for (int32_t j : l.jobs) {
          PixBounds pb = JobScanBounds(ctx->h_geo[j]);
          w += (double)std::max(bound[j], 1) * std::max(pb.b - pb.t, 1);
        }
      }
      work.push_back(w);
      work_all += w;
    }
    if (all_items.empty()) return HIPR_SUCCESS;
    if ((rr = Upload(&ctx->sub_list, all_items.data(), all_items.size(), st)) != HIPR_SUCCESS)
      return rr;
    if (persistent && ctx->queues.Ensure(sizeof(int32_t) * (plan.size() + 1)) != hipSuccess)
      return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
    int32_t* q = persistent ? ctx->queues.as<int32_t>() : nullptr;
    hipStream_t streams[3] = {st, ctx->side, ctx->side2};
    hipEvent_t joins[3] = {nullptr, ctx->ev_join, ctx->ev_join2};
    bool used[3] = {true, false, false};
    for (size_t i = 0; i < plan.size(); i++)
      if (cnt[i]) used[plan[i].stream] = true;
    const bool timed = !use_graph;
    auto launches = [&]() -> hipr_result {
      if (q)
        hipLaunchKernelGGL(FillKernel, dim3(1), dim3(kWave), 0, st, q, (int64_t)plan.size(), 0);
      HIP_TRY(hipEventRecord(ctx->ev_fork, st));
      for (int k = 1; k < 3; k++)
        if (used[k]) HIP_TRY(hipStreamWaitEvent(streams[k], ctx->ev_fork, 0));
      for (size_t i = 0; i < plan.size(); i++) {
        if (!cnt[i]) continue;
        const WalkLaunch& l = plan[i];
        WalkParams w = wp;
        w.walk_list = ctx->sub_list.as<int32_t>() + off[i];
        w.n_walk = (int32_t)cnt[i];
        w.cov = l.kmode == kWalkWrite ? ctx->wcov.as<JobCov>() : ctx->cov.as<JobCov>();
        int32_t* qi = q ? q + i : nullptr;
        double share = work_all > 0 ? work[i] / work_all : 1.0;
        hipStream_t ls = streams[l.stream];
        hipr_result lr;
#define HIPR_WALK_K(KK)                                                             \
  (l.kmode == kWalkCount   ? LaunchWalk<KK, kWalkCount>(ctx, w, ls, timed, qi, share) \
   : l.kmode == kWalkWrite ? LaunchWalk<KK, kWalkWrite>(ctx, w, ls, timed, qi, share) \
                           : LaunchWalk<KK, kWalkSchedule>(ctx, w, ls, timed, qi, share))
        lr = l.K == 4    ? HIPR_WALK_K(4)
             : l.K == 8  ? HIPR_WALK_K(8)
             : l.K == 16 ? HIPR_WALK_K(16)
                         : LaunchWalkWide(ctx, l.K, l.kmode, w, ls, timed);
#undef HIPR_WALK_K
        if (lr != HIPR_SUCCESS) return lr;
      }
      for (int k = 1; k < 3; k++) {
        if (!used[k]) continue;
        HIP_TRY(hipEventRecord(joins[k], streams[k]));
        HIP_TRY(hipStreamWaitEvent(st, joins[k], 0));
      }
      return HIPR_SUCCESS;
    };
    if (!use_graph) return launches();
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    rr = launches();
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(st, &g);
    if (g) ctx->graphs.push_back(g);
    if (rr != HIPR_SUCCESS) return rr;
    HIP_TRY(ce);
    hipGraphExec_t ge = nullptr;
    HIP_TRY(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    ctx->graph_execs.push_back(ge);
    int stage = HIPR_STAGE_WALK;
    for (const WalkLaunch& l : plan)
      if (l.kmode != kWalkWrite)
        stage = l.kmode == kWalkCount ? HIPR_STAGE_WALK_COUNT : HIPR_STAGE_WALK_SCHEDULE;
    bool any_write = false;
    for (const WalkLaunch& l : plan) any_write |= l.kmode == kWalkWrite;
    if (any_write) stage = HIPR_STAGE_WALK;  // a pass that writes planes counts as the walk
    hipEvent_t b = StageBegin(ctx, st);
    HIP_TRY(hipGraphLaunch(ge, st));
    StageEnd(ctx, stage, b, st);
    return HIPR_SUCCESS;
  };
  // One walk pass of one mode: the 4-slot jobs on the main stream, the 8-
  // and 16-slot jobs on a side stream.
  auto walk_pass = [&](int kmode, const std::vector<int32_t>& sm,
                       const std::vector<int32_t>& wd) -> hipr_result {
    std::vector<int32_t> j8, j16, j32, j64;
    for (int32_t j : wd)
      (tier[j] <= 8 ? j8 : tier[j] == 16 ? j16 : tier[j] == 32 ? j32 : j64).push_back(j);
    return run_plan({{kmode, 64, 2, j64},
                     {kmode, 32, 2, j32},
                     {kmode, 8, 1, j8},
                     {kmode, 16, 1, j16},
                     {kmode, 4, 0, sm}});
  };
  // A 4-slot job whose list still outgrew 256 edges (flag 2) is counted again
  // by the 16-slot walk (recording its snapshots there); the bound makes this
  // a safety net.  `from` holds the flags of the pass that overflowed.
  auto move_overflowed = [&](const JobCov* from, std::vector<int32_t>* moved_out) -> hipr_result {
    std::vector<int32_t> moved, keep;
    for (int32_t j : small) (from[j].flags & 2 ? moved : keep).push_back(j);
    if (moved.empty()) return HIPR_SUCCESS;
    small.swap(keep);
    for (int32_t j : moved) tier[j] = 16;
    assign_snaps(moved);
    if (ctx->snaps.EnsureKeep(std::max<int64_t>(snap_total, 1), ctx->snaps.cap, st) !=
        hipSuccess)
      return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
    wp.snaps = ctx->snaps.as<uint8_t>();
    HIP_TRY(hipMemcpyAsync(ctx->snap_base.p, sb.data(), sizeof(int64_t) * nj,
                           hipMemcpyHostToDevice, st));
    hipr_result rr;
    if ((rr = walk_pass(kWalkCount, {}, moved)) != HIPR_SUCCESS) return rr;
    if ((rr = read_cov(&ctx->cov, ctx->h_cov)) != HIPR_SUCCESS) return rr;
    if ((rr = limit_error(ctx->h_cov, moved, 2 | 16 | 32)) != HIPR_SUCCESS) return rr;
    wide.insert(wide.end(), moved.begin(), moved.end());
    if (moved_out) *moved_out = moved;
    return HIPR_SUCCESS;
  };
  // Before a write pass over `js`: maps at -1 (all jobs) and the jobs'
  // write-pass coverage records at zero.
  auto prepare_write = [&](const std::vector<int32_t>* js) -> hipr_result {
    if (!js) {
      hipLaunchKernelGGL(FillKernel, dim3(1024), dim3(256), 0, st, ctx->map0.as<int32_t>(),
                         total_map, -1);
      hipLaunchKernelGGL(FillKernel, dim3(1024), dim3(256), 0, st, ctx->map1.as<int32_t>(),
                         total_map, -1);
      hipLaunchKernelGGL(FillKernel, dim3(256), dim3(256), 0, st, (int32_t*)ctx->wcov.p,
                         (int64_t)nj * (int64_t)(sizeof(JobCov) / 4), 0);
      hipLaunchKernelGGL(FillKernel, dim3(1), dim3(kWave), 0, st, ctx->pool_ctr.as<int32_t>(),
                         (int64_t)1, 0);
    } else if (!js->empty()) {
      hipr_result rr = Upload(&ctx->gather_list, js->data(), js->size(), st);
      if (rr != HIPR_SUCCESS) return rr;
      hipLaunchKernelGGL(ResetCovKernel, dim3((unsigned)((js->size() + 255) / 256)), dim3(256), 0,
                         st, ctx->gather_list.as<int32_t>(), (int)js->size(),
                         ctx->wcov.as<JobCov>());
    }
    HIP_TRY(hipGetLastError());
    return HIPR_SUCCESS;
  };
  JobCov* hw = ctx->h_wcov;
  if (mode == 0) {
    // count pass over every job (recording the band snapshots), then planes
    // sized exactly
    if ((r = walk_pass(kWalkCount, small, wide)) != HIPR_SUCCESS) return r;
    if ((r = read_cov(&ctx->cov, ctx->h_cov)) != HIPR_SUCCESS) return r;
    if ((r = limit_error(ctx->h_cov, wide, 2)) != HIPR_SUCCESS) return r;
    if ((r = limit_error(ctx->h_cov, all, 16 | 32)) != HIPR_SUCCESS) return r;
    if ((r = move_overflowed(ctx->h_cov, nullptr)) != HIPR_SUCCESS) return r;
    for (int32_t j : all) pcap[j] = ctx->h_cov[j].planes;
  } else {
    // planes from the estimates, at most what the job can use: three planes
    // (a keep-zero triple) per visible tile, and per band that meets the
    // tile
    for (int32_t j : all) {
      const Canvas& cv = f.canvases[f.jobs[j].canvas];
      const TileRange tr = JobTiles(JobScanBounds(ctx->h_geo[j]), cv.width, cv.height);
      const int64_t bands_per_tile =
          nbands[j] > 1 ? 1 + (kTile + std::max(band_rows, 1) - 1) / std::max(band_rows, 1) : 1;
      const int64_t most = 3 * (int64_t)tr.ntc * tr.ntr * bands_per_tile;
      pcap[j] = (int32_t)std::max<int64_t>(std::min<int64_t>(test[j], most), 1);
    }
  }
  assign_planes(all);
  // Estimate modes: a shared pool after the estimated ranges serves the
  // planes of under-estimated jobs; only a job that finds the pool spent is
  // walked again.
  const int64_t pool_base = plane_total;
  const int64_t pool_cap = mode == 0 ? 0 : std::max<int64_t>(plane_total / 4, 1 << 16);
  plane_total += pool_cap;
  if (ctx->pool_ctr.Ensure(sizeof(int32_t)) != hipSuccess) return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  if ((r = upload_planes()) != HIPR_SUCCESS) return r;
  wp.pool_ctr = ctx->pool_ctr.as<int32_t>();
  wp.pool_base = (int32_t)pool_base;
  wp.pool_cap = (int32_t)pool_cap;
  wp.plane_base = ctx->plane_base.as<int64_t>();
  wp.plane_cap = ctx->plane_cap.as<int32_t>();
  wp.map0 = ctx->map0.as<int32_t>();
  wp.map1 = ctx->map1.as<int32_t>();
  wp.planes = ctx->planes.as<uint8_t>();
  wp.cov = ctx->wcov.as<JobCov>();
  if ((r = prepare_write(nullptr)) != HIPR_SUCCESS) return r;
  if (mode == 0) {
    if ((r = walk_pass(kWalkWrite, small, wide)) != HIPR_SUCCESS) return r;
  } else {
    // The banded (longest) walks need a schedule pass before their write
    // pass; both run on side streams while the main stream writes the other
    // jobs, so the schedule pass overlaps the bulk of the writing.
    std::vector<int32_t> b4, u4, b8, u8, b16, u16, u32, u64;
    for (int32_t j : small) (nbands[j] > 1 ? b4 : u4).push_back(j);
    for (int32_t j : wide) {
      if (tier[j] <= 8) {
        (nbands[j] > 1 ? b8 : u8).push_back(j);
      } else if (tier[j] == 16) {
        (nbands[j] > 1 ? b16 : u16).push_back(j);
      } else {
        (tier[j] == 32 ? u32 : u64).push_back(j);
      }
    }
    if ((r = run_plan({{kWalkWrite, 64, 1, u64},
                       {kWalkWrite, 32, 1, u32},
                       {kWalkSchedule, 8, 2, b8},
                       {kWalkWrite, 8, 2, b8},
                       {kWalkSchedule, 16, 2, b16},
                       {kWalkWrite, 16, 2, b16},
                       {kWalkWrite, 8, 2, u8},
                       {kWalkWrite, 16, 2, u16},
                       {kWalkSchedule, 4, 1, b4},
                       {kWalkWrite, 4, 1, b4},
                       {kWalkWrite, 4, 0, u4}})) != HIPR_SUCCESS)
      return r;
  }
  if ((r = read_cov(&ctx->wcov, hw)) != HIPR_SUCCESS) return r;
  if (mode == 0) {
    if ((r = limit_error(hw, all, 1 | 2 | 16 | 32)) != HIPR_SUCCESS) return r;
  } else {
    if ((r = limit_error(hw, wide, 2)) != HIPR_SUCCESS) return r;
    if ((r = limit_error(hw, all, 16 | 32)) != HIPR_SUCCESS) return r;
    // a one-band 4-slot job (not scheduled) whose list outgrew 256 edges is
    // counted by the 16-slot walk and redone below with its exact planes
    std::vector<int32_t> moved;
    if ((r = move_overflowed(hw, &moved)) != HIPR_SUCCESS) return r;
    for (int32_t j : moved) {
      hw[j].planes = ctx->h_cov[j].planes;
      pcap[j] = -1;
    }
    // Under-estimated jobs took their extra planes from the pool.  Exact
    // fallback for the jobs that found it spent (flag 1) and the moved
    // ones: the write pass counted their planes; they get exactly that many,
    // appended, and are walked again.
    for (int32_t j : all)
      if (hw[j].planes > test[j]) f.stats.underestimates++;
    std::vector<int32_t> redo_small, redo_wide;
    for (int32_t j : small)
      if ((hw[j].flags & 1) || pcap[j] < 0) redo_small.push_back(j);
    for (int32_t j : wide)
      if ((hw[j].flags & 1) || pcap[j] < 0) redo_wide.push_back(j);
    if (!redo_small.empty() || !redo_wide.empty()) {
      f.stats.fallbacks += (int64_t)(redo_small.size() + redo_wide.size());
      std::vector<int32_t> redo = redo_small;
      redo.insert(redo.end(), redo_wide.begin(), redo_wide.end());
      for (int32_t j : redo) pcap[j] = hw[j].planes;
      assign_planes(redo);
      if ((r = upload_planes()) != HIPR_SUCCESS) return r;
      wp.planes = ctx->planes.as<uint8_t>();
      if ((r = prepare_write(&redo)) != HIPR_SUCCESS) return r;
      if ((r = walk_pass(kWalkWrite, redo_small, redo_wide)) != HIPR_SUCCESS) return r;
      if ((r = read_cov(&ctx->wcov, hw)) != HIPR_SUCCESS) return r;
      if ((r = limit_error(hw, redo, 1 | 2 | 16 | 32)) != HIPR_SUCCESS) return r;
    }
  }
  int64_t max_active = 0;
  for (int j = 0; j < nj; j++) {
    if (f.jobs[j].kind == kJobRect) continue;
    f.stats.tiles += hw[j].tiles;
    f.stats.tile_planes += hw[j].planes;
    max_active = std::max<int64_t>(max_active, hw[j].max_active);
  }
  f.stats.max_active = max_active;
  surf->job_edges.assign(nj, -1);
  surf->job_planes.assign(nj, -1);
  surf->job_steps.assign(nj, 0);
  surf->job_time.assign(nj, 0);
  surf->job_phase.assign((size_t)nj * 5, 0);
  surf->job_active.assign(nj, 0);
  surf->draw_serial = ++ctx->draw_serial;
  surf->job_path.assign(nj, 0);
  surf->job_kind.assign(nj, 0);
  surf->job_canvas.assign(nj, 0);
  surf->job_ebase.assign(nj, 0);
  surf->job_mbase.assign(nj, 0);
  surf->job_tiles.assign((size_t)nj * 5, 0);
  for (int j = 0; j < nj; j++) {
    surf->job_path[j] = f.jobs[j].path;
    surf->job_kind[j] = f.jobs[j].kind;
    surf->job_canvas[j] = f.jobs[j].canvas;
    if (f.jobs[j].kind == kJobRect) continue;
    surf->job_ebase[j] = ebase[j];
    surf->job_mbase[j] = mbase[j];
    const Canvas& cv = f.canvases[f.jobs[j].canvas];
    TileRange tr = JobTiles(JobScanBounds(ctx->h_geo[j]), cv.width, cv.height);
    int32_t* t = &surf->job_tiles[(size_t)j * 5];
    t[0] = tr.tc0;
    t[1] = tr.tr0;
    t[2] = tr.ntc;
    t[3] = tr.ntr;
    t[4] = cv.tiles_x;
  }
  for (int j = 0; j < nj; j++) {
    if (f.jobs[j].kind == kJobRect) continue;
    surf->job_edges[j] = ctx->h_geo[j].edges;
    surf->job_planes[j] = hw[j].planes;
    surf->job_steps[j] = hw[j].steps;
    surf->job_time[j] = hw[j].time_10ns;
    surf->job_active[j] = hw[j].max_active;
    for (int k = 0; k < 5; k++) surf->job_phase[(size_t)j * 5 + k] = hw[j].phase[k];
  }
  // ---- composite: layers, blur, canvas ----
  CompositeParams cp{};
  cp.n_canvases = (int)f.canvases.size();
  cp.jobs = f.d_jobs;
  cp.cov = ctx->wcov.as<JobCov>();
  cp.map_base = ctx->map_base.as<int64_t>();
  cp.map0 = ctx->map0.as<int32_t>();
  cp.map1 = ctx->map1.as<int32_t>();
  cp.planes = ctx->planes.as<uint8_t>();
  cp.paints = pl->d_paints.as<PaintDev>();
  cp.grads = pl->d_grads.as<GradDev>();
  cp.srgb = ctx->srgb.as<uint8_t>();
  cp.images = pl->d_images.as<ImageDev>();
  cp.image_pixels = pl->image_pixels.as<uint8_t>();
  // per-job tile ranges for the composite's scan
  if (ctx->job_tiles.Ensure(sizeof(int4) * std::max(nj, 1)) != hipSuccess)
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  // (ctx->canvases holds every canvas of the frame, as the walk read them)
  hipLaunchKernelGGL(JobTileRangeKernel, dim3((nj + 255) / 256), dim3(256), 0, st, f.d_jobs,
                     ctx->wcov.as<JobCov>(), pl->d_paints.as<PaintDev>(),
                     ctx->canvases.as<Canvas>(), nj, ctx->job_tiles.as<int4>());
  HIP_TRY(hipGetLastError());
  cp.job_tiles = ctx->job_tiles.as<int4>();
  // the composite variant with colour filters and pixmap sampling only when
  // a paint of the frame needs them
  bool rich = false;
  for (const PaintDev& pd : pl->paints)
    rich |= pd.cf_type != HIPR_COLOR_FILTER_NONE || pd.kind == kPaintPixmap;
  auto composite = [&](const CompositeParams& p, int tiles) {
    if (rich) {
      hipLaunchKernelGGL(CompositeKernel<true>, dim3(tiles), dim3(64), 0, st, p);
    } else {
      hipLaunchKernelGGL(CompositeKernel<false>, dim3(tiles), dim3(64), 0, st, p);
    }
  };
  int main_tiles = 0;
  for (int t = 0; t < n_targets; t++) main_tiles += f.canvases[t].tiles_x * f.canvases[t].tiles_y;
  if ((int)f.canvases.size() > n_targets) {
    std::vector<Canvas> lc(f.canvases.begin() + n_targets, f.canvases.end());
    int layer_tiles = 0;
    for (auto& c : lc) {
      c.tile_begin -= main_tiles;
      layer_tiles += c.tiles_x * c.tiles_y;
    }
    if ((r = Upload(&ctx->layer_canvases, lc.data(), lc.size(), st)) != HIPR_SUCCESS) return r;
    CompositeParams lp = cp;
    lp.canvases = ctx->layer_canvases.as<Canvas>();
    lp.n_canvases = (int)lc.size();
    lp.total_tiles = layer_tiles;
    STAGE_LAUNCH(ctx, HIPR_STAGE_LAYERS, composite(lp, layer_tiles));
    HIP_TRY(hipGetLastError());
    // Layers of the blur styles other than kNormal keep their unblurred
    // pixels for the style pass.
    std::vector<BlurStyleJob> styled;
    int64_t raw_bytes = 0;
    for (size_t b = 0; b < pl->blurs.size(); b++)
      if (pl->blur_style[b] != HIPR_BLUR_NORMAL)
        raw_bytes += (int64_t)pl->blurs[b].width * pl->blurs[b].height * 4;
    if (raw_bytes > 0) {
      if (ctx->blur_raw.Ensure(raw_bytes) != hipSuccess) return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
      int64_t roff = 0;
      for (size_t b = 0; b < pl->blurs.size(); b++) {
        if (pl->blur_style[b] == HIPR_BLUR_NORMAL) continue;
        const int64_t n = (int64_t)pl->blurs[b].width * pl->blurs[b].height;
        uint8_t* raw = ctx->blur_raw.as<uint8_t>() + roff;
        uint8_t* layer_px = f.canvases[pl->layer_of_blur[b]].pixels;
        HIP_TRY(hipMemcpyAsync(raw, layer_px, n * 4, hipMemcpyDeviceToDevice, st));
        styled.push_back(BlurStyleJob{raw, layer_px, n, pl->blur_style[b], 0});
        roff += n * 4;
      }
    }
    // StackBlur each layer: rows into a temporary, columns back
    std::vector<BlurJob> rows, cols;
    std::vector<int64_t> rb{0}, cb{0};
    int64_t tmp_bytes = 0;
    for (auto& b : pl->blurs) tmp_bytes += (int64_t)b.width * b.height * 4;
    if (ctx->blur_tmp.Ensure(std::max<int64_t>(tmp_bytes, 1)) != hipSuccess)
      return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
    int64_t off = 0;
    for (size_t b = 0; b < pl->blurs.size(); b++) {
      BlurJob bj = pl->blurs[b];
      uint8_t* layer_px = f.canvases[pl->layer_of_blur[b]].pixels;
      uint8_t* tmp = ctx->blur_tmp.as<uint8_t>() + off;
      off += (int64_t)bj.width * bj.height * 4;
      if (bj.radius <= 1) continue;  // SWStackBlur copies for radius <= 1
      BlurJob h = bj, v = bj;
      h.src = layer_px;
      h.dst = tmp;
      v.src = tmp;
      v.dst = layer_px;
      rows.push_back(h);
      cols.push_back(v);
      rb.push_back(rb.back() + bj.height);
      cb.push_back(cb.back() + bj.width);
    }
    if (!rows.empty() && blur_impl == HIPR_BLUR_IMPL_MFMA) {
      if ((r = LaunchBlurMfma(ctx, rows, cols, &ctx->blur_rows, &ctx->blur_cols, st, true)) !=
          HIPR_SUCCESS)
        return r;
    } else if (!rows.empty()) {
      if ((r = Upload(&ctx->blur_rows, rows.data(), rows.size(), st)) != HIPR_SUCCESS) return r;
      if ((r = Upload(&ctx->blur_cols, cols.data(), cols.size(), st)) != HIPR_SUCCESS) return r;
      if ((r = Upload(&ctx->blur_rb, rb.data(), rb.size(), st)) != HIPR_SUCCESS) return r;
      if ((r = Upload(&ctx->blur_cb, cb.data(), cb.size(), st)) != HIPR_SUCCESS) return r;
      int64_t nl = rb.back(), nc = cb.back();
      STAGE_LAUNCH(ctx, HIPR_STAGE_BLUR,
    hipLaunchKernelGGL(StackBlurLineKernel, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, st,
                         ctx->blur_rows.as<BlurJob>(), ctx->blur_rb.as<int64_t>(),
                         (int)rows.size(), nl, 0));
      HIP_TRY(hipGetLastError());
      STAGE_LAUNCH(ctx, HIPR_STAGE_BLUR,
    hipLaunchKernelGGL(StackBlurLineKernel, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, st,
                         ctx->blur_cols.as<BlurJob>(), ctx->blur_cb.as<int64_t>(),
                         (int)cols.size(), nc, 1));
      HIP_TRY(hipGetLastError());
    }
    if (!styled.empty()) {
      if ((r = Upload(&ctx->blur_styles, styled.data(), styled.size(), st)) != HIPR_SUCCESS)
        return r;
      dim3 grid(256, (unsigned)std::min<size_t>(styled.size(), 65535));
      STAGE_LAUNCH(ctx, HIPR_STAGE_BLUR,
    hipLaunchKernelGGL(BlurStyleKernel, grid, dim3(256), 0, st,
                         ctx->blur_styles.as<BlurStyleJob>(), (int)styled.size()));
      HIP_TRY(hipGetLastError());
    }
  }
  CompositeParams mp = cp;
  mp.canvases = ctx->canvases.as<Canvas>();
  mp.n_canvases = n_targets;
  mp.total_tiles = main_tiles;
  if (ctx->counters.Ensure(2 * sizeof(unsigned long long)) != hipSuccess)
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  HIP_TRY(hipMemsetAsync(ctx->counters.p, 0, 2 * sizeof(unsigned long long), st));
  mp.counters = ctx->counters.as<unsigned long long>();
  STAGE_LAUNCH(ctx, HIPR_STAGE_COMPOSITE, composite(mp, main_tiles));
  HIP_TRY(hipGetLastError());
  unsigned long long counts[2] = {0, 0};
  HIP_TRY(hipMemcpyAsync(counts, ctx->counters.p, sizeof(counts), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipEventRecord(ctx->ev1, st));
  HIP_TRY(hipEventSynchronize(ctx->ev1));
  f.stats.blends = (int64_t)counts[0];
  f.stats.pixels = (int64_t)counts[1];
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  f.stats.ms_total = ms;
  for (const auto& m : ctx->marks) {
    float sm = 0;
    if (hipEventElapsedTime(&sm, m.begin, m.end) == hipSuccess) {
      surf->stage_ms[m.stage] += sm;
      surf->stage_launches[m.stage]++;
    }
  }
  surf->stats = f.stats;
  return HIPR_SUCCESS;
}

hipr_result LaunchBlurMfma(hipr_context ctx, const std::vector<BlurJob>& rows,
                           const std::vector<BlurJob>& cols, DevBuf* d_rows, DevBuf* d_cols,
                           hipStream_t st, bool timed) {
  auto tile_jobs = [](const std::vector<BlurJob>& js, int pass, int64_t* total) {
    std::vector<BlurTileJob> out;
    int64_t t = 0;
    for (const BlurJob& b : js) {
      BlurTileJob tj{};
      tj.b = b;
      tj.tile_begin = t;
      int n_lines = pass == 0 ? b.height : b.width, len = pass == 0 ? b.width : b.height;
      tj.line_tiles = (n_lines + kBlurLines - 1) / kBlurLines;
      tj.out_tiles = (len + kBlurOut - 1) / kBlurOut;
      t += (int64_t)tj.line_tiles * tj.out_tiles;
      out.push_back(tj);
    }
    *total = t;
    return out;
  };
  int64_t tr = 0, tc = 0;
  std::vector<BlurTileJob> jr = tile_jobs(rows, 0, &tr), jc = tile_jobs(cols, 1, &tc);
  hipr_result r;
  if ((r = Upload(d_rows, jr.data(), jr.size(), st)) != HIPR_SUCCESS) return r;
  if ((r = Upload(d_cols, jc.data(), jc.size(), st)) != HIPR_SUCCESS) return r;
  unsigned gr = (unsigned)std::min<int64_t>(std::max<int64_t>(tr, 1), 1 << 16);
  unsigned gc = (unsigned)std::min<int64_t>(std::max<int64_t>(tc, 1), 1 << 16);
  hipEvent_t b0 = timed ? StageBegin(ctx, st) : nullptr;
  hipLaunchKernelGGL(StackBlurMfmaKernel<0>, dim3(gr), dim3(64), 0, st,
                     d_rows->as<BlurTileJob>(), (int)jr.size(), tr);
  HIP_TRY(hipGetLastError());
  if (timed) StageEnd(ctx, HIPR_STAGE_BLUR, b0, st);
  hipEvent_t b1 = timed ? StageBegin(ctx, st) : nullptr;
  hipLaunchKernelGGL(StackBlurMfmaKernel<1>, dim3(gc), dim3(64), 0, st,
                     d_cols->as<BlurTileJob>(), (int)jc.size(), tc);
  HIP_TRY(hipGetLastError());
  if (timed) StageEnd(ctx, HIPR_STAGE_BLUR, b1, st);
  return HIPR_SUCCESS;
}

// Splits a plan into sub-plans of consecutive jobs within the budgets.  Each
// sub-plan borrows the plan's geometry and paints, keeps the global path
// indices (per-path estimates and paints apply unchanged) and draws on the
// same targets; drawn in order, the sub-plans blend every pixel in the
// plan's order.  Plans with mask-filter layers are not split: a layer is
// composited, blurred and drawn within one run.
hipr_result BuildChunks(hipr_prepared_s* pl, int64_t budget_points) {
  const int64_t kChunkPoints = budget_points;
  const int64_t kChunkTiles = kChunkTilesPerPoint * budget_points;
  pl->chunks_built = true;
  pl->chunk_budget = budget_points;
  pl->chunks.clear();
  const int nj = (int)pl->jobs.size();
  if (pl->layer_bytes > 0 || nj == 0 || (int)pl->job_points.size() != nj) return HIPR_SUCCESS;
  int64_t tp = 0, tt = 0;
  for (int j = 0; j < nj; j++) {
    tp += pl->job_points[j];
    tt += pl->job_tiles[j];
  }
  if (tp <= kChunkPoints && tt <= kChunkTiles) return HIPR_SUCCESS;
  std::vector<int> cut{0};
  int64_t cp = 0, ct = 0;
  for (int j = 0; j < nj; j++) {
    if (j > cut.back() && (cp + pl->job_points[j] > kChunkPoints ||
                           ct + pl->job_tiles[j] > kChunkTiles)) {
      cut.push_back(j);
      cp = ct = 0;
    }
    cp += pl->job_points[j];
    ct += pl->job_tiles[j];
  }
  cut.push_back(nj);
  hipStream_t st = pl->ctx->stream;
  for (size_t c = 0; c + 1 < cut.size(); c++) {
    const int j0 = cut[c], j1 = cut[c + 1];
    auto sub = std::make_unique<hipr_prepared_s>();
    sub->ctx = pl->ctx;
    sub->n_paths = pl->n_paths;
    sub->n_targets = pl->n_targets;
    sub->borrowed = true;
    sub->chunks_built = true;
    sub->verbs = pl->verbs;
    sub->pts = pl->pts;
    sub->weights = pl->weights;
    sub->scene_paints = pl->scene_paints;
    sub->d_paints = pl->d_paints;
    sub->d_grads = pl->d_grads;
    sub->d_images = pl->d_images;
    sub->image_pixels = pl->image_pixels;
    sub->images = pl->images;
    sub->image_bytes = pl->image_bytes;
    sub->paints = pl->paints;
    sub->n_verbs = pl->n_verbs;
    sub->n_pts = pl->n_pts;
    sub->n_weights = pl->n_weights;
    sub->job_offset = j0;
    sub->jobs.assign(pl->jobs.begin() + j0, pl->jobs.begin() + j1);
    sub->job_points.assign(pl->job_points.begin() + j0, pl->job_points.begin() + j1);
    sub->job_tiles.assign(pl->job_tiles.begin() + j0, pl->job_tiles.begin() + j1);
    sub->canvases.assign(pl->canvases.begin(), pl->canvases.begin() + pl->n_targets);
    for (Canvas& cv : sub->canvases) {
      cv.job_begin = std::min(std::max(cv.job_begin, j0), j1) - j0;
      cv.job_end = std::min(std::max(cv.job_end, j0), j1) - j0;
    }
    for (int32_t j : pl->walk_jobs)
      if (j >= j0 && j < j1) sub->walk_jobs.push_back(j - j0);
    for (size_t k = 0; k < pl->stroke_jobs.size(); k++) {
      const int32_t j = pl->stroke_jobs[k];
      if (j < j0 || j >= j1) continue;
      sub->stroke_jobs.push_back(j - j0);
      if (k < pl->stroke_quads.size()) sub->stroke_quads.push_back(pl->stroke_quads[k]);
    }
    hipr_result r;
    if ((r = Upload(&sub->d_jobs, sub->jobs.data(), sub->jobs.size(), st)) != HIPR_SUCCESS ||
        (r = Upload(&sub->d_walk, sub->walk_jobs.data(), sub->walk_jobs.size(), st)) !=
            HIPR_SUCCESS ||
        (r = Upload(&sub->d_stroke, sub->stroke_jobs.data(), sub->stroke_jobs.size(), st)) !=
            HIPR_SUCCESS)
      return r;
    pl->chunks.push_back(std::move(sub));
  }
  HIP_TRY(hipStreamSynchronize(st));
  return HIPR_SUCCESS;
}

// Draws a plan, in one pipeline run or, when it exceeds the sub-plan
// budgets, one run per sub-plan; the statistics and stage times are summed
// over the runs (`pixels` then counts each run's touched pixels).
hipr_result RunPlan(const std::vector<hipr_surface>& targets, hipr_surface surf,
                    hipr_prepared_s* pl, const hipr_options* opt) {
  hipr_result r;
  const int64_t budget = opt && opt->chunk_points > 0 ? opt->chunk_points : kChunkPointsDefault;
  if ((!pl->chunks_built || pl->chunk_budget != budget) &&
      (r = BuildChunks(pl, budget)) != HIPR_SUCCESS)
    return r;
  if (pl->chunks.empty()) return RunPipeline(targets, surf, pl, opt);
  hipr_stats sum{};
  double ms[HIPR_STAGE_COUNT] = {};
  int32_t launches[HIPR_STAGE_COUNT] = {};
  // per-job records of the whole plan, gathered from the sub-plan runs
  const size_t nj = pl->jobs.size();
  std::vector<int32_t> edges(nj, -1), planes(nj, -1), steps(nj, 0), tms(nj, 0), act(nj, 0);
  std::vector<int32_t> phase(nj * 5, 0);
  for (auto& sub : pl->chunks) {
    if ((r = RunPipeline(targets, surf, sub.get(), opt)) != HIPR_SUCCESS) return r;
    const size_t o = (size_t)sub->job_offset, n = sub->jobs.size();
    if (surf->job_edges.size() == n) {
      std::copy(surf->job_edges.begin(), surf->job_edges.end(), edges.begin() + o);
      std::copy(surf->job_planes.begin(), surf->job_planes.end(), planes.begin() + o);
      std::copy(surf->job_steps.begin(), surf->job_steps.end(), steps.begin() + o);
      std::copy(surf->job_time.begin(), surf->job_time.end(), tms.begin() + o);
      std::copy(surf->job_active.begin(), surf->job_active.end(), act.begin() + o);
      std::copy(surf->job_phase.begin(), surf->job_phase.end(), phase.begin() + 5 * o);
    }
    const hipr_stats& st = surf->stats;
    sum.raster_paths += st.raster_paths;
    sum.edges += st.edges;
    sum.tiles += st.tiles;
    sum.tile_planes += st.tile_planes;
    sum.fallbacks += st.fallbacks;
    sum.max_active = std::max(sum.max_active, st.max_active);
    sum.ms_total += st.ms_total;
    sum.blends += st.blends;
    sum.pixels += st.pixels;
    sum.underestimates += st.underestimates;
    for (int k = 0; k < HIPR_STAGE_COUNT; k++) {
      ms[k] += surf->stage_ms[k];
      launches[k] += surf->stage_launches[k];
    }
  }
  sum.paths = pl->n_paths;
  surf->stats = sum;
  for (int k = 0; k < HIPR_STAGE_COUNT; k++) {
    surf->stage_ms[k] = ms[k];
    surf->stage_launches[k] = launches[k];
  }
  surf->job_edges.swap(edges);
  surf->job_planes.swap(planes);
  surf->job_steps.swap(steps);
  surf->job_time.swap(tms);
  surf->job_active.swap(act);
  surf->job_phase.swap(phase);
  // edge records and tile maps of earlier sub-plans are gone
  surf->draw_serial = 0;
  return HIPR_SUCCESS;
}

hipr_result Prepare(hipr_context ctx, const hipr_scene* scene, hipr_prepared_s** out) {
  hipr_result vr = ValidateScene(scene);
  if (vr != HIPR_SUCCESS) return vr;
  hipr_prepared_s* pl = new (std::nothrow) hipr_prepared_s();
  if (!pl) return HIPR_ERROR_OUT_OF_HOST_MEMORY;
  pl->ctx = ctx;
  pl->n_paths = scene->n_paths;
  hipStream_t st = ctx->stream;
  hipr_result r = HIPR_SUCCESS;
  if (scene->n_paths > 0) {
    BuildPlan(scene, pl);
    int n = scene->n_paths;
    int nv = scene->verb_offsets[n], np = scene->point_offsets[n], nw = scene->weight_offsets[n];
    for (int32_t j : pl->stroke_jobs) pl->stroke_quads.push_back(QuadPathSegments(scene, pl->jobs[j].path));
    LongestStrokesFirst(pl);
    // the scene's geometry, then the rectangles the plan added
    std::vector<uint8_t> all_verbs;
    std::vector<float> all_pts;
    const uint8_t* vsrc = scene->verbs;
    const float* psrc = scene->points;
    if (!pl->x_verbs.empty()) {
      all_verbs.assign(scene->verbs, scene->verbs + nv);
      all_verbs.insert(all_verbs.end(), pl->x_verbs.begin(), pl->x_verbs.end());
      all_pts.assign(scene->points, scene->points + 2 * (size_t)np);
      all_pts.insert(all_pts.end(), pl->x_pts.begin(), pl->x_pts.end());
      vsrc = all_verbs.data();
      psrc = all_pts.data();
      nv += (int)pl->x_verbs.size();
      np += (int)(pl->x_pts.size() / 2);
    }
    pl->n_verbs = nv;
    pl->n_pts = np;
    pl->n_weights = nw;
    if ((r = Upload(&pl->verbs, vsrc, nv, st)) != HIPR_SUCCESS ||
        (r = Upload(&pl->pts, psrc, 2 * (size_t)np, st)) != HIPR_SUCCESS ||
        (r = Upload(&pl->weights, scene->weights, nw, st)) != HIPR_SUCCESS ||
        (r = Upload(&pl->scene_paints, scene->paints, n, st)) != HIPR_SUCCESS ||
        (r = Upload(&pl->d_jobs, pl->jobs.data(), pl->jobs.size(), st)) != HIPR_SUCCESS ||
        (r = Upload(&pl->d_paints, pl->paints.data(), pl->paints.size(), st)) != HIPR_SUCCESS ||
        (r = Upload(&pl->d_grads, pl->grads.data(), pl->grads.size(), st)) != HIPR_SUCCESS ||
        (r = Upload(&pl->d_walk, pl->walk_jobs.data(), pl->walk_jobs.size(), st)) != HIPR_SUCCESS ||
        (r = Upload(&pl->d_stroke, pl->stroke_jobs.data(), pl->stroke_jobs.size(), st)) != HIPR_SUCCESS ||
        (r = Upload(&pl->d_images, pl->images.data(), pl->images.size(), st)) != HIPR_SUCCESS) {
      delete pl;
      return r;
    }
    if (pl->image_bytes > 0) {
      if (pl->image_pixels.Ensure(pl->image_bytes) != hipSuccess) {
        delete pl;
        return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
      }
      for (size_t i = 0; i < pl->images.size(); i++) {
        const ImageDev& d = pl->images[i];
        if (hipMemcpyAsync(pl->image_pixels.as<uint8_t>() + d.offset, scene->images[i].pixels,
                           (size_t)d.width * d.height * 4, hipMemcpyHostToDevice,
                           st) != hipSuccess) {
          delete pl;
          return HIPR_ERROR_DEVICE;
        }
      }
    }
    if (hipStreamSynchronize(st) != hipSuccess) {
      delete pl;
      return HIPR_ERROR_DEVICE;
    }
  }
  ctx->refs++;  // released by hipr_prepared_destroy
  *out = pl;
  return HIPR_SUCCESS;
}

// Concatenates the plans of a batch's pairs into b->merged: job, paint,
// gradient, canvas and path indices and geometry offsets are rebased; the
// geometry is gathered with device-to-device copies.
hipr_result MergeBatch(hipr_batch b) {
  hipr_context ctx = b->ctx;
  hipStream_t st = ctx->stream;
  hipr_prepared_s& m = b->merged;
  m.ctx = ctx;
  m.jobs.clear();
  m.canvases.clear();
  m.paints.clear();
  m.grads.clear();
  m.blurs.clear();
  m.layer_of_blur.clear();
  m.blur_style.clear();
  m.job_points.clear();
  m.job_tiles.clear();
  m.chunks.clear();
  m.chunks_built = false;
  m.walk_jobs.clear();
  m.stroke_jobs.clear();
  m.stroke_quads.clear();
  const int T = (int)b->pairs.size();
  m.n_targets = T;
  std::vector<int> layer_base(T);
  int n_layers = 0;
  for (int p = 0; p < T; p++) {
    layer_base[p] = T + n_layers;
    const hipr_prepared_s* s = b->pairs[p].second;
    n_layers += s->canvases.empty() ? 0 : (int)s->canvases.size() - 1;
  }
  m.canvases.assign(T + n_layers, Canvas{});
  int64_t voff = 0, poff = 0, woff = 0, lbytes = 0;
  int path_off = 0, paint_off = 0, grad_off = 0, job_off = 0, image_off = 0;
  int64_t image_byte_off = 0;
  m.images.clear();
  for (int p = 0; p < T; p++) {
    const hipr_prepared_s* s = b->pairs[p].second;
    auto canv = [&](int c) { return c == 0 ? p : layer_base[p] + c - 1; };
    if (voff + s->n_verbs > INT32_MAX || poff + s->n_pts > INT32_MAX ||
        woff + s->n_weights > INT32_MAX)
      return HIPR_ERROR_CAPACITY;
    for (const Job& src : s->jobs) {
      Job jb = src;
      jb.path += path_off;
      jb.canvas = canv(jb.canvas);
      if (jb.kind != kJobRect) {
        jb.verb_begin += (int32_t)voff;
        jb.verb_end += (int32_t)voff;
        jb.pt_begin += (int32_t)poff;
        jb.w_begin += (int32_t)woff;
      }
      if (jb.image >= 0) jb.image = canv(jb.image);
      jb.paint += paint_off;
      m.jobs.push_back(jb);
    }
    for (const PaintDev& src : s->paints) {
      PaintDev q = src;
      if (q.kind == kPaintGradient) q.grad += grad_off;
      if (q.kind == kPaintImage) q.layer = canv(q.layer);
      if (q.kind == kPaintPixmap) q.pix_img += image_off;
      m.paints.push_back(q);
    }
    for (ImageDev d : s->images) {
      d.offset += image_byte_off;
      m.images.push_back(d);
    }
    m.grads.insert(m.grads.end(), s->grads.begin(), s->grads.end());
    for (size_t c = 0; c < s->canvases.size(); c++) {
      Canvas cv = s->canvases[c];
      cv.job_begin += job_off;
      cv.job_end += job_off;
      if (c > 0) cv.pixels = (uint8_t*)((intptr_t)cv.pixels + lbytes);
      m.canvases[canv((int)c)] = cv;
    }
    if (s->canvases.empty()) {  // an empty scene draws nothing on its target
      m.canvases[p].job_begin = m.canvases[p].job_end = job_off;
    }
    m.blurs.insert(m.blurs.end(), s->blurs.begin(), s->blurs.end());
    for (int32_t l : s->layer_of_blur) m.layer_of_blur.push_back(canv(l));
    m.blur_style.insert(m.blur_style.end(), s->blur_style.begin(), s->blur_style.end());
    m.job_points.insert(m.job_points.end(), s->job_points.begin(), s->job_points.end());
    m.job_tiles.insert(m.job_tiles.end(), s->job_tiles.begin(), s->job_tiles.end());
    for (int32_t j : s->walk_jobs) m.walk_jobs.push_back(j + job_off);
    for (int32_t j : s->stroke_jobs) m.stroke_jobs.push_back(j + job_off);
    m.stroke_quads.insert(m.stroke_quads.end(), s->stroke_quads.begin(), s->stroke_quads.end());
    voff += s->n_verbs;
    poff += s->n_pts;
    woff += s->n_weights;
    lbytes += s->layer_bytes;
    path_off += s->n_paths;
    paint_off += (int)s->paints.size();
    grad_off += (int)s->grads.size();
    job_off += (int)s->jobs.size();
    image_off += (int)s->images.size();
    image_byte_off += s->image_bytes;
  }
  m.image_bytes = image_byte_off;
  if (image_byte_off > 0) {
    if (m.image_pixels.Ensure(image_byte_off) != hipSuccess)
      return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
    int64_t o = 0;
    for (int p = 0; p < T; p++) {
      const hipr_prepared_s* s = b->pairs[p].second;
      if (s->image_bytes > 0)
        HIP_TRY(hipMemcpyAsync(m.image_pixels.as<uint8_t>() + o, s->image_pixels.p, s->image_bytes,
                               hipMemcpyDeviceToDevice, st));
      o += s->image_bytes;
    }
  }
  m.n_paths = path_off;
  m.layer_bytes = lbytes;
  m.n_verbs = voff;
  m.n_pts = poff;
  m.n_weights = woff;
  if (m.verbs.Ensure(std::max<int64_t>(voff, 1)) != hipSuccess ||
      m.pts.Ensure(8 * std::max<int64_t>(poff, 1)) != hipSuccess ||
      m.weights.Ensure(4 * std::max<int64_t>(woff, 1)) != hipSuccess ||
      m.scene_paints.Ensure(sizeof(hipr_paint) * std::max(path_off, 1)) != hipSuccess)
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  int64_t v = 0, q = 0, w = 0, pa = 0;
  for (int p = 0; p < T; p++) {
    const hipr_prepared_s* s = b->pairs[p].second;
    if (s->n_verbs)
      HIP_TRY(hipMemcpyAsync(m.verbs.as<uint8_t>() + v, s->verbs.p, s->n_verbs,
                             hipMemcpyDeviceToDevice, st));
    if (s->n_pts)
      HIP_TRY(hipMemcpyAsync(m.pts.as<uint8_t>() + 8 * q, s->pts.p, 8 * s->n_pts,
                             hipMemcpyDeviceToDevice, st));
    if (s->n_weights)
      HIP_TRY(hipMemcpyAsync(m.weights.as<uint8_t>() + 4 * w, s->weights.p, 4 * s->n_weights,
                             hipMemcpyDeviceToDevice, st));
    if (s->n_paths)
      HIP_TRY(hipMemcpyAsync(m.scene_paints.as<hipr_paint>() + pa, s->scene_paints.p,
                             sizeof(hipr_paint) * s->n_paths, hipMemcpyDeviceToDevice, st));
    v += s->n_verbs;
    q += s->n_pts;
    w += s->n_weights;
    pa += s->n_paths;
  }
  LongestStrokesFirst(&m);
  hipr_result r;
  if ((r = Upload(&m.d_jobs, m.jobs.data(), m.jobs.size(), st)) != HIPR_SUCCESS ||
      (r = Upload(&m.d_paints, m.paints.data(), m.paints.size(), st)) != HIPR_SUCCESS ||
      (r = Upload(&m.d_grads, m.grads.data(), m.grads.size(), st)) != HIPR_SUCCESS ||
      (r = Upload(&m.d_images, m.images.data(), m.images.size(), st)) != HIPR_SUCCESS ||
      (r = Upload(&m.d_walk, m.walk_jobs.data(), m.walk_jobs.size(), st)) != HIPR_SUCCESS ||
      (r = Upload(&m.d_stroke, m.stroke_jobs.data(), m.stroke_jobs.size(), st)) != HIPR_SUCCESS)
    return r;
  HIP_TRY(hipStreamSynchronize(st));
  b->dirty = false;
  return HIPR_SUCCESS;
}

constexpr int kMlpSamplesPerWave = 256;

// One forward/backward pass over n device-resident samples into
// ctx->mlp_grads / mlp_loss; returns the number of waves used.
hipr_result LaunchGrad(hipr_context ctx, const float* d_feats, const float* d_targets, int64_t n,
                       int* waves_out) {
  int waves = (int)((n + kMlpSamplesPerWave - 1) / kMlpSamplesPerWave);
  waves = std::max(waves, 1);
  if (ctx->mlp_grads.Ensure(sizeof(float) * kMlpTrained * waves) != hipSuccess ||
      ctx->mlp_loss.Ensure(sizeof(float) * waves) != hipSuccess)
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  hipLaunchKernelGGL(MlpGradKernel, dim3(waves), dim3(64), 0, ctx->stream,
                     ctx->mlp_ops.as<_Float16>(), ctx->mlp_prm.as<float>(), d_feats, d_targets, n,
                     kMlpSamplesPerWave, 1.0f / (float)std::max<int64_t>(n, 1),
                     ctx->mlp_grads.as<float>(), ctx->mlp_loss.as<float>());
  HIP_TRY(hipGetLastError());
  *waves_out = waves;
  return HIPR_SUCCESS;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

HIPR_API hipr_result hipr_context_create(int32_t device, hipr_context* out) {
  if (!out) return HIPR_ERROR_INVALID_ARGUMENT;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return HIPR_ERROR_INITIALIZATION_FAILED;
  if (device < 0 || device >= n) return HIPR_ERROR_INVALID_ARGUMENT;
  hipr_context ctx = new (std::nothrow) hipr_context_s();
  if (!ctx) return HIPR_ERROR_OUT_OF_HOST_MEMORY;
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess ||
      hipDeviceGetAttribute(&ctx->n_cus, hipDeviceAttributeMultiprocessorCount, device) !=
          hipSuccess ||
      hipDeviceGetAttribute(&ctx->max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) !=
          hipSuccess ||
      hipStreamCreate(&ctx->stream) != hipSuccess ||
      hipStreamCreate(&ctx->side) != hipSuccess || hipStreamCreate(&ctx->side2) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->ev_join2, hipEventDisableTiming) != hipSuccess ||
      hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
    delete ctx;
    return HIPR_ERROR_INITIALIZATION_FAILED;
  }
  // The sRGB tables of SRGBGammaColorFilter (color_filter.cc:128-166) from
  // their definition: floor(255 f(i / 255)) with f the sRGB encoding
  // (linear -> sRGB) or decoding, in double (a 1e-6 nudge keeps f(1) = 1 at
  // 255).
  uint8_t tab[512];
  for (int i = 0; i < 256; i++) {
    double x = i / 255.0;
    double e = x <= 0.0031308 ? 12.92 * x : 1.055 * std::pow(x, 1 / 2.4) - 0.055;
    double d = x <= 0.04045 ? x / 12.92 : std::pow((x + 0.055) / 1.055, 2.4);
    tab[i] = (uint8_t)std::floor(255 * e + 1e-6);
    tab[256 + i] = (uint8_t)std::floor(255 * d + 1e-6);
  }
  if (ctx->srgb.Ensure(sizeof(tab)) != hipSuccess ||
      hipMemcpy(ctx->srgb.p, tab, sizeof(tab), hipMemcpyHostToDevice) != hipSuccess) {
    delete ctx;
    return HIPR_ERROR_INITIALIZATION_FAILED;
  }
  *out = ctx;
  return HIPR_SUCCESS;
}

static void FreeContext(hipr_context ctx);

// A context is released when its handle and every surface and prepared scene
// made from it are destroyed, in any order.
static void ReleaseContext(hipr_context ctx) {
  if (ctx && --ctx->refs == 0) FreeContext(ctx);
}

HIPR_API void hipr_context_destroy(hipr_context ctx) { ReleaseContext(ctx); }

static void FreeContext(hipr_context ctx) {
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  DevBuf* bufs[] = {&ctx->job_tiles, &ctx->canvases, &ctx->layer_canvases, &ctx->geo, &ctx->cov, &ctx->edge_base,
                    &ctx->edge_cap, &ctx->edges, &ctx->wedges, &ctx->gkeys, &ctx->plane_base,
                    &ctx->plane_cap, &ctx->planes, &ctx->map_base, &ctx->map0, &ctx->map1,
                    &ctx->sub_list, &ctx->big_list, &ctx->bound, &ctx->gather_list, &ctx->stroke_out, &ctx->o_verbs, &ctx->o_pts, &ctx->o_w,
                    &ctx->s_verbs, &ctx->s_pts, &ctx->s_w, &ctx->layers, &ctx->blur_rows,
                    &ctx->blur_cols, &ctx->blur_rb, &ctx->blur_cb, &ctx->blur_tmp,
                    &ctx->blur_jobs, &ctx->blur_lines, &ctx->d_verbs, &ctx->d_pts, &ctx->d_w,
                    &ctx->d_jobs, &ctx->run_jobs, &ctx->counters, &ctx->wcov,
                    &ctx->qstate, &ctx->q_base, &ctx->snaps, &ctx->snap_base,
                    &ctx->band_count, &ctx->queues, &ctx->flatten_redo, &ctx->pool_ctr, &ctx->stroke_redo,
                    &ctx->stroke_redo_out, &ctx->blur_raw, &ctx->blur_styles};
  ReleaseGraphs(ctx);
  for (DevBuf* b : bufs) b->Free();
  DevBuf* mlp[] = {&ctx->mlp_prm, &ctx->mlp_ops, &ctx->mlp_feats, &ctx->mlp_preds,
                   &ctx->mlp_targets, &ctx->mlp_grads, &ctx->mlp_loss, &ctx->mlp_m,
                   &ctx->mlp_v, &ctx->mlp_gsum, &ctx->mlp_curve};
  for (DevBuf* b : mlp) b->Free();
  if (ctx->h_geo) (void)hipHostFree(ctx->h_geo);
  if (ctx->h_cov) (void)hipHostFree(ctx->h_cov);
  if (ctx->h_wcov) (void)hipHostFree(ctx->h_wcov);
  if (ctx->h_so) (void)hipHostFree(ctx->h_so);
  for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
  (void)hipStreamSynchronize(ctx->side);
  (void)hipEventDestroy(ctx->ev0);
  (void)hipEventDestroy(ctx->ev1);
  (void)hipEventDestroy(ctx->ev_fork);
  (void)hipEventDestroy(ctx->ev_join);
  (void)hipEventDestroy(ctx->ev_join2);
  (void)hipStreamSynchronize(ctx->side2);
  (void)hipStreamDestroy(ctx->side);
  (void)hipStreamDestroy(ctx->side2);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

HIPR_API hipr_result hipr_surface_create(hipr_context ctx, int32_t width, int32_t height,
                                         hipr_surface* out) {
  return hipr_surface_create_ex(ctx, width, height, HIPR_ALPHA_PREMUL, HIPR_COLOR_RGBA, out);
}

HIPR_API hipr_result hipr_surface_create_ex(hipr_context ctx, int32_t width, int32_t height,
                                            int32_t alpha_type, int32_t color_type,
                                            hipr_surface* out) {
  if (!out) return HIPR_ERROR_INVALID_ARGUMENT;
  *out = nullptr;
  if (!ctx) return HIPR_ERROR_INVALID_HANDLE;
  if (width <= 0 || height <= 0) return HIPR_ERROR_INVALID_ARGUMENT;
  if (width > kMaxCanvasSide || height > kMaxCanvasSide) return HIPR_ERROR_NOT_SUPPORTED;
  if (alpha_type < HIPR_ALPHA_OPAQUE || alpha_type > HIPR_ALPHA_UNPREMUL ||
      (color_type != HIPR_COLOR_RGBA && color_type != HIPR_COLOR_BGRA))
    return HIPR_ERROR_INVALID_ARGUMENT;
  hipr_surface s = new (std::nothrow) hipr_surface_s();
  if (!s) return HIPR_ERROR_OUT_OF_HOST_MEMORY;
  s->ctx = ctx;
  s->format = (alpha_type != HIPR_ALPHA_PREMUL ? kFormatUnpremul : 0) |
              (color_type == HIPR_COLOR_BGRA ? kFormatBGRA : 0);
  s->width = width;
  s->height = height;
  if (hipSetDevice(ctx->device) != hipSuccess ||
      hipMalloc(&s->pixels, (size_t)width * height * 4) != hipSuccess) {
    delete s;
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  }
  if (hipMemsetAsync(s->pixels, 0, (size_t)width * height * 4, ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess) {
    (void)hipFree(s->pixels);
    delete s;
    return HIPR_ERROR_DEVICE;
  }
  ctx->refs++;  // released by hipr_surface_destroy
  *out = s;
  return HIPR_SUCCESS;
}

HIPR_API void hipr_surface_destroy(hipr_surface s) {
  if (!s) return;
  hipr_context ctx = s->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(s->pixels);
  delete s;
  ReleaseContext(ctx);
}

HIPR_API hipr_result hipr_surface_clear(hipr_surface s, uint32_t rgba) {
  if (!s) return HIPR_ERROR_INVALID_HANDLE;
  if (rgba == 0) {
    HIP_TRY(hipMemsetAsync(s->pixels, 0, (size_t)s->width * s->height * 4, s->ctx->stream));
  } else {
    HIP_TRY(hipMemsetD32Async((hipDeviceptr_t)s->pixels, rgba, (size_t)s->width * s->height,
                              s->ctx->stream));
  }
  return HIPR_SUCCESS;
}

HIPR_API hipr_result hipr_surface_read_pixels(hipr_surface s, uint8_t* dst, size_t bytes) {
  if (!s) return HIPR_ERROR_INVALID_HANDLE;
  if (!dst || bytes < (size_t)s->width * s->height * 4) return HIPR_ERROR_INVALID_ARGUMENT;
  HIP_TRY(hipMemcpyAsync(dst, s->pixels, (size_t)s->width * s->height * 4,
                         hipMemcpyDeviceToHost, s->ctx->stream));
  HIP_TRY(hipStreamSynchronize(s->ctx->stream));
  return HIPR_SUCCESS;
}

HIPR_API hipr_result hipr_surface_write_pixels(hipr_surface s, const uint8_t* src, size_t bytes) {
  if (!s) return HIPR_ERROR_INVALID_HANDLE;
  if (!src || bytes < (size_t)s->width * s->height * 4) return HIPR_ERROR_INVALID_ARGUMENT;
  HIP_TRY(hipMemcpyAsync(s->pixels, src, (size_t)s->width * s->height * 4,
                         hipMemcpyHostToDevice, s->ctx->stream));
  HIP_TRY(hipStreamSynchronize(s->ctx->stream));
  return HIPR_SUCCESS;
}

HIPR_API hipr_result hipr_surface_finish(hipr_surface s) {
  if (!s) return HIPR_ERROR_INVALID_HANDLE;
  HIP_TRY(hipStreamSynchronize(s->ctx->stream));
  return HIPR_SUCCESS;
}

HIPR_API hipr_result hipr_surface_get_stats(hipr_surface s, hipr_stats* out) {
  if (!s) return HIPR_ERROR_INVALID_HANDLE;
  if (!out) return HIPR_ERROR_INVALID_ARGUMENT;
  *out = s->stats;
  return HIPR_SUCCESS;
}

HIPR_API hipr_result hipr_surface_get_stage_times(hipr_surface s, double* ms,
                                                  int32_t* launches, int32_t n_stages) {
  if (!s) return HIPR_ERROR_INVALID_HANDLE;
  if (n_stages < 0 || (n_stages > 0 && !ms && !launches)) return HIPR_ERROR_INVALID_ARGUMENT;
  for (int k = 0; k < n_stages; k++) {
    bool in = k < HIPR_STAGE_COUNT;
    if (ms) ms[k] = in ? s->stage_ms[k] : 0.0;
    if (launches) launches[k] = in ? s->stage_launches[k] : 0;
  }
  return HIPR_SUCCESS;
}

HIPR_API hipr_result hipr_prepared_job_count(hipr_prepared p, int32_t* n_jobs) {
  if (!p) return HIPR_ERROR_INVALID_HANDLE;
  if (!n_jobs) return HIPR_ERROR_INVALID_ARGUMENT;
  *n_jobs = (int32_t)p->jobs.size();
  return HIPR_SUCCESS;
}

HIPR_API hipr_result hipr_prepared_features(hipr_prepared p, int32_t width, int32_t height,
                                            float* features, int32_t n_jobs) {
  if (!p) return HIPR_ERROR_INVALID_HANDLE;
  if (!features || n_jobs != (int32_t)p->jobs.size() || width <= 0 || height <= 0)
    return HIPR_ERROR_INVALID_ARGUMENT;
  hipr_context ctx = p->ctx;
  HIP_TRY(hipSetDevice(ctx->device));
  if (n_jobs == 0) return HIPR_SUCCESS;
  std::vector<Canvas> cvs = p->canvases;
  cvs[0].width = width;
  cvs[0].height = height;
  SetTiles(&cvs);
  hipr_result r;
  if ((r = Upload(&ctx->canvases, cvs.data(), cvs.size(), ctx->stream)) != HIPR_SUCCESS) return r;
  if ((r = LaunchFeatures(ctx, p, ctx->canvases.as<Canvas>())) != HIPR_SUCCESS) return r;
  HIP_TRY(hipMemcpyAsync(features, ctx->mlp_feats.p, sizeof(float) * kMlpIn * n_jobs,
                         hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return HIPR_SUCCESS;
}

HIPR_API hipr_result hipr_surface_get_job_counts(hipr_surface s, int32_t* edges,
                                                 int32_t* planes, int32_t n_jobs) {
  if (!s) return HIPR_ERROR_INVALID_HANDLE;
  if (n_jobs != (int32_t)s->job_edges.size()) return HIPR_ERROR_INVALID_ARGUMENT;
  for (int j = 0; j < n_jobs; j++) {
    if (edges) edges[j] = s->job_edges[j];
    if (planes) planes[j] = s->job_planes[j];
  }
  return HIPR_SUCCESS;
}

HIPR_API hipr_result hipr_surface_get_job_profile(hipr_surface s, int32_t* steps,
                                                  int32_t* time_10ns, int32_t* phases,
                                                  int32_t* max_active, int32_t n_jobs) {
  if (!s) return HIPR_ERROR_INVALID_HANDLE;
  if (n_jobs != (int32_t)s->job_steps.size()) return HIPR_ERROR_INVALID_ARGUMENT;
  for (int j = 0; j < n_jobs; j++) {
    if (steps) steps[j] = s->job_steps[j];
    if (time_10ns) time_10ns[j] = s->job_time[j];
    if (max_active) max_active[j] = s->job_active[j];
    if (phases)
      for (int k = 0; k < 5; k++) phases[(size_t)j * 5 + k] = s->job_phase[(size_t)j * 5 + k];
  }
  return HIPR_SUCCESS;
}

HIPR_API hipr_result hipr_surface_get_jobs(hipr_surface s, int32_t* path, int32_t* kind,
                                          int32_t* canvas, int32_t cap, int32_t* n_jobs) {
  if (!s) return HIPR_ERROR_INVALID_HANDLE;
  if (!n_jobs) return HIPR_ERROR_INVALID_ARGUMENT;
  *n_jobs = (int32_t)s->job_path.size();
  if (cap < 0) return HIPR_ERROR_INVALID_ARGUMENT;
  for (int j = 0; j < std::min(cap, *n_jobs); j++) {
    if (path) path[j] = s->job_path[j];
    if (kind) kind[j] = s->job_kind[j];
    if (canvas) canvas[j] = s->job_canvas[j];
  }
  return HIPR_SUCCESS;
}

HIPR_API hipr_result hipr_surface_get_job_edges(hipr_surface s, int32_t job, int32_t* edges,
                                               int64_t cap, int64_t* n_edges) {
  if (!s) return HIPR_ERROR_INVALID_HANDLE;
  if (!n_edges || job < 0 || job >= (int32_t)s->job_path.size())
    return HIPR_ERROR_INVALID_ARGUMENT;
  hipr_context ctx = s->ctx;
  if (s->draw_serial == 0 || s->draw_serial != ctx->draw_serial)
    return HIPR_ERROR_INVALID_ARGUMENT;  // another draw reused the buffers
  const int64_t n = s->job_kind[job] == kJobRect ? 0 : s->job_edges[job];
  *n_edges = n;
  if (!edges || n == 0) return HIPR_SUCCESS;
  if (cap < n) return HIPR_ERROR_CAPACITY;
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  HIP_TRY(hipMemcpy(edges, ctx->edges.as<EdgeRec>() + s->job_ebase[job], sizeof(EdgeRec) * n,
                    hipMemcpyDeviceToHost));
  return HIPR_SUCCESS;
}

HIPR_API hipr_result hipr_surface_get_tile_bins(hipr_surface s, int64_t* offsets, int32_t* bins,
                                               int64_t cap, int32_t n_jobs) {
  if (!s) return HIPR_ERROR_INVALID_HANDLE;
  if (!offsets || n_jobs != (int32_t)s->job_path.size()) return HIPR_ERROR_INVALID_ARGUMENT;
  hipr_context ctx = s->ctx;
  if (s->draw_serial == 0 || s->draw_serial != ctx->draw_serial)
    return HIPR_ERROR_INVALID_ARGUMENT;
  int64_t total_map = 0;
  for (int j = 0; j < n_jobs; j++) {
    const int32_t* t = &s->job_tiles[(size_t)j * 5];
    total_map = std::max<int64_t>(total_map, s->job_mbase[j] + (int64_t)t[2] * t[3]);
  }
  std::vector<int32_t> m0(total_map), m1(total_map);
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (total_map > 0) {
    HIP_TRY(hipMemcpy(m0.data(), ctx->map0.p, sizeof(int32_t) * total_map, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(m1.data(), ctx->map1.p, sizeof(int32_t) * total_map, hipMemcpyDeviceToHost));
  }
  int64_t k = 0;
  for (int j = 0; j < n_jobs; j++) {
    offsets[j] = k;
    if (s->job_kind[j] == kJobRect) continue;
    const int32_t* t = &s->job_tiles[(size_t)j * 5];
    for (int r = 0; r < t[3]; r++) {
      for (int c = 0; c < t[2]; c++) {
        int64_t i = s->job_mbase[j] + (int64_t)r * t[2] + c;
        int32_t code = ((t[1] + r) * t[4] + (t[0] + c)) << 1;
        if (m0[i] >= 0) {
          if (bins && k < cap) bins[k] = code;
          k++;
        }
        if (m1[i] >= 0) {
          if (bins && k < cap) bins[k] = code | 1;
          k++;
        }
      }
    }
  }
  offsets[n_jobs] = k;
  return bins && k > cap ? HIPR_ERROR_CAPACITY : HIPR_SUCCESS;
}

HIPR_API hipr_result hipr_presizer_predict(hipr_context ctx, const float* params,
                                           const float* features, int64_t n, float* out) {
  if (!ctx) return HIPR_ERROR_INVALID_HANDLE;
  if (!params || n < 0 || (n > 0 && (!features || !out))) return HIPR_ERROR_INVALID_ARGUMENT;
  HIP_TRY(hipSetDevice(ctx->device));
  if (n == 0) return HIPR_SUCCESS;
  hipr_result r;
  if ((r = UploadPresizer(ctx, params)) != HIPR_SUCCESS) return r;
  if ((r = Upload(&ctx->mlp_feats, features, (size_t)n * kMlpIn, ctx->stream)) != HIPR_SUCCESS)
    return r;
  if (ctx->mlp_preds.Ensure(sizeof(float) * kMlpOut * n) != hipSuccess)
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  int blocks = (int)std::min<int64_t>((n + 15) / 16, 8192);
  hipLaunchKernelGGL(MlpInferKernel, dim3(blocks), dim3(64), 0, ctx->stream,
                     ctx->mlp_ops.as<_Float16>(), ctx->mlp_prm.as<float>(),
                     ctx->mlp_feats.as<float>(), n, ctx->mlp_preds.as<float>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, ctx->mlp_preds.p, sizeof(float) * kMlpOut * n,
                         hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx->mlp_dev_current = false;  // the installed pre-sizer is re-uploaded on use
  return HIPR_SUCCESS;
}

HIPR_API hipr_result hipr_presizer_gradients(hipr_context ctx, const float* params,
                                             const float* features, const float* targets,
                                             int64_t n, float* grads, float* loss) {
  if (!ctx) return HIPR_ERROR_INVALID_HANDLE;
  if (!params || !features || !targets || !grads || n <= 0) return HIPR_ERROR_INVALID_ARGUMENT;
  HIP_TRY(hipSetDevice(ctx->device));
  hipr_result r;
  if ((r = UploadPresizer(ctx, params)) != HIPR_SUCCESS) return r;
  if ((r = Upload(&ctx->mlp_feats, features, (size_t)n * kMlpIn, ctx->stream)) != HIPR_SUCCESS ||
      (r = Upload(&ctx->mlp_targets, targets, (size_t)n * kMlpOut, ctx->stream)) != HIPR_SUCCESS)
    return r;
  int waves = 0;
  if ((r = LaunchGrad(ctx, ctx->mlp_feats.as<float>(), ctx->mlp_targets.as<float>(), n, &waves)) !=
      HIPR_SUCCESS)
    return r;
  if (ctx->mlp_gsum.Ensure(sizeof(float) * (kMlpTrained + 1)) != hipSuccess)
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  float* gsum = ctx->mlp_gsum.as<float>();
  hipLaunchKernelGGL(MlpAdamKernel, dim3((kMlpTrained + 255) / 256), dim3(256), 0, ctx->stream,
                     ctx->mlp_prm.as<float>(), (float*)nullptr, (float*)nullptr,
                     ctx->mlp_grads.as<float>(), waves, gsum, 0.f, 0.f, 0.f, 0.f, 1.f, 1.f,
                     ctx->mlp_loss.as<float>(), 1.0f / (float)n, gsum + kMlpTrained);
  HIP_TRY(hipGetLastError());
  std::vector<float> h(kMlpTrained + 1);
  HIP_TRY(hipMemcpyAsync(h.data(), gsum, sizeof(float) * h.size(), hipMemcpyDeviceToHost,
                         ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  std::memcpy(grads, h.data(), sizeof(float) * kMlpTrained);
  if (loss) *loss = h[kMlpTrained];
  ctx->mlp_dev_current = false;
  return HIPR_SUCCESS;
}

HIPR_API hipr_result hipr_presizer_train(hipr_context ctx, float* params, const float* features,
                                         const float* targets, int64_t n, int32_t steps,
                                         int32_t batch, float lr, uint64_t seed,
                                         float* loss_curve) {
  if (!ctx) return HIPR_ERROR_INVALID_HANDLE;
  if (!params || !features || !targets || n <= 0 || steps < 0 || batch <= 0 || !(lr > 0))
    return HIPR_ERROR_INVALID_ARGUMENT;
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  hipr_result r;
  if ((r = UploadPresizer(ctx, params)) != HIPR_SUCCESS) return r;
  if (ctx->mlp_m.Ensure(sizeof(float) * kMlpTrained) != hipSuccess ||
      ctx->mlp_v.Ensure(sizeof(float) * kMlpTrained) != hipSuccess ||
      ctx->mlp_curve.Ensure(sizeof(float) * std::max(steps, 1)) != hipSuccess ||
      ctx->mlp_feats.Ensure(sizeof(float) * kMlpIn * n) != hipSuccess ||
      ctx->mlp_targets.Ensure(sizeof(float) * kMlpOut * n) != hipSuccess)
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  HIP_TRY(hipMemsetAsync(ctx->mlp_m.p, 0, sizeof(float) * kMlpTrained, st));
  HIP_TRY(hipMemsetAsync(ctx->mlp_v.p, 0, sizeof(float) * kMlpTrained, st));
  int64_t bsz = std::min<int64_t>(batch, n);
  int64_t n_batches = std::max<int64_t>(n / bsz, 1);
  std::vector<int64_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::mt19937_64 rng(seed);
  std::vector<float> hf((size_t)n * kMlpIn), ht((size_t)n * kMlpOut);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int32_t t = 0; t < steps; t++) {
    int64_t b = t % n_batches;
    if (b == 0) {  // new epoch: shuffle and upload
      std::shuffle(perm.begin(), perm.end(), rng);
      for (int64_t i = 0; i < n; i++) {
        std::memcpy(&hf[i * kMlpIn], features + perm[i] * kMlpIn, sizeof(float) * kMlpIn);
        std::memcpy(&ht[i * kMlpOut], targets + perm[i] * kMlpOut, sizeof(float) * kMlpOut);
      }
      HIP_TRY(hipMemcpyAsync(ctx->mlp_feats.p, hf.data(), sizeof(float) * hf.size(),
                             hipMemcpyHostToDevice, st));
      HIP_TRY(hipMemcpyAsync(ctx->mlp_targets.p, ht.data(), sizeof(float) * ht.size(),
                             hipMemcpyHostToDevice, st));
      HIP_TRY(hipStreamSynchronize(st));
    }
    int waves = 0;
    if ((r = LaunchGrad(ctx, ctx->mlp_feats.as<float>() + b * bsz * kMlpIn,
                        ctx->mlp_targets.as<float>() + b * bsz * kMlpOut, bsz, &waves)) !=
        HIPR_SUCCESS)
      return r;
    float corr1 = 1.f - std::pow(b1, (float)(t + 1)), corr2 = 1.f - std::pow(b2, (float)(t + 1));
    hipLaunchKernelGGL(MlpAdamKernel, dim3((kMlpTrained + 255) / 256), dim3(256), 0, st,
                       ctx->mlp_prm.as<float>(), ctx->mlp_m.as<float>(), ctx->mlp_v.as<float>(),
                       ctx->mlp_grads.as<float>(), waves, (float*)nullptr, lr, b1, b2, eps, corr1,
                       corr2, ctx->mlp_loss.as<float>(), 1.0f / (float)bsz,
                       ctx->mlp_curve.as<float>() + t);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(MlpPackKernel, dim3(kOpBlocks), dim3(256), 0, st, ctx->mlp_prm.as<float>(),
                       ctx->mlp_ops.as<_Float16>());
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipMemcpyAsync(params, ctx->mlp_prm.p, sizeof(float) * kMlpTrained,
                         hipMemcpyDeviceToHost, st));
  if (loss_curve && steps > 0)
    HIP_TRY(hipMemcpyAsync(loss_curve, ctx->mlp_curve.p, sizeof(float) * steps,
                           hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  ctx->mlp_dev_current = false;
  return HIPR_SUCCESS;
}

HIPR_API hipr_result hipr_context_set_presizer(hipr_context ctx, const float* params,
                                               float margin_edges, float margin_planes) {
  if (!ctx) return HIPR_ERROR_INVALID_HANDLE;
  if (!params || !std::isfinite(margin_edges) || !std::isfinite(margin_planes))
    return HIPR_ERROR_INVALID_ARGUMENT;
  for (int i = 0; i < kMlpTotal; i++)
    if (!std::isfinite(params[i])) return HIPR_ERROR_INVALID_ARGUMENT;
  HIP_TRY(hipSetDevice(ctx->device));
  hipr_result r;
  if ((r = UploadPresizer(ctx, params)) != HIPR_SUCCESS) return r;
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx->presizer_host.assign(params, params + kMlpTotal);
  ctx->mlp_dev_current = true;
  ctx->has_presizer = true;
  ctx->margin_edges = margin_edges;
  ctx->margin_planes = margin_planes;
  return HIPR_SUCCESS;
}

HIPR_API hipr_result hipr_scene_prepare(hipr_context ctx, const hipr_scene* scene,
                                        hipr_prepared* out) {
  if (!out) return HIPR_ERROR_INVALID_ARGUMENT;
  *out = nullptr;
  if (!ctx) return HIPR_ERROR_INVALID_HANDLE;
  HIP_TRY(hipSetDevice(ctx->device));
  return Prepare(ctx, scene, out);
}

HIPR_API void hipr_prepared_destroy(hipr_prepared p) {
  if (!p) return;
  hipr_context ctx = p->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete p;
  ReleaseContext(ctx);
}

HIPR_API hipr_result hipr_surface_draw_prepared(hipr_surface surf, hipr_prepared p,
                                                const hipr_options* options) {
  if (!surf || !p) return HIPR_ERROR_INVALID_HANDLE;
  if (p->ctx != surf->ctx) return HIPR_ERROR_INVALID_ARGUMENT;
  HIP_TRY(hipSetDevice(surf->ctx->device));
  return RunPlan({surf}, surf, p, options);
}

HIPR_API hipr_result hipr_batch_create(hipr_context ctx, hipr_batch* out) {
  if (!out) return HIPR_ERROR_INVALID_ARGUMENT;
  *out = nullptr;
  if (!ctx) return HIPR_ERROR_INVALID_HANDLE;
  hipr_batch b = new (std::nothrow) hipr_batch_s();
  if (!b) return HIPR_ERROR_OUT_OF_HOST_MEMORY;
  b->ctx = ctx;
  b->merged.ctx = ctx;
  b->sink.ctx = ctx;
  ctx->refs++;  // released by hipr_batch_destroy
  *out = b;
  return HIPR_SUCCESS;
}

HIPR_API void hipr_batch_destroy(hipr_batch b) {
  if (!b) return;
  hipr_context ctx = b->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete b;
  ReleaseContext(ctx);
}

HIPR_API hipr_result hipr_batch_add(hipr_batch b, hipr_surface surf, hipr_prepared scene) {
  if (!b || !surf || !scene) return HIPR_ERROR_INVALID_HANDLE;
  if (surf->ctx != b->ctx || scene->ctx != b->ctx) return HIPR_ERROR_INVALID_ARGUMENT;
  for (const auto& pr : b->pairs)
    if (pr.first == surf) return HIPR_ERROR_INVALID_ARGUMENT;  // one scene per surface
  if (b->pairs.size() >= (size_t)(1 << 20)) return HIPR_ERROR_CAPACITY;
  b->pairs.push_back({surf, scene});
  b->dirty = true;
  return HIPR_SUCCESS;
}

HIPR_API hipr_result hipr_batch_clear(hipr_batch b) {
  if (!b) return HIPR_ERROR_INVALID_HANDLE;
  b->pairs.clear();
  b->dirty = true;
  return HIPR_SUCCESS;
}

HIPR_API hipr_result hipr_batch_draw(hipr_batch b, const hipr_options* options) {
  if (!b) return HIPR_ERROR_INVALID_HANDLE;
  if (b->pairs.empty()) return HIPR_SUCCESS;
  HIP_TRY(hipSetDevice(b->ctx->device));
  hipr_result r;
  if (b->dirty && (r = MergeBatch(b)) != HIPR_SUCCESS) return r;
  hipr_options opt{};
  if (options) opt = *options;
  if (opt.band_rows == 0) opt.band_rows = b->tuned_band_rows;
  if (opt.band_min_edges == 0) opt.band_min_edges = b->tuned_band_min_edges;
  if (opt.flags == 0) opt.flags = b->tuned_flags;
  std::vector<hipr_surface> targets;
  for (const auto& pr : b->pairs) targets.push_back(pr.first);
  return RunPlan(targets, &b->sink, &b->merged, &opt);
}

HIPR_API hipr_result hipr_batch_set_split(hipr_batch b, int32_t band_rows, int32_t band_min_edges,
                                          int32_t flags) {
  if (!b) return HIPR_ERROR_INVALID_HANDLE;
  b->tuned_band_rows = band_rows;
  b->tuned_band_min_edges = band_min_edges;
  b->tuned_flags = flags;
  return HIPR_SUCCESS;
}

HIPR_API hipr_result hipr_batch_autotune(hipr_batch b, int32_t mode, int32_t repeats,
                                         int32_t* out_band_rows, int32_t* out_band_min_edges,
                                         int32_t* out_flags) {
  if (!b) return HIPR_ERROR_INVALID_HANDLE;
  if (repeats < 1 || repeats > 100 || mode < HIPR_MODE_EXACT || mode > HIPR_MODE_LEARNED ||
      mode == HIPR_MODE_ESTIMATES)
    return HIPR_ERROR_INVALID_ARGUMENT;
  if (b->pairs.empty()) return HIPR_SUCCESS;
  // Work splits: rows per walk band, and (estimate modes) which paths are
  // banded.  Each candidate is drawn once to warm up, then `repeats` times;
  // the device time of its fastest draw decides.
  const int32_t rows[] = {-1, 64, 128, 256};
  const int32_t mins[] = {INT32_MAX, 4 * kWave + 1, 3 * kWave + 1, 2 * kWave + 1, -1};
  struct Cand {
    int32_t rows, min_edges, flags;
  };
  std::vector<Cand> cands;
  for (int32_t f : {0, (int32_t)HIPR_FLAG_PERSISTENT}) {
    for (int32_t r : rows) {
      if (mode == HIPR_MODE_EXACT || r < 0) {
        cands.push_back({r, 0, f});
      } else {
        for (int32_t m : mins) cands.push_back({r, m, f});
      }
    }
  }
  double best_ms = 1e30;
  Cand best = cands[0];
  for (const Cand& c : cands) {
    hipr_options opt{};
    opt.mode = mode;
    opt.band_rows = c.rows;
    opt.band_min_edges = c.min_edges;
    opt.flags = c.flags;
    double t = 1e30;
    for (int k = 0; k <= repeats; k++) {
      hipr_result r = hipr_batch_draw(b, &opt);
      if (r != HIPR_SUCCESS) return r;
      if (k > 0) t = std::min(t, b->sink.stats.ms_total);
    }
    if (t < best_ms) {
      best_ms = t;
      best = c;
    }
  }
  b->tuned_band_rows = best.rows;
  b->tuned_band_min_edges = best.min_edges;
  b->tuned_flags = best.flags;
  if (out_band_rows) *out_band_rows = best.rows;
  if (out_band_min_edges) *out_band_min_edges = best.min_edges;
  if (out_flags) *out_flags = best.flags;
  return HIPR_SUCCESS;
}

HIPR_API hipr_result hipr_batch_get_stats(hipr_batch b, hipr_stats* out) {
  if (!b) return HIPR_ERROR_INVALID_HANDLE;
  return hipr_surface_get_stats(&b->sink, out);
}

HIPR_API hipr_result hipr_batch_get_stage_times(hipr_batch b, double* ms, int32_t* launches,
                                                int32_t n_stages) {
  if (!b) return HIPR_ERROR_INVALID_HANDLE;
  return hipr_surface_get_stage_times(&b->sink, ms, launches, n_stages);
}

HIPR_API hipr_result hipr_surface_draw_scene(hipr_surface surf, const hipr_scene* scene,
                                             const hipr_options* options) {
  if (!surf) return HIPR_ERROR_INVALID_HANDLE;
  hipr_prepared p = nullptr;
  HIP_TRY(hipSetDevice(surf->ctx->device));
  hipr_result r = Prepare(surf->ctx, scene, &p);
  if (r != HIPR_SUCCESS) return r;
  r = RunPlan({surf}, surf, p, options);
  hipr_prepared_destroy(p);
  return r;
}

HIPR_API hipr_result hipr_debug_build_edges(hipr_context ctx, const hipr_scene* scene,
                                            int32_t* edge_offsets, int32_t* edges,
                                            int64_t edge_cap) {
  if (!ctx) return HIPR_ERROR_INVALID_HANDLE;
  if (!edge_offsets) return HIPR_ERROR_INVALID_ARGUMENT;
  hipr_result vr = ValidateScene(scene);
  if (vr != HIPR_SUCCESS && vr != HIPR_ERROR_NOT_SUPPORTED) return vr;
  hipStream_t st = ctx->stream;
  HIP_TRY(hipSetDevice(ctx->device));
  int n = scene->n_paths;
  std::vector<Job> jobs(n);
  for (int p = 0; p < n; p++) {
    Job jb{};
    jb.path = p;
    jb.kind = kJobFill;
    jb.verb_begin = scene->verb_offsets[p];
    jb.verb_end = scene->verb_offsets[p + 1];
    jb.pt_begin = scene->point_offsets[p];
    jb.w_begin = scene->weight_offsets[p];
    jb.m = ToAffine(DrawMatrix(scene, p));
    jobs[p] = jb;
  }
  int nv = n ? scene->verb_offsets[n] : 0, np = n ? scene->point_offsets[n] : 0;
  int nw = n ? scene->weight_offsets[n] : 0;
  hipr_result r;
  if ((r = Upload(&ctx->d_verbs, scene->verbs, nv, st)) != HIPR_SUCCESS) return r;
  if ((r = Upload(&ctx->d_pts, scene->points, 2 * (size_t)np, st)) != HIPR_SUCCESS) return r;
  if ((r = Upload(&ctx->d_w, scene->weights, nw, st)) != HIPR_SUCCESS) return r;
  if ((r = Upload(&ctx->d_jobs, jobs.data(), jobs.size(), st)) != HIPR_SUCCESS) return r;
  if (ctx->geo.Ensure(sizeof(JobGeo) * std::max(n, 1)) != hipSuccess ||
      ctx->edge_base.Ensure(sizeof(int64_t) * std::max(n, 1)) != hipSuccess ||
      ctx->edge_cap.Ensure(sizeof(int32_t) * std::max(n, 1)) != hipSuccess)
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  if ((r = HostStage(ctx, n)) != HIPR_SUCCESS) return r;
  GeomSrc gs{ctx->d_verbs.as<uint8_t>(), ctx->d_pts.as<float>(), ctx->d_w.as<float>()};
  if (n > 0) {
    hipLaunchKernelGGL((FlattenKernel<0>), dim3(n), dim3(64), 0, st, ctx->d_jobs.as<Job>(), n, gs,
                       gs, ctx->edge_base.as<int64_t>(), ctx->edge_cap.as<int32_t>(),
                       (EdgeRec*)nullptr, ctx->geo.as<JobGeo>(), nullptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(ctx->h_geo, ctx->geo.p, sizeof(JobGeo) * n, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
  }
  std::vector<int64_t> base(n + 1, 0);
  std::vector<int32_t> cap(n, 0);
  for (int p = 0; p < n; p++) {
    cap[p] = ctx->h_geo[p].edges;
    base[p + 1] = base[p] + cap[p];
  }
  for (int p = 0; p <= n; p++) edge_offsets[p] = (int32_t)base[p];
  int64_t total = base[n];
  if (!edges || total > edge_cap) return total > edge_cap ? HIPR_ERROR_CAPACITY : HIPR_SUCCESS;
  if (ctx->edges.Ensure(sizeof(EdgeRec) * std::max<int64_t>(total, 1)) != hipSuccess)
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  HIP_TRY(hipMemcpyAsync(ctx->edge_base.p, base.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(ctx->edge_cap.p, cap.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
  if (n > 0) {
    hipLaunchKernelGGL((FlattenKernel<1>), dim3(n), dim3(64), 0, st, ctx->d_jobs.as<Job>(), n, gs,
                       gs, ctx->edge_base.as<int64_t>(), ctx->edge_cap.as<int32_t>(),
                       ctx->edges.as<EdgeRec>(), ctx->geo.as<JobGeo>(), nullptr);
    HIP_TRY(hipGetLastError());
  }
  if (total > 0)
    HIP_TRY(hipMemcpyAsync(edges, ctx->edges.p, sizeof(EdgeRec) * total, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return HIPR_SUCCESS;
}

HIPR_API hipr_result hipr_debug_wave_ops(hipr_context ctx, const int32_t* in, int32_t* out,
                                         int32_t n_waves) {
  if (!ctx) return HIPR_ERROR_INVALID_HANDLE;
  if (!in || !out || n_waves <= 0) return HIPR_ERROR_INVALID_ARGUMENT;
  HIP_TRY(hipSetDevice(ctx->device));
  hipr_result r;
  if ((r = Upload(&ctx->d_verbs, (const uint8_t*)in, sizeof(int32_t) * 64 * n_waves,
                  ctx->stream)) != HIPR_SUCCESS)
    return r;
  if (ctx->d_pts.Ensure(sizeof(int32_t) * 6 * 64 * n_waves) != hipSuccess)
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  hipLaunchKernelGGL(WaveOpsCheckKernel, dim3(n_waves), dim3(64), 0, ctx->stream,
                     ctx->d_verbs.as<int>(), ctx->d_pts.as<int>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, ctx->d_pts.p, sizeof(int32_t) * 6 * 64 * n_waves,
                         hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return HIPR_SUCCESS;
}

HIPR_API hipr_result hipr_stackblur_ex(hipr_context ctx, const uint8_t* src_host,
                                       uint8_t* dst_host, int32_t width, int32_t height,
                                       int32_t radius, int32_t impl) {
  if (impl == HIPR_BLUR_IMPL_LINES) return hipr_stackblur(ctx, src_host, dst_host, width, height, radius);
  if (!ctx) return HIPR_ERROR_INVALID_HANDLE;
  if (impl != HIPR_BLUR_IMPL_MFMA || !src_host || !dst_host || width <= 0 || height <= 0)
    return HIPR_ERROR_INVALID_ARGUMENT;
  hipStream_t st = ctx->stream;
  HIP_TRY(hipSetDevice(ctx->device));
  size_t bytes = (size_t)width * height * 4;
  radius = std::min(radius, 254);
  if (ctx->blur_tmp.Ensure(bytes * 2) != hipSuccess) return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  uint8_t* a = ctx->blur_tmp.as<uint8_t>();
  uint8_t* b = a + bytes;
  HIP_TRY(hipMemcpyAsync(a, src_host, bytes, hipMemcpyHostToDevice, st));
  if (radius > 1) {
    std::vector<BlurJob> rows{{a, b, width, height, radius}}, cols{{b, a, width, height, radius}};
    hipr_result r;
    if ((r = LaunchBlurMfma(ctx, rows, cols, &ctx->blur_jobs, &ctx->blur_lines, st, false)) !=
        HIPR_SUCCESS)
      return r;
  }
  HIP_TRY(hipMemcpyAsync(dst_host, a, bytes, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return HIPR_SUCCESS;
}

HIPR_API hipr_result hipr_stackblur_device_noalloc(hipr_context ctx, const uint8_t* src,
                                                   uint8_t* dst, int32_t width, int32_t height,
                                                   int32_t radius) {
  if (!ctx) return HIPR_ERROR_INVALID_HANDLE;
  if (!src || !dst || width <= 0 || height <= 0) return HIPR_ERROR_INVALID_ARGUMENT;
  if ((int64_t)width * height > kNoAllocMaxPixels) {
    std::fprintf(stderr, "skity_hip: hipr_stackblur_device_noalloc takes at most %lld pixels\n",
                 (long long)kNoAllocMaxPixels);
    return HIPR_ERROR_CAPACITY;
  }
  HIP_TRY(hipSetDevice(ctx->device));
  uint8_t* tmp = nullptr;
  HIP_TRY(NoAllocScratch(ctx->device, &tmp));
  hipStream_t st = ctx->stream;
  const size_t bytes = (size_t)width * height * 4;
  radius = std::min(radius, 254);
  if (radius <= 1) {  // SWStackBlur copies for radius <= 1
    if (src != dst) HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st));
  } else {
    // rows: src -> scratch, columns: scratch -> dst
    BlurTileJob rj{}, cj{};
    rj.b = BlurJob{src, tmp, width, height, radius};
    rj.line_tiles = (height + kBlurLines - 1) / kBlurLines;
    rj.out_tiles = (width + kBlurOut - 1) / kBlurOut;
    cj.b = BlurJob{tmp, dst, width, height, radius};
    cj.line_tiles = (width + kBlurLines - 1) / kBlurLines;
    cj.out_tiles = (height + kBlurOut - 1) / kBlurOut;
    const int64_t tr = (int64_t)rj.line_tiles * rj.out_tiles;
    const int64_t tc = (int64_t)cj.line_tiles * cj.out_tiles;
    hipLaunchKernelGGL(StackBlurMfmaOneKernel<0>, dim3((unsigned)std::min<int64_t>(tr, 1 << 16)),
                       dim3(64), 0, st, rj, tr);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(StackBlurMfmaOneKernel<1>, dim3((unsigned)std::min<int64_t>(tc, 1 << 16)),
                       dim3(64), 0, st, cj, tc);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipStreamSynchronize(st));
  return HIPR_SUCCESS;
}

HIPR_API hipr_result hipr_stackblur(hipr_context ctx, const uint8_t* src_host, uint8_t* dst_host,
                                    int32_t width, int32_t height, int32_t radius) {
  if (!ctx) return HIPR_ERROR_INVALID_HANDLE;
  if (!src_host || !dst_host || width <= 0 || height <= 0) return HIPR_ERROR_INVALID_ARGUMENT;
  hipStream_t st = ctx->stream;
  HIP_TRY(hipSetDevice(ctx->device));
  size_t bytes = (size_t)width * height * 4;
  radius = std::min(radius, 254);
  if (ctx->blur_tmp.Ensure(bytes * 2) != hipSuccess) return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  uint8_t* a = ctx->blur_tmp.as<uint8_t>();
  uint8_t* b = a + bytes;
  HIP_TRY(hipMemcpyAsync(a, src_host, bytes, hipMemcpyHostToDevice, st));
  if (radius > 1) {
    BlurJob jobs2[2] = {{a, b, width, height, radius}, {b, a, width, height, radius}};
    int64_t lb[1] = {0};
    hipr_result r;
    if ((r = Upload(&ctx->blur_jobs, jobs2, 2, st)) != HIPR_SUCCESS) return r;
    if ((r = Upload(&ctx->blur_lines, lb, 1, st)) != HIPR_SUCCESS) return r;
    hipLaunchKernelGGL(StackBlurLineKernel, dim3((height + 255) / 256), dim3(256), 0, st,
                       ctx->blur_jobs.as<BlurJob>(), ctx->blur_lines.as<int64_t>(), 1,
                       (int64_t)height, 0);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(StackBlurLineKernel, dim3((width + 255) / 256), dim3(256), 0, st,
                       ctx->blur_jobs.as<BlurJob>() + 1, ctx->blur_lines.as<int64_t>(), 1,
                       (int64_t)width, 1);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipMemcpyAsync(dst_host, a, bytes, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return HIPR_SUCCESS;
}

}  // extern "C"
