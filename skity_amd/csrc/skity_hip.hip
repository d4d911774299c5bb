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
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/skity_hip.h"
#include "hipr_blur.hpp"
#include "hipr_composite.hpp"
#include "hipr_flatten.hpp"
#include "hipr_sort.hpp"
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
GradDev MakeGradDev(const hipr_gradient& g, const Aff& ctm) {
  GradDev d{};
  d.type = g.type;
  d.tile = g.tile_mode;
  d.count = g.color_count;
  d.has_offsets = g.has_offsets;
  Aff ptu;
  if (g.type == HIPR_GRADIENT_LINEAR) {
    float ssx = g.points[2] - g.points[0], ssy = g.points[3] - g.points[1];
    float length = std::sqrt(ssx * ssx + ssy * ssy);
    float scale = length > 0 ? 1.0f / length : 0;
    float ux = ssx * scale, uy = ssy * scale;
    Aff rot;
    rot.sx = ux;
    rot.kx = uy;  // -sine with sine = -uy
    rot.ky = -uy;
    rot.sy = ux;
    ptu = Mul(Mul(rot, Scale(scale, scale)), Translate(-g.points[0], -g.points[1]));
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
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // device scene
  DevBuf verbs, pts, weights;
  // pipeline buffers
  DevBuf jobs, geo, cov, canvases, paints, grads;
  DevBuf edge_base, edge_cap, edges, wedges, gkeys;
  DevBuf plane_base, plane_cap, planes, map_base, map0, map1;
  DevBuf walk_list, layers, blur_jobs, blur_lines, blur_tmp;
  // pinned staging for readbacks
  JobGeo* h_geo = nullptr;
  size_t h_geo_cap = 0;
  JobCov* h_cov = nullptr;
  size_t h_cov_cap = 0;
};

struct hipr_surface_s {
  hipr_context ctx = nullptr;
  int32_t width = 0, height = 0;
  uint8_t* pixels = nullptr;
  hipr_stats stats{};
};

namespace {

template <class T>
hipr_result Upload(DevBuf* b, const T* src, size_t n, hipStream_t s) {
  if (b->Ensure(std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess)
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  if (n) HIP_TRY(hipMemcpyAsync(b->p, src, n * sizeof(T), hipMemcpyHostToDevice, s));
  return HIPR_SUCCESS;
}

hipr_result HostStage(hipr_context ctx, size_t n) {
  if (n > ctx->h_geo_cap) {
    if (ctx->h_geo) hipHostFree(ctx->h_geo);
    if (ctx->h_cov) hipHostFree(ctx->h_cov);
    size_t want = n + n / 2 + 64;
    HIP_TRY(hipHostMalloc((void**)&ctx->h_geo, want * sizeof(JobGeo)));
    HIP_TRY(hipHostMalloc((void**)&ctx->h_cov, want * sizeof(JobCov)));
    ctx->h_geo_cap = ctx->h_cov_cap = want;
  }
  return HIPR_SUCCESS;
}

hipr_result ValidateScene(const hipr_scene* s) {
  if (!s) return HIPR_ERROR_INVALID_ARGUMENT;
  if (s->n_paths < 0) return HIPR_ERROR_INVALID_ARGUMENT;
  if (s->n_paths == 0) return HIPR_SUCCESS;
  if (!s->verbs || !s->verb_offsets || !s->points || !s->point_offsets ||
      !s->weight_offsets || !s->paints)
    return HIPR_ERROR_INVALID_ARGUMENT;
  for (int p = 0; p < s->n_paths; p++) {
    int v0 = s->verb_offsets[p], v1 = s->verb_offsets[p + 1];
    if (v1 < v0) return HIPR_ERROR_INVALID_ARGUMENT;
    int need_pts = 0, need_w = 0;
    bool open = false;  // a contour is open (a move was seen)
    for (int v = v0; v < v1; v++) {
      int verb = s->verbs[v];
      if (verb < 0 || verb > HIPR_VERB_CLOSE) return HIPR_ERROR_INVALID_ARGUMENT;
      if (verb == HIPR_VERB_MOVE) {
        open = true;
      } else if (verb != HIPR_VERB_CLOSE && !open) {
        return HIPR_ERROR_INVALID_ARGUMENT;  // segment without a move
      } else if (verb == HIPR_VERB_CLOSE) {
        open = false;
      }
      need_pts += verb == HIPR_VERB_MOVE || verb == HIPR_VERB_LINE ? 1
                  : verb == HIPR_VERB_QUAD || verb == HIPR_VERB_CONIC ? 2
                  : verb == HIPR_VERB_CUBIC ? 3 : 0;
      need_w += verb == HIPR_VERB_CONIC ? 1 : 0;
      if (verb == HIPR_VERB_MOVE && v > v0 && s->verbs[v - 1] == HIPR_VERB_MOVE)
        return HIPR_ERROR_INVALID_ARGUMENT;  // the Path builder merges moves
    }
    if (s->point_offsets[p + 1] - s->point_offsets[p] != need_pts)
      return HIPR_ERROR_INVALID_ARGUMENT;
    if (s->weight_offsets[p + 1] - s->weight_offsets[p] != need_w)
      return HIPR_ERROR_INVALID_ARGUMENT;
    if (need_w && !s->weights) return HIPR_ERROR_INVALID_ARGUMENT;
    for (int i = s->point_offsets[p]; i < s->point_offsets[p + 1]; i++) {
      if (!std::isfinite(s->points[2 * i]) || !std::isfinite(s->points[2 * i + 1]))
        return HIPR_ERROR_INVALID_ARGUMENT;
    }
    const hipr_paint& pt = s->paints[p];
    if (pt.blend != HIPR_BLEND_SRC_OVER) return HIPR_ERROR_NOT_SUPPORTED;
    if (pt.style < 0 || pt.style > 3) return HIPR_ERROR_INVALID_ARGUMENT;
    if (pt.style != HIPR_STYLE_FILL) return HIPR_ERROR_NOT_SUPPORTED;
    if (pt.gradient >= 0) {
      if (!s->gradients || pt.gradient >= s->n_gradients) return HIPR_ERROR_INVALID_ARGUMENT;
      const hipr_gradient& g = s->gradients[pt.gradient];
      if (g.color_count < 2 || g.color_count > HIPR_MAX_GRADIENT_STOPS)
        return HIPR_ERROR_INVALID_ARGUMENT;
      if (g.type < HIPR_GRADIENT_LINEAR || g.type > HIPR_GRADIENT_SWEEP)
        return HIPR_ERROR_NOT_SUPPORTED;
    }
    if (pt.blur_radius > 0 && pt.blur_style != HIPR_BLUR_NORMAL)
      return HIPR_ERROR_NOT_SUPPORTED;
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

struct Plan {
  std::vector<Job> jobs;
  std::vector<Canvas> canvases;
  std::vector<PaintDev> paints;
  std::vector<GradDev> grads;
  std::vector<BlurJob> blurs;
  std::vector<int32_t> layer_of_blur;  // canvas index of each blur job
  std::vector<int32_t> walk_jobs;      // non-image jobs
  int64_t layer_bytes = 0;
};

void AddShapeJobs(const hipr_scene* s, int p, int canvas, float tx, float ty, int paint,
                  std::vector<Job>* out) {
  Job jb{};
  jb.path = p;
  jb.canvas = canvas;
  jb.fill = s->fill_types ? s->fill_types[p] : HIPR_FILL_WINDING;
  jb.geom = 0;
  jb.verb_begin = s->verb_offsets[p];
  jb.verb_end = s->verb_offsets[p + 1];
  jb.pt_begin = s->point_offsets[p];
  jb.w_begin = s->weight_offsets[p];
  jb.tx = tx;
  jb.ty = ty;
  jb.image = -1;
  jb.paint = paint;
  jb.kind = kJobFill;  // strokes are rejected by ValidateScene
  out->push_back(jb);
}

PaintDev SolidOrGradient(const hipr_scene* s, const hipr_paint& pt, const Aff& ctm,
                         std::vector<GradDev>* grads) {
  PaintDev pd{};
  pd.ga = 255;
  if (pt.gradient >= 0) {
    pd.kind = kPaintGradient;
    pd.grad = (int32_t)grads->size();
    grads->push_back(MakeGradDev(s->gradients[pt.gradient], ctm));
  } else {
    pd.kind = kPaintSolid;
    pd.color = PremulRGBAHost(pt.color);
  }
  return pd;
}

void BuildPlan(const hipr_scene* s, hipr_surface surf, Plan* pl) {
  // canvas 0 is the surface; layer canvases follow.
  std::vector<std::vector<Job>> canvas_jobs(1);
  Canvas main{};
  main.pixels = surf->pixels;
  main.width = surf->width;
  main.height = surf->height;
  pl->canvases.push_back(main);
  for (int p = 0; p < s->n_paths; p++) {
    const hipr_paint& pt = s->paints[p];
    if (pt.blur_radius > 0) {
      // HandleFilter (sw_canvas.cc:797-826): local bounds, stroke-expanded,
      // +- radius, floored/ceiled (image_filter.cc:33-41).
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
      lc.pixels = (uint8_t*)(intptr_t)pl->layer_bytes;  // offset, fixed later
      pl->layer_bytes += (int64_t)W * H * 4;
      pl->canvases.push_back(lc);
      canvas_jobs.emplace_back();
      Aff ctm = Translate(-L, -T);
      int pi = (int)pl->paints.size();
      pl->paints.push_back(SolidOrGradient(s, pt, ctm, &pl->grads));
      AddShapeJobs(s, p, layer, -L, -T, pi, &canvas_jobs[layer]);
      int radius = std::min((int)std::round(r), 254);
      BlurJob bj{};
      bj.width = W;
      bj.height = H;
      bj.radius = radius;
      pl->blurs.push_back(bj);
      pl->layer_of_blur.push_back(layer);
      // the composite draw: OnDrawImageRect with a decal nearest pixmap shader
      PaintDev ip{};
      ip.kind = kPaintImage;
      ip.layer = layer;
      ip.ga = (int32_t)(uint8_t)(255 * pt.color[3]);
      ip.L = (int)L;
      ip.T = (int)T;
      ip.R = (int)R;
      ip.B = (int)B;
      // Scale(1/W, 1/H) * inverse(Translate(L, T))
      Aff inv;
      Invert(Translate(L, T), &inv);
      Aff m = Mul(Scale(1.f / W, 1.f / H), inv);
      ip.isx = m.sx;
      ip.itx = m.tx;
      ip.isy = m.sy;
      ip.ity = m.ty;
      int ipi = (int)pl->paints.size();
      pl->paints.push_back(ip);
      Job jb{};
      jb.path = p;
      jb.canvas = 0;
      jb.kind = kJobImage;
      jb.image = layer;
      jb.paint = ipi;
      canvas_jobs[0].push_back(jb);
    } else {
      int pi = (int)pl->paints.size();
      pl->paints.push_back(SolidOrGradient(s, pt, Aff(), &pl->grads));
      AddShapeJobs(s, p, 0, 0.f, 0.f, pi, &canvas_jobs[0]);
    }
  }
  // layers first (they are composited and blurred before the canvas)
  int total_tiles = 0;
  for (size_t c = 1; c <= canvas_jobs.size(); c++) {
    size_t ci = c % canvas_jobs.size();  // 1..n-1 then 0
    Canvas& cv = pl->canvases[ci];
    cv.job_begin = (int)pl->jobs.size();
    for (auto& jb : canvas_jobs[ci]) pl->jobs.push_back(jb);
    cv.job_end = (int)pl->jobs.size();
  }
  for (auto& cv : pl->canvases) {
    cv.tiles_x = (cv.width + kTile - 1) / kTile;
    cv.tiles_y = (cv.height + kTile - 1) / kTile;
    cv.tile_begin = total_tiles;
    total_tiles += cv.tiles_x * cv.tiles_y;
  }
  for (size_t i = 0; i < pl->jobs.size(); i++)
    if (pl->jobs[i].kind != kJobImage) pl->walk_jobs.push_back((int32_t)i);
}

template <int K, int kMode>
hipr_result LaunchWalk(const WalkParams& P, int lds_width, hipStream_t st) {
  if (P.n_walk == 0) return HIPR_SUCCESS;
  size_t lds = (size_t)(2 * (lds_width + 1) + 2 * (lds_width / kTile + 2) + 5 * K * kWave) *
               sizeof(int32_t);
  hipLaunchKernelGGL((WalkKernel<K, kMode>), dim3(P.n_walk), dim3(64), lds, st, P);
  HIP_TRY(hipGetLastError());
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
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess ||
      hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
    delete ctx;
    return HIPR_ERROR_INITIALIZATION_FAILED;
  }
  *out = ctx;
  return HIPR_SUCCESS;
}

HIPR_API void hipr_context_destroy(hipr_context ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  DevBuf* bufs[] = {&ctx->verbs, &ctx->pts, &ctx->weights, &ctx->jobs, &ctx->geo,
                    &ctx->cov, &ctx->canvases, &ctx->paints, &ctx->grads, &ctx->edge_base,
                    &ctx->edge_cap, &ctx->edges, &ctx->wedges, &ctx->gkeys,
                    &ctx->plane_base, &ctx->plane_cap, &ctx->planes, &ctx->map_base,
                    &ctx->map0, &ctx->map1, &ctx->walk_list, &ctx->layers,
                    &ctx->blur_jobs, &ctx->blur_lines, &ctx->blur_tmp};
  for (DevBuf* b : bufs) b->Free();
  if (ctx->h_geo) hipHostFree(ctx->h_geo);
  if (ctx->h_cov) hipHostFree(ctx->h_cov);
  hipEventDestroy(ctx->ev0);
  hipEventDestroy(ctx->ev1);
  hipStreamDestroy(ctx->stream);
  delete ctx;
}

HIPR_API hipr_result hipr_surface_create(hipr_context ctx, int32_t width, int32_t height,
                                         hipr_surface* out) {
  if (!out) return HIPR_ERROR_INVALID_ARGUMENT;
  *out = nullptr;
  if (!ctx) return HIPR_ERROR_INVALID_HANDLE;
  if (width <= 0 || height <= 0) return HIPR_ERROR_INVALID_ARGUMENT;
  hipr_surface s = new (std::nothrow) hipr_surface_s();
  if (!s) return HIPR_ERROR_OUT_OF_HOST_MEMORY;
  s->ctx = ctx;
  s->width = width;
  s->height = height;
  HIP_TRY(hipSetDevice(ctx->device));
  if (hipMalloc(&s->pixels, (size_t)width * height * 4) != hipSuccess) {
    delete s;
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  }
  HIP_TRY(hipMemsetAsync(s->pixels, 0, (size_t)width * height * 4, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  *out = s;
  return HIPR_SUCCESS;
}

HIPR_API void hipr_surface_destroy(hipr_surface s) {
  if (!s) return;
  hipStreamSynchronize(s->ctx->stream);
  hipFree(s->pixels);
  delete s;
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

HIPR_API hipr_result hipr_surface_write_pixels(hipr_surface s, const uint8_t* src,
                                               size_t bytes) {
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

HIPR_API hipr_result hipr_surface_draw_scene(hipr_surface surf, const hipr_scene* scene,
                                             const hipr_options* options) {
  if (!surf) return HIPR_ERROR_INVALID_HANDLE;
  hipr_result vr = ValidateScene(scene);
  if (vr != HIPR_SUCCESS) return vr;
  hipr_context ctx = surf->ctx;
  hipStream_t st = ctx->stream;
  HIP_TRY(hipSetDevice(ctx->device));
  hipr_stats stats{};
  stats.paths = scene->n_paths;
  if (scene->n_paths == 0) {
    surf->stats = stats;
    return HIPR_SUCCESS;
  }
  HIP_TRY(hipEventRecord(ctx->ev0, st));
  Plan pl;
  BuildPlan(scene, surf, &pl);
  int nj = (int)pl.jobs.size();
  int nwalk = (int)pl.walk_jobs.size();
  stats.raster_paths = nwalk;
  // ---- upload scene and plan ----
  int nv = scene->verb_offsets[scene->n_paths];
  int np = scene->point_offsets[scene->n_paths];
  int nw = scene->weight_offsets[scene->n_paths];
  hipr_result r;
  if ((r = Upload(&ctx->verbs, scene->verbs, nv, st)) != HIPR_SUCCESS) return r;
  if ((r = Upload(&ctx->pts, scene->points, 2 * (size_t)np, st)) != HIPR_SUCCESS) return r;
  if ((r = Upload(&ctx->weights, scene->weights, nw, st)) != HIPR_SUCCESS) return r;
  // layer pixel storage
  if (pl.layer_bytes > 0) {
    if (ctx->layers.Ensure(pl.layer_bytes) != hipSuccess) return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
    for (size_t c = 1; c < pl.canvases.size(); c++)
      pl.canvases[c].pixels = ctx->layers.as<uint8_t>() + (intptr_t)pl.canvases[c].pixels;
    HIP_TRY(hipMemsetAsync(ctx->layers.p, 0, pl.layer_bytes, st));
  }
  if ((r = Upload(&ctx->jobs, pl.jobs.data(), nj, st)) != HIPR_SUCCESS) return r;
  if ((r = Upload(&ctx->canvases, pl.canvases.data(), pl.canvases.size(), st)) != HIPR_SUCCESS)
    return r;
  if ((r = Upload(&ctx->paints, pl.paints.data(), pl.paints.size(), st)) != HIPR_SUCCESS)
    return r;
  if ((r = Upload(&ctx->grads, pl.grads.data(), pl.grads.size(), st)) != HIPR_SUCCESS)
    return r;
  if ((r = Upload(&ctx->walk_list, pl.walk_jobs.data(), nwalk, st)) != HIPR_SUCCESS) return r;
  if (ctx->geo.Ensure(sizeof(JobGeo) * std::max(nj, 1)) != hipSuccess ||
      ctx->cov.Ensure(sizeof(JobCov) * std::max(nj, 1)) != hipSuccess ||
      ctx->edge_base.Ensure(sizeof(int64_t) * std::max(nj, 1)) != hipSuccess ||
      ctx->edge_cap.Ensure(sizeof(int32_t) * std::max(nj, 1)) != hipSuccess ||
      ctx->plane_base.Ensure(sizeof(int64_t) * std::max(nj, 1)) != hipSuccess ||
      ctx->plane_cap.Ensure(sizeof(int32_t) * std::max(nj, 1)) != hipSuccess ||
      ctx->map_base.Ensure(sizeof(int64_t) * std::max(nj, 1)) != hipSuccess)
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  if ((r = HostStage(ctx, nj)) != HIPR_SUCCESS) return r;
  GeomSrc gs{ctx->verbs.as<uint8_t>(), ctx->pts.as<float>(), ctx->weights.as<float>()};
  GeomSrc go = gs;  // stroke outlines: not produced yet

  // ---- stage (a): flatten, count pass ----
  if (nj > 0) {
    hipLaunchKernelGGL((FlattenKernel<0>), dim3(nj), dim3(64), 0, st, ctx->jobs.as<Job>(), nj,
                       gs, go, ctx->edge_base.as<int64_t>(), ctx->edge_cap.as<int32_t>(),
                       (EdgeRec*)nullptr, ctx->geo.as<JobGeo>());
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(ctx->h_geo, ctx->geo.p, sizeof(JobGeo) * nj, hipMemcpyDeviceToHost,
                           st));
    HIP_TRY(hipStreamSynchronize(st));
  }
  std::vector<int64_t> ebase(nj + 1, 0);
  std::vector<int32_t> ecap(nj, 0);
  int lds_width = 256;
  for (int j = 0; j < nj; j++) {
    ecap[j] = ctx->h_geo[j].edges;
    ebase[j + 1] = ebase[j] + ecap[j];
    if (pl.jobs[j].kind != kJobImage) {
      PixBounds pb = JobPixBounds(ctx->h_geo[j]);
      if (!pb.empty) lds_width = std::max(lds_width, pb.r - pb.l);
    }
  }
  lds_width = (lds_width + 255) & ~255;
  int64_t total_edges = ebase[nj];
  stats.edges = total_edges;
  if (ctx->edges.Ensure(sizeof(EdgeRec) * std::max<int64_t>(total_edges, 1)) != hipSuccess ||
      ctx->wedges.Ensure(sizeof(EdgeRec) * std::max<int64_t>(total_edges, 1)) != hipSuccess ||
      ctx->gkeys.Ensure(sizeof(uint64_t) * 2 * std::max<int64_t>(total_edges, 1)) != hipSuccess)
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  HIP_TRY(hipMemcpyAsync(ctx->edge_base.p, ebase.data(), sizeof(int64_t) * nj,
                         hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(ctx->edge_cap.p, ecap.data(), sizeof(int32_t) * nj,
                         hipMemcpyHostToDevice, st));
  if (nj > 0) {
    hipLaunchKernelGGL((FlattenKernel<1>), dim3(nj), dim3(64), 0, st, ctx->jobs.as<Job>(), nj,
                       gs, go, ctx->edge_base.as<int64_t>(), ctx->edge_cap.as<int32_t>(),
                       ctx->edges.as<EdgeRec>(), ctx->geo.as<JobGeo>());
    HIP_TRY(hipGetLastError());
  }
  // ---- sort + gather ----
  auto sort_gather = [&]() -> hipr_result {
    if (nwalk == 0) return HIPR_SUCCESS;
    hipLaunchKernelGGL(SortGatherKernel, dim3(nwalk), dim3(kSortThreads), 0, st,
                       ctx->walk_list.as<int32_t>(), nwalk, ctx->geo.as<JobGeo>(),
                       ctx->edge_base.as<int64_t>(), ctx->edges.as<EdgeRec>(),
                       ctx->wedges.as<EdgeRec>(), ctx->gkeys.as<uint64_t>());
    HIP_TRY(hipGetLastError());
    return HIPR_SUCCESS;
  };
  if ((r = sort_gather()) != HIPR_SUCCESS) return r;
  // ---- stage (b): walk, count pass ----
  WalkParams wp{};
  wp.jobs = ctx->jobs.as<Job>();
  wp.walk_list = ctx->walk_list.as<int32_t>();
  wp.n_walk = nwalk;
  wp.geo = ctx->geo.as<JobGeo>();
  wp.edge_base = ctx->edge_base.as<int64_t>();
  wp.wedges = ctx->wedges.as<EdgeRec>();
  wp.canvases = ctx->canvases.as<Canvas>();
  wp.cov = ctx->cov.as<JobCov>();
  wp.lds_width = lds_width;
  if ((r = LaunchWalk<4, 0>(wp, lds_width, st)) != HIPR_SUCCESS) return r;
  if (nj > 0) {
    HIP_TRY(hipMemcpyAsync(ctx->h_cov, ctx->cov.p, sizeof(JobCov) * nj, hipMemcpyDeviceToHost,
                           st));
    HIP_TRY(hipStreamSynchronize(st));
  }
  std::vector<int64_t> pbase(nj + 1, 0), mbase(nj + 1, 0);
  std::vector<int32_t> pcap(nj, 0);
  int64_t max_active = 0;
  for (int j = 0; j < nj; j++) {
    const JobCov& c = ctx->h_cov[j];
    bool walked = pl.jobs[j].kind != kJobImage;
    if (walked && (c.flags & (2 | 16 | 32))) {
      std::fprintf(stderr, "skity_hip: job %d exceeds a device limit (flags %d)\n", j,
                   c.flags);
      return HIPR_ERROR_CAPACITY;
    }
    pcap[j] = walked ? c.planes : 0;
    pbase[j + 1] = pbase[j] + pcap[j];
    int64_t msz = 0;
    if (walked) {
      const Canvas& cv = pl.canvases[pl.jobs[j].canvas];
      PixBounds pb{c.left, c.top, c.right, c.bottom, !(c.left < c.right && c.top < c.bottom)};
      TileRange tr = JobTiles(pb, cv.width, cv.height);
      msz = (int64_t)tr.ntc * tr.ntr;
    }
    mbase[j + 1] = mbase[j] + msz;
    stats.tiles += walked ? c.tiles : 0;
    stats.tile_planes += pcap[j];
    max_active = std::max<int64_t>(max_active, c.max_active);
  }
  stats.max_active = max_active;
  int64_t total_planes = pbase[nj], total_map = mbase[nj];
  if (ctx->planes.Ensure((size_t)kTilePlaneBytes * std::max<int64_t>(total_planes, 1)) !=
          hipSuccess ||
      ctx->map0.Ensure(sizeof(int32_t) * std::max<int64_t>(total_map, 1)) != hipSuccess ||
      ctx->map1.Ensure(sizeof(int32_t) * std::max<int64_t>(total_map, 1)) != hipSuccess)
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  HIP_TRY(hipMemsetAsync(ctx->map0.p, 0xFF, sizeof(int32_t) * std::max<int64_t>(total_map, 1), st));
  HIP_TRY(hipMemsetAsync(ctx->map1.p, 0xFF, sizeof(int32_t) * std::max<int64_t>(total_map, 1), st));
  HIP_TRY(hipMemcpyAsync(ctx->plane_base.p, pbase.data(), sizeof(int64_t) * nj,
                         hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(ctx->plane_cap.p, pcap.data(), sizeof(int32_t) * nj,
                         hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(ctx->map_base.p, mbase.data(), sizeof(int64_t) * nj,
                         hipMemcpyHostToDevice, st));
  // ---- stage (b): walk, write pass (on a fresh copy of the sorted edges) ----
  if ((r = sort_gather()) != HIPR_SUCCESS) return r;
  wp.plane_base = ctx->plane_base.as<int64_t>();
  wp.plane_cap = ctx->plane_cap.as<int32_t>();
  wp.planes = ctx->planes.as<uint8_t>();
  wp.map_base = ctx->map_base.as<int64_t>();
  wp.map0 = ctx->map0.as<int32_t>();
  wp.map1 = ctx->map1.as<int32_t>();
  if ((r = LaunchWalk<4, 1>(wp, lds_width, st)) != HIPR_SUCCESS) return r;
  // ---- composite: layers, blur, canvas ----
  CompositeParams cp{};
  cp.canvases = ctx->canvases.as<Canvas>();
  cp.n_canvases = (int)pl.canvases.size();
  cp.jobs = ctx->jobs.as<Job>();
  cp.cov = ctx->cov.as<JobCov>();
  cp.map_base = ctx->map_base.as<int64_t>();
  cp.map0 = ctx->map0.as<int32_t>();
  cp.map1 = ctx->map1.as<int32_t>();
  cp.plane_base = ctx->plane_base.as<int64_t>();
  cp.planes = ctx->planes.as<uint8_t>();
  cp.paints = ctx->paints.as<PaintDev>();
  cp.grads = ctx->grads.as<GradDev>();
  int main_tiles = pl.canvases[0].tiles_x * pl.canvases[0].tiles_y;
  int total_tiles = 0;
  for (auto& cv : pl.canvases) total_tiles += cv.tiles_x * cv.tiles_y;
  if (pl.canvases.size() > 1) {
    // layer tiles come after the main canvas's tiles in tile numbering;
    // composite them first by launching over the layer range only.
    CompositeParams lp = cp;
    lp.canvases = ctx->canvases.as<Canvas>() + 1;
    lp.n_canvases = cp.n_canvases - 1;
    // re-base: the layer canvases' tile_begin starts at main_tiles
    std::vector<Canvas> lc(pl.canvases.begin() + 1, pl.canvases.end());
    for (auto& c : lc) c.tile_begin -= main_tiles;
    DevBuf& tmpc = ctx->blur_jobs;  // reuse as a small scratch for layer canvases
    (void)tmpc;
    static thread_local DevBuf layer_canvases;
    if ((r = Upload(&layer_canvases, lc.data(), lc.size(), st)) != HIPR_SUCCESS) return r;
    lp.canvases = layer_canvases.as<Canvas>();
    lp.total_tiles = total_tiles - main_tiles;
    hipLaunchKernelGGL(CompositeKernel, dim3(lp.total_tiles), dim3(64), 0, st, lp);
    HIP_TRY(hipGetLastError());
    // blur each layer: rows into tmp, then columns back into the layer
    std::vector<BlurJob> rows, cols;
    std::vector<int64_t> rb{0}, cb{0};
    int64_t tmp_bytes = 0;
    for (size_t b = 0; b < pl.blurs.size(); b++) tmp_bytes += (int64_t)pl.blurs[b].width * pl.blurs[b].height * 4;
    if (ctx->blur_tmp.Ensure(std::max<int64_t>(tmp_bytes, 1)) != hipSuccess)
      return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
    int64_t off = 0;
    std::vector<std::pair<uint8_t*, int64_t>> copies;
    for (size_t b = 0; b < pl.blurs.size(); b++) {
      BlurJob bj = pl.blurs[b];
      uint8_t* layer_px = pl.canvases[pl.layer_of_blur[b]].pixels;
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
    if (!rows.empty()) {
      static thread_local DevBuf d_rows, d_cols, d_rb, d_cb;
      if ((r = Upload(&d_rows, rows.data(), rows.size(), st)) != HIPR_SUCCESS) return r;
      if ((r = Upload(&d_cols, cols.data(), cols.size(), st)) != HIPR_SUCCESS) return r;
      if ((r = Upload(&d_rb, rb.data(), rb.size(), st)) != HIPR_SUCCESS) return r;
      if ((r = Upload(&d_cb, cb.data(), cb.size(), st)) != HIPR_SUCCESS) return r;
      int64_t nl = rb.back();
      hipLaunchKernelGGL(StackBlurLineKernel, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0,
                         st, d_rows.as<BlurJob>(), d_rb.as<int64_t>(), (int)rows.size(), nl, 0);
      HIP_TRY(hipGetLastError());
      int64_t nc = cb.back();
      hipLaunchKernelGGL(StackBlurLineKernel, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0,
                         st, d_cols.as<BlurJob>(), d_cb.as<int64_t>(), (int)cols.size(), nc, 1);
      HIP_TRY(hipGetLastError());
    }
  }
  {
    CompositeParams mp = cp;
    mp.n_canvases = 1;
    mp.total_tiles = main_tiles;
    hipLaunchKernelGGL(CompositeKernel, dim3(main_tiles), dim3(64), 0, st, mp);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipEventRecord(ctx->ev1, st));
  HIP_TRY(hipEventSynchronize(ctx->ev1));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  stats.ms_total = ms;
  surf->stats = stats;
  return HIPR_SUCCESS;
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
  std::vector<Job> jobs;
  for (int p = 0; p < n; p++) AddShapeJobs(scene, p, 0, 0.f, 0.f, 0, &jobs);
  int nv = n ? scene->verb_offsets[n] : 0, np = n ? scene->point_offsets[n] : 0;
  int nw = n ? scene->weight_offsets[n] : 0;
  hipr_result r;
  if ((r = Upload(&ctx->verbs, scene->verbs, nv, st)) != HIPR_SUCCESS) return r;
  if ((r = Upload(&ctx->pts, scene->points, 2 * (size_t)np, st)) != HIPR_SUCCESS) return r;
  if ((r = Upload(&ctx->weights, scene->weights, nw, st)) != HIPR_SUCCESS) return r;
  if ((r = Upload(&ctx->jobs, jobs.data(), jobs.size(), st)) != HIPR_SUCCESS) return r;
  if (ctx->geo.Ensure(sizeof(JobGeo) * std::max(n, 1)) != hipSuccess ||
      ctx->edge_base.Ensure(sizeof(int64_t) * std::max(n, 1)) != hipSuccess ||
      ctx->edge_cap.Ensure(sizeof(int32_t) * std::max(n, 1)) != hipSuccess)
    return HIPR_ERROR_OUT_OF_DEVICE_MEMORY;
  if ((r = HostStage(ctx, n)) != HIPR_SUCCESS) return r;
  GeomSrc gs{ctx->verbs.as<uint8_t>(), ctx->pts.as<float>(), ctx->weights.as<float>()};
  if (n > 0) {
    hipLaunchKernelGGL((FlattenKernel<0>), dim3(n), dim3(64), 0, st, ctx->jobs.as<Job>(), n, gs,
                       gs, ctx->edge_base.as<int64_t>(), ctx->edge_cap.as<int32_t>(),
                       (EdgeRec*)nullptr, ctx->geo.as<JobGeo>());
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
    hipLaunchKernelGGL((FlattenKernel<1>), dim3(n), dim3(64), 0, st, ctx->jobs.as<Job>(), n, gs,
                       gs, ctx->edge_base.as<int64_t>(), ctx->edge_cap.as<int32_t>(),
                       ctx->edges.as<EdgeRec>(), ctx->geo.as<JobGeo>());
    HIP_TRY(hipGetLastError());
  }
  if (total > 0)
    HIP_TRY(hipMemcpyAsync(edges, ctx->edges.p, sizeof(EdgeRec) * total, hipMemcpyDeviceToHost, st));
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
    BlurJob h{a, b, width, height, radius}, v{b, a, width, height, radius};
    hipr_result r;
    int64_t lb[2] = {0, 0};
    if ((r = Upload(&ctx->blur_jobs, &h, 1, st)) != HIPR_SUCCESS) return r;
    if ((r = Upload(&ctx->blur_lines, lb, 1, st)) != HIPR_SUCCESS) return r;
    hipLaunchKernelGGL(StackBlurLineKernel, dim3((height + 255) / 256), dim3(256), 0, st,
                       ctx->blur_jobs.as<BlurJob>(), ctx->blur_lines.as<int64_t>(), 1,
                       (int64_t)height, 0);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    if ((r = Upload(&ctx->blur_jobs, &v, 1, st)) != HIPR_SUCCESS) return r;
    hipLaunchKernelGGL(StackBlurLineKernel, dim3((width + 255) / 256), dim3(256), 0, st,
                       ctx->blur_jobs.as<BlurJob>(), ctx->blur_lines.as<int64_t>(), 1,
                       (int64_t)width, 1);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipMemcpyAsync(dst_host, a, bytes, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return HIPR_SUCCESS;
}

}  // extern "C"
