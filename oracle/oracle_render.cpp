// oracle_render.cpp -- TEST INFRASTRUCTURE ONLY (parity checker).
//
// CPU restatement of skity's software brushes, blending, gradients, StackBlur
// and the mask-filter draw (SWCanvas::OnDrawPath / DoBrush / GenerateBrush /
// HandleFilter).  Exposed through the C functions in oracle.h, which only
// tests/, smoke() and the CPU-baseline leg of bench.py call.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "oracle.h"
#include "oracle_raster.hpp"

namespace orc {

// Stroke::StrokePath restatement lives in oracle_stroke.cpp.
PathData StrokeOutline(const PathData& quad_path, const hipr_paint& paint);

namespace {

// ---- colour helpers: src/graphic/color.cc:44-59, color_priv.hpp:20-84 -----
inline uint32_t SetARGB(uint32_t a, uint32_t r, uint32_t g, uint32_t b) {
  return (a << 24) | (r << 16) | (g << 8) | b;
}
inline uint32_t GetA(uint32_t c) { return (c >> 24) & 0xFF; }
inline uint32_t GetR(uint32_t c) { return (c >> 16) & 0xFF; }
inline uint32_t GetG(uint32_t c) { return (c >> 8) & 0xFF; }
inline uint32_t GetB(uint32_t c) { return c & 0xFF; }

inline uint8_t UnitToByte(float v) {
  float s = v * 255;
  s = std::min(std::max(s, 0.f), 255.f);
  return (uint8_t)s;
}
// Color4fToColor: truncating (color.cc:53-59).
inline uint32_t Color4fToColor(const float c[4]) {
  return SetARGB(UnitToByte(c[3]), UnitToByte(c[0]), UnitToByte(c[1]),
                 UnitToByte(c[2]));
}
inline uint32_t MulDiv255Round(uint32_t a, uint32_t b) {
  uint32_t prod = a * b + 128;
  return (prod + (prod >> 8)) >> 8;
}
// ColorToPMColor (color_priv.cc:103-106).
inline uint32_t ColorToPMColor(uint32_t c) {
  uint32_t a = GetA(c), r = GetR(c), g = GetG(c), b = GetB(c);
  if (a != 255) {
    r = MulDiv255Round(r, a);
    g = MulDiv255Round(g, a);
    b = MulDiv255Round(b, a);
  }
  return SetARGB(a, r, g, b);
}
inline uint32_t AlphaMulQ(uint32_t c, uint32_t scale) {
  const uint32_t mask = 0xFF00FF;
  uint32_t rb = ((c & mask) * scale) >> 8;
  uint32_t ag = ((c >> 8) & mask) * scale;
  return (rb & mask) | (ag & ~mask);
}
inline uint32_t PMSrcOver(uint32_t src, uint32_t dst) {
  return src + AlphaMulQ(dst, 256 - GetA(src));
}

// ---- premultiplied RGBA8 target: SWRenderTarget::BlendPixel with FastBlend
// and PorterDuffBlend for kSrcOver (sw_render_target.cc:12-141;
// blend_mode.cc:125-191); Bitmap::Get/SetPixel for kRGBA (bitmap.cc:25-85).
struct Target {
  uint8_t* px;
  int w, h;
  uint32_t Get(int x, int y) const {
    const uint8_t* p = px + ((size_t)y * w + x) * 4;
    return SetARGB(p[3], p[0], p[1], p[2]);
  }
  void Set(int x, int y, uint32_t c) {
    uint8_t* p = px + ((size_t)y * w + x) * 4;
    p[0] = (uint8_t)GetR(c);
    p[1] = (uint8_t)GetG(c);
    p[2] = (uint8_t)GetB(c);
    p[3] = (uint8_t)GetA(c);
  }
  void Blend(int x, int y, uint32_t src) {
    if ((unsigned)x >= (unsigned)w || (unsigned)y >= (unsigned)h) return;
    uint32_t a = GetA(src);
    if (a == 0) return;
    if (a == 255) {
      Set(x, y, src);
      return;
    }
    Set(x, y, PMSrcOver(src, Get(x, y)));
  }
};

// ---- affine matrices as skity's 4x4 Matrix restricted to the plane -------
// Products follow glm's column order (sum of columns left to right), which
// for affine operands reduces to the expressions below.
struct Aff {
  float sx = 1, kx = 0, tx = 0, ky = 0, sy = 1, ty = 0;
  bool IsIdentity() const {
    return sx == 1 && kx == 0 && tx == 0 && ky == 0 && sy == 1 && ty == 0;
  }
  bool OnlyScaleTranslate() const { return kx == 0 && ky == 0; }
};
Aff Mul(const Aff& a, const Aff& b) {
  if (a.IsIdentity()) return b;
  if (b.IsIdentity()) return a;
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
// Matrix::Invert (src/geometry/matrix.cc:134-142,242-279).  The general case
// uses glm::inverse in the reference; glm is not available, so the inverse is
// computed in double and rounded (within the 1/255 gradient tolerance).
bool Invert(const Aff& m, Aff* out) {
  if (m.IsIdentity()) {
    *out = Aff();
    return true;
  }
  if (m.OnlyScaleTranslate()) {
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
inline void MapPoint(const Aff& m, float x, float y, float* ox, float* oy) {
  *ox = x * m.sx + y * m.kx + m.tx;
  *oy = x * m.ky + y * m.sy + m.ty;
}

// ---- gradients: src/render/sw/sw_span_brush.cc:21-32,164-219,239-383 -----
float RemapFloatTile(float t, int tile) {
  if (tile == HIPR_TILE_CLAMP) {
    t = std::clamp(t, 0.0f, 1.0f);
  } else if (tile == HIPR_TILE_REPEAT) {
    t = t - std::floor(t);
  } else if (tile == HIPR_TILE_MIRROR) {
    float t1 = t - 1;
    float t2 = t1 - 2 * std::floor(t1 * 0.5) - 1;
    t = std::abs(t2);
  }
  return t;
}

struct GradientBrush {
  hipr_gradient g;
  std::vector<float> offsets;  // after EnsureColorOffsetsNonDescending
  Aff points_to_unit;

  void LerpColor(float current, float out[4]) const {
    if (std::abs(current) <= 1.f / 4096) {
      current = 0.0f;
    } else if (std::abs(current - 1.0f) <= 1.f / 4096) {
      current = 1.0f;
    }
    if (g.tile_mode == HIPR_TILE_DECAL && (current < 0.0 || current >= 1.0)) {
      out[0] = out[1] = out[2] = out[3] = 0;
      return;
    }
    current = RemapFloatTile(current, g.tile_mode);
    int color_count = g.color_count;
    int stop_count = g.has_offsets ? color_count : 0;
    int start_index = 0, end_index = 1;
    float step = 1.f / (color_count - 1);
    if (stop_count > 0 && current <= offsets[0]) {
      std::memcpy(out, g.colors[0], sizeof(float) * 4);
      return;
    }
    int i = 0;
    float start = 0.f, end = 0.f;
    for (i = 0; i < color_count - 1; i++) {
      if (stop_count > 0) {
        start = offsets[i];
        end = offsets[i + 1];
      } else {
        start = step * i;
        end = step * (i + 1);
      }
      if (current >= start && current <= end) {
        start_index = i;
        end_index = i + 1;
        break;
      }
    }
    if (i == color_count - 1 && color_count > 0) {
      std::memcpy(out, g.colors[color_count - 1], sizeof(float) * 4);
      return;
    }
    float total = end - start;
    float value = current - start;
    float mix = 0.5f;
    if (total > 0) mix = value / total;
    for (int c = 0; c < 4; c++)
      out[c] = g.colors[start_index][c] * (1 - mix) + g.colors[end_index][c] * mix;
  }

  uint32_t Color(int x, int y) const {
    float sx = x + 0.5f, sy = y + 0.5f, mx, my;
    MapPoint(points_to_unit, sx, sy, &mx, &my);
    float t;
    if (g.type == HIPR_GRADIENT_LINEAR) {
      t = mx;
    } else if (g.type == HIPR_GRADIENT_RADIAL) {
      t = std::sqrt(mx * mx + my * my);
    } else {
      float angle = std::atan2(-my, -mx);
      float bias = g.radius[0], scale = g.radius[1];
      constexpr static float k1Over2Pi = 0.1591549430918;
      t = (angle * k1Over2Pi + 0.5 + bias) * scale;
    }
    float c[4];
    LerpColor(t, c);
    return ColorToPMColor(Color4fToColor(c));
  }
};

// PointsToUnit (sw_span_brush.cc:164-198) composed with device_to_local
// (sw_canvas.cc:736-751): device_to_local = inverse(local) * inverse(CTM).
Aff GradientPointsToUnit(const hipr_gradient& g, const Aff& ctm) {
  Aff ptu;
  if (g.type == HIPR_GRADIENT_LINEAR) {
    float ssx = g.points[2] - g.points[0], ssy = g.points[3] - g.points[1];
    float length = std::sqrt(ssx * ssx + ssy * ssy);
    float scale = length > 0 ? 1.0f / length : 0;
    float ux = ssx * scale, uy = ssy * scale;
    float sine = -uy, cosine = ux;
    Aff rot;
    rot.sx = cosine;
    rot.kx = -sine;
    rot.ky = sine;
    rot.sy = cosine;
    ptu = Mul(Mul(rot, Scale(scale, scale)), Translate(-g.points[0], -g.points[1]));
  } else if (g.type == HIPR_GRADIENT_RADIAL) {
    float r = g.radius[0];
    float scale = r > 0 ? 1.0f / r : 0;
    ptu = Mul(Scale(scale, scale), Translate(-g.points[0], -g.points[1]));
  } else {
    ptu = Translate(-g.points[0], -g.points[1]);
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
  return Mul(ptu, Mul(d2l, l2l));
}

void PrepareGradient(const hipr_gradient& g, const Aff& ctm, GradientBrush* b) {
  b->g = g;
  b->offsets.assign(g.offsets, g.offsets + g.color_count);
  if (g.has_offsets) {
    for (int i = 0; i < g.color_count - 1; i++)
      if (b->offsets[i] > b->offsets[i + 1]) b->offsets[i + 1] = b->offsets[i];
  }
  b->points_to_unit = GradientPointsToUnit(g, ctm);
}

// A brush: solid colour, gradient, or the premultiplied image of a blurred
// draw sampled nearest through a decal pixmap shader (PixmapBrush,
// sw_span_brush.cc:555-579; BitmapSampler, bitmap_sampler.cc:26-40,85-108).
struct Brush {
  int kind = 0;  // 0 solid, 1 gradient, 2 image
  uint32_t solid = 0;
  GradientBrush grad;
  const uint8_t* img = nullptr;
  int img_w = 0, img_h = 0;
  Aff img_ptu;
  uint8_t global_alpha = 255;

  uint32_t Color(int x, int y) const {
    if (kind == 0) return solid;
    if (kind == 1) return grad.Color(x, y);
    float u, v;
    MapPoint(img_ptu, x + 0.5f, y + 0.5f, &u, &v);
    if (u < 0.0 || u >= 1.0 || v < 0.0 || v >= 1.0) return 0;
    float fx = u * (uint32_t)img_w, fy = v * (uint32_t)img_h;
    uint32_t ix = std::min<uint32_t>(std::max<uint32_t>((uint32_t)fx, 0), img_w - 1);
    uint32_t iy = std::min<uint32_t>(std::max<uint32_t>((uint32_t)fy, 0), img_h - 1);
    const uint8_t* p = img + ((size_t)iy * img_w + ix) * 4;
    // Color4fFromColor then Color4fToColor (may lose one step per channel).
    float c[4] = {p[0] / 255.f, p[1] / 255.f, p[2] / 255.f, p[3] / 255.f};
    return Color4fToColor(c);
  }
};

// SWSpanBrush::Brush / BrushH (sw_span_brush.cc:66-138).
void ApplySpans(const std::vector<Span>& spans, const Brush& br, Target* t) {
  for (const Span& s : spans) {
    int x = s.x, y = s.y, len = s.len;
    if (y < 0 || y >= t->h) continue;
    if (x >= t->w || x + len < 0) continue;
    if (x < 0) {
      len += x;
      x = 0;
    }
    if (x + len >= t->w) len = t->w - x;
    if (len <= 0) continue;
    uint32_t alpha = (uint8_t)(s.cover & br.global_alpha);
    for (int l = 0; l < len; l++) {
      uint32_t c = br.Color(x + l, y);
      if (alpha != 255) c = AlphaMulQ(c, alpha);
      t->Blend(x + l, y, c);
    }
  }
}

// ---- StackBlur (src/render/sw/sw_stack_blur.cc:15-338) as the triangle
// convolution it computes: horizontal then vertical, clamp-to-edge, weights
// (r + 1 - |i|), normalised by (sum * mul[r]) >> shr[r].
void StackBlurTables(int r, uint32_t* mul, uint32_t* shr);

}  // namespace

void StackBlurTablesPublic(int r, uint32_t* mul, uint32_t* shr) {
  StackBlurTables(r, mul, shr);
}

namespace {

// The published StackBlur normalisation: shr is the smallest shift with
// 2^shr / (r+1)^2 > 256 and mul = ceil(2^shr / (r+1)^2) (the tables at
// sw_stack_blur.cc:286-338 list exactly these values for r in [0, 254]).
void StackBlurTables(int r, uint32_t* mul, uint32_t* shr) {
  uint64_t d = (uint64_t)(r + 1) * (r + 1);
  int s = 0;
  while (!(((uint64_t)1 << s) > 256 * d)) s++;
  *shr = (uint32_t)s;
  *mul = (uint32_t)((((uint64_t)1 << s) + d - 1) / d);
}

void StackBlurPass(const uint8_t* src, uint8_t* dst, int w, int h, int r,
                   bool horizontal) {
  uint32_t mul, shr;
  StackBlurTables(r, &mul, &shr);
  int n = horizontal ? w : h;
  int lines = horizontal ? h : w;
  for (int line = 0; line < lines; line++) {
    for (int i = 0; i < n; i++) {
      for (int c = 0; c < 4; c++) {
        uint64_t sum = 0;
        for (int k = -r; k <= r; k++) {
          int j = std::min(std::max(i + k, 0), n - 1);
          size_t idx = horizontal ? ((size_t)line * w + j) : ((size_t)j * w + line);
          sum += (uint64_t)(r + 1 - std::abs(k)) * src[idx * 4 + c];
        }
        size_t o = horizontal ? ((size_t)line * w + i) : ((size_t)i * w + line);
        dst[o * 4 + c] = (uint8_t)((sum * mul) >> shr);
      }
    }
  }
}

}  // namespace

void StackBlur(const uint8_t* src, uint8_t* dst, int w, int h, int radius) {
  radius = std::min(radius, 254);
  if (radius <= 1) {
    std::memcpy(dst, src, (size_t)w * h * 4);
    return;
  }
  std::vector<uint8_t> tmp((size_t)w * h * 4);
  StackBlurPass(src, tmp.data(), w, h, radius, true);
  StackBlurPass(tmp.data(), dst, w, h, radius, false);
}

namespace {

// GenerateBrush (sw_canvas.cc:727-795): gradient shader, else the stroke or
// fill colour.
Brush MakeBrush(const hipr_scene* s, const hipr_paint& paint, const Aff& ctm,
                bool stroke) {
  Brush br;
  if (paint.gradient >= 0 && s->gradients && paint.gradient < s->n_gradients) {
    br.kind = 1;
    PrepareGradient(s->gradients[paint.gradient], ctm, &br.grad);
  } else {
    br.kind = 0;
    br.solid = ColorToPMColor(Color4fToColor(stroke ? paint.stroke_color : paint.color));
  }
  return br;
}

// One fill or stroke rasterisation plus brush (sw_canvas.cc:374-407).
void DrawShape(const hipr_scene* s, const PathData& path, const hipr_paint& paint,
               bool stroke, const Aff& ctm, Target* t) {
  Brush br = MakeBrush(s, paint, ctm, stroke);
  std::vector<Span> spans;
  float bounds[4];
  if (!stroke) {
    RastePath(path, &spans, bounds, nullptr, ctm.tx, ctm.ty);
  } else {
    PathData quad = QuadPath(path);
    PathData outline = StrokeOutline(quad, paint);
    RastePath(outline, &spans, bounds, nullptr, ctm.tx, ctm.ty);
  }
  ApplySpans(spans, br, t);
}

// The canvas matrix is a translation here (identity on the main canvas,
// -filter_bounds.LT on a mask-filter layer).
void DrawPathNoFilter(const hipr_scene* s, const PathData& path,
                      const hipr_paint& paint, const Aff& ctm, Target* t) {
  bool need_fill = paint.style != HIPR_STYLE_STROKE;
  bool need_stroke = paint.style != HIPR_STYLE_FILL;
  if (paint.style == HIPR_STYLE_STROKE_THEN_FILL) {
    if (need_stroke) DrawShape(s, path, paint, true, ctm, t);
    if (need_fill) DrawShape(s, path, paint, false, ctm, t);
  } else {
    if (need_fill) DrawShape(s, path, paint, false, ctm, t);
    if (need_stroke) DrawShape(s, path, paint, true, ctm, t);
  }
}

}  // namespace

// SWCanvas::HandleFilter + MaskFilterOnFilter(kNormal) + OnDrawImageRect
// (sw_canvas.cc:135-144,641-677,797-826; mask_filter.cc:51-58;
// image_filter.cc:33-41,184-194).  Returns the filter bounds and leaves the
// blurred premultiplied image in *blurred (exposed for stage parity).
void BlurredLayer(const hipr_scene* s, int p, std::vector<uint8_t>* blurred,
                  int fb[4]) {
  const hipr_paint& paint = s->paints[p];
  PathData path = ScenePath(s, p);
  float b[4];
  PathBounds(path, b);
  if (paint.style != HIPR_STYLE_FILL) {
    float w = paint.stroke_width;
    b[0] = std::floor(b[0] - w);
    b[1] = std::floor(b[1] - w);
    b[2] = std::floor(b[2] + w);
    b[3] = std::floor(b[3] + w);
  }
  float r = paint.blur_radius;
  float l = std::floor(b[0] - r), tp = std::floor(b[1] - r);
  float rt = std::ceil(b[2] + r), bt = std::ceil(b[3] + r);
  fb[0] = (int)l;
  fb[1] = (int)tp;
  fb[2] = (int)rt;
  fb[3] = (int)bt;
  int W = (int)(rt - l), H = (int)(bt - tp);
  std::vector<uint8_t> layer((size_t)W * H * 4, 0);
  Target lt{layer.data(), W, H};
  hipr_paint work = paint;
  work.blur_radius = 0;
  // The temporary canvas is translated by -filter_bounds.LT; its points go
  // through the canvas matrix as x + (-L) after the quad conversion
  // (sw_raster.cc:739-740; matrix.cc:462-465).
  DrawPathNoFilter(s, path, work, Translate(-l, -tp), &lt);
  blurred->assign((size_t)W * H * 4, 0);
  int radius = (int)std::round(r);
  StackBlur(layer.data(), blurred->data(), W, H, radius);
}

namespace {

void DrawPathBlurred(const hipr_scene* s, int p, Target* t) {
  const hipr_paint& paint = s->paints[p];
  std::vector<uint8_t> blurred;
  int fb[4];
  BlurredLayer(s, p, &blurred, fb);
  int W = fb[2] - fb[0], H = fb[3] - fb[1];
  if (W <= 0 || H <= 0) return;
  // OnDrawImageRect: rect path of the filter bounds filled with a decal
  // nearest pixmap shader; paint alpha applied again as global alpha.
  PathData rect;
  float L = (float)fb[0], T = (float)fb[1], R = (float)fb[2], B = (float)fb[3];
  rect.MoveTo(Pt{L, T});
  rect.LineTo(Pt{R, T});
  rect.LineTo(Pt{R, B});
  rect.LineTo(Pt{L, B});
  rect.Close();
  std::vector<Span> spans;
  float bounds[4];
  RastePath(rect, &spans, bounds, nullptr);
  Brush br;
  br.kind = 2;
  br.img = blurred.data();
  br.img_w = W;
  br.img_h = H;
  Aff local = Translate(L, T);
  Aff inv;
  Invert(local, &inv);
  br.img_ptu = Mul(Scale(1.f / W, 1.f / H), inv);
  // Paint::GetAlphaF: the stroke colour's alpha for kStroke_Style, else the
  // fill colour's (paint.cc:96-98).
  float alpha_f = paint.style == HIPR_STYLE_STROKE ? paint.stroke_color[3] : paint.color[3];
  br.global_alpha = (uint8_t)(255 * alpha_f);
  ApplySpans(spans, br, t);
}

}  // namespace

void RenderScene(const hipr_scene* s, uint8_t* px, int w, int h) {
  Target t{px, w, h};
  Aff ident;
  for (int p = 0; p < s->n_paths; p++) {
    const hipr_paint& paint = s->paints[p];
    if (paint.blur_radius > 0) {
      DrawPathBlurred(s, p, &t);
    } else {
      PathData path = ScenePath(s, p);
      DrawPathNoFilter(s, path, paint, ident, &t);
    }
  }
}

// RenderScene with the work split over threads; the result is the same image.
// Rasterising a draw reads nothing but its own path and paint, so the draws
// of a chunk are rasterised concurrently into span lists (and blurred
// layers); the spans are then applied in draw order, each thread owning a
// band of rows, which keeps every pixel's blend sequence (rows are
// independent, and a stable sort by row keeps each row's emission order).
namespace {
struct SpanOp {
  std::vector<Span> spans;  // sorted by row, emission order within a row
  std::vector<size_t> row_start;  // first span of each row offset from y0
  int y0 = 0;
  Brush br;
  std::vector<uint8_t> layer;  // pixels of an image brush
};

void FinishOp(SpanOp* op) {
  std::stable_sort(op->spans.begin(), op->spans.end(),
                   [](const Span& a, const Span& b) { return a.y < b.y; });
  if (op->spans.empty()) return;
  op->y0 = op->spans.front().y;
  int y1 = op->spans.back().y + 1;
  op->row_start.assign((size_t)(y1 - op->y0) + 1, 0);
  size_t i = 0;
  for (int y = op->y0; y <= y1; y++) {
    while (i < op->spans.size() && op->spans[i].y < y) i++;
    op->row_start[y - op->y0] = i;
  }
}

void ShapeOps(const hipr_scene* s, int p, std::vector<std::unique_ptr<SpanOp>>* ops) {
  const hipr_paint& paint = s->paints[p];
  if (paint.blur_radius > 0) {
    auto op = std::make_unique<SpanOp>();
    int fb[4];
    BlurredLayer(s, p, &op->layer, fb);
    int W = fb[2] - fb[0], H = fb[3] - fb[1];
    if (W <= 0 || H <= 0) return;
    PathData rect;
    float L = (float)fb[0], T = (float)fb[1], R = (float)fb[2], B = (float)fb[3];
    rect.MoveTo(Pt{L, T});
    rect.LineTo(Pt{R, T});
    rect.LineTo(Pt{R, B});
    rect.LineTo(Pt{L, B});
    rect.Close();
    float bounds[4];
    RastePath(rect, &op->spans, bounds, nullptr);
    op->br.kind = 2;
    op->br.img = op->layer.data();
    op->br.img_w = W;
    op->br.img_h = H;
    Aff inv;
    Invert(Translate(L, T), &inv);
    op->br.img_ptu = Mul(Scale(1.f / W, 1.f / H), inv);
    float alpha_f = paint.style == HIPR_STYLE_STROKE ? paint.stroke_color[3] : paint.color[3];
    op->br.global_alpha = (uint8_t)(255 * alpha_f);
    FinishOp(op.get());
    ops->push_back(std::move(op));
    return;
  }
  PathData path = ScenePath(s, p);
  Aff ident;
  auto shape = [&](bool stroke) {
    auto op = std::make_unique<SpanOp>();
    op->br = MakeBrush(s, paint, ident, stroke);
    float bounds[4];
    if (!stroke) {
      RastePath(path, &op->spans, bounds, nullptr);
    } else {
      PathData outline = StrokeOutline(QuadPath(path), paint);
      RastePath(outline, &op->spans, bounds, nullptr);
    }
    FinishOp(op.get());
    ops->push_back(std::move(op));
  };
  bool need_fill = paint.style != HIPR_STYLE_STROKE;
  bool need_stroke = paint.style != HIPR_STYLE_FILL;
  if (paint.style == HIPR_STYLE_STROKE_THEN_FILL) {
    if (need_stroke) shape(true);
    if (need_fill) shape(false);
  } else {
    if (need_fill) shape(false);
    if (need_stroke) shape(true);
  }
}

template <class F>
void ParallelFor(int n, int threads, F&& fn) {
  std::atomic<int> next{0};
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; t++)
    pool.emplace_back([&]() {
      for (int i; (i = next.fetch_add(1)) < n;) fn(i);
    });
  for (auto& th : pool) th.join();
}
}  // namespace

void RenderSceneParallel(const hipr_scene* s, uint8_t* px, int w, int h, int threads) {
  if (threads <= 1) {
    RenderScene(s, px, w, h);
    return;
  }
  Target t{px, w, h};
  const int chunk = 8 * threads;
  constexpr int kBand = 16;
  const int bands = (h + kBand - 1) / kBand;
  for (int p0 = 0; p0 < s->n_paths; p0 += chunk) {
    const int n = std::min(chunk, s->n_paths - p0);
    std::vector<std::vector<std::unique_ptr<SpanOp>>> ops(n);
    ParallelFor(n, threads, [&](int i) { ShapeOps(s, p0 + i, &ops[i]); });
    ParallelFor(bands, threads, [&](int b) {
      const int y0 = b * kBand, y1 = std::min(h, y0 + kBand);
      for (auto& draw : ops) {
        for (auto& op : draw) {
          if (op->spans.empty()) continue;
          const int oy1 = op->y0 + (int)op->row_start.size() - 1;
          const int a = std::max(y0, op->y0), e = std::min(y1, oy1);
          if (a >= e) continue;
          std::vector<Span> rows(op->spans.begin() + op->row_start[a - op->y0],
                                 op->spans.begin() + op->row_start[e - op->y0]);
          ApplySpans(rows, op->br, &t);
        }
      }
    });
  }
}

void PathSpans(const hipr_scene* s, int p, int stroke,
               std::vector<Span>* spans, float bounds[4],
               std::vector<Edge>* edges) {
  PathData path = ScenePath(s, p);
  if (!stroke) {
    RastePath(path, spans, bounds, edges);
  } else {
    PathData quad = QuadPath(path);
    PathData outline = StrokeOutline(quad, s->paints[p]);
    RastePath(outline, spans, bounds, edges);
  }
}

uint32_t GradientColorAt(const hipr_gradient& g, int x, int y) {
  GradientBrush b;
  PrepareGradient(g, Aff(), &b);
  return b.Color(x, y);
}

}  // namespace orc
