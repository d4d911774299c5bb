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

// ---- targets: SWRenderTarget::BlendPixel = FastBlend, else PorterDuffBlend
// with the target converted to and from premultiplied form when it is not
// premultiplied (sw_render_target.cc:12-141; blend_mode.cc:93-191;
// color_priv.cc:97-108); Bitmap::Get/SetPixel for kRGBA and kBGRA
// (bitmap.cc:25-85).
inline uint32_t Alpha255To256(uint32_t a) { return a + 1; }
inline uint32_t PMColorMul(uint32_t s, uint32_t d) {
  return SetARGB(MulDiv255Round(GetA(s), GetA(d)), MulDiv255Round(GetR(s), GetR(d)),
                 MulDiv255Round(GetG(s), GetG(d)), MulDiv255Round(GetB(s), GetB(d)));
}
// PMColorToColor: per channel (scale * c + 2^23) >> 24 in 32-bit arithmetic,
// scale = 255 * 2^24 / a rounded to nearest (the values of the reference's
// UnPreMultiply table, color_priv.cc:16-96).
inline uint32_t UnpremulScale(uint32_t a) {
  return a == 0 ? 0u : (uint32_t)((((uint64_t)255 << 24) + a / 2) / a);
}
inline uint32_t PMColorToColor(uint32_t c) {
  uint32_t a = GetA(c), sc = UnpremulScale(a);
  auto ap = [&](uint32_t v) { return (uint32_t)(sc * v + (1u << 23)) >> 24; };
  return SetARGB(a, ap(GetR(c)), ap(GetG(c)), ap(GetB(c)));
}
inline float SoftLightComponent(float sx, float sy, float dx, float dy) {
  // blend_mode.cc:93-110, with the reference's order of float operations
  if (2.f * sx <= sy) {
    return dx * dx * (sy - 2 * sx) / dy + (1 - dy) * sx + dx * (-sy + 2 * sx + 1);
  } else if (4.f * dx <= dy) {
    float DSqd = dx * dx;
    float DCub = DSqd * dx;
    float DaSqd = dy * dy;
    float DaCub = DaSqd * dy;
    return (DaSqd * (sx - dx * (3 * sy - 6 * sx - 1)) + 12 * dy * DSqd * (sy - 2 * sx) -
            16 * DCub * (sy - 2 * sx) - DaCub * sx) /
           DaSqd;
  } else {
    return dx * (sy - 2 * sx + 1) + sx - std::sqrt(dy * dx) * (sy - 2 * sx) - dy * sx;
  }
}
inline uint32_t SoftLight(uint32_t src, uint32_t dst) {
  if (GetA(dst) == 0) return src;
  float s[4] = {GetR(src) / 255.f, GetG(src) / 255.f, GetB(src) / 255.f, GetA(src) / 255.f};
  float d[4] = {GetR(dst) / 255.f, GetG(dst) / 255.f, GetB(dst) / 255.f, GetA(dst) / 255.f};
  float c[4] = {SoftLightComponent(s[0], s[3], d[0], d[3]),
                SoftLightComponent(s[1], s[3], d[1], d[3]),
                SoftLightComponent(s[2], s[3], d[2], d[3]), s[3] + (1 - s[3]) * d[3]};
  return Color4fToColor(c);
}
uint32_t PorterDuffBlend(uint32_t src, uint32_t dst, int mode) {
  if (mode > HIPR_BLEND_SCREEN && mode != HIPR_BLEND_SOFT_LIGHT) mode = HIPR_BLEND_SRC_OVER;
  const uint32_t sa = GetA(src), da = GetA(dst);
  switch (mode) {
    case HIPR_BLEND_CLEAR: return 0;
    case HIPR_BLEND_SRC: return src;
    case HIPR_BLEND_DST: return dst;
    case HIPR_BLEND_SRC_OVER: return sa == 0 ? dst : PMSrcOver(src, dst);
    case HIPR_BLEND_DST_OVER: return da == 255 ? dst : PMSrcOver(dst, src);
    case HIPR_BLEND_SRC_IN: return da == 255 ? src : AlphaMulQ(src, Alpha255To256(da));
    case HIPR_BLEND_DST_IN: return sa == 255 ? dst : AlphaMulQ(dst, Alpha255To256(sa));
    case HIPR_BLEND_SRC_OUT: return da == 0 ? src : AlphaMulQ(src, Alpha255To256(255 - da));
    case HIPR_BLEND_DST_OUT: return sa == 0 ? dst : AlphaMulQ(dst, Alpha255To256(255 - sa));
    case HIPR_BLEND_SRC_ATOP:
      return AlphaMulQ(src, Alpha255To256(da)) + AlphaMulQ(dst, Alpha255To256(255 - sa));
    case HIPR_BLEND_DST_ATOP:
      return AlphaMulQ(dst, Alpha255To256(sa)) + AlphaMulQ(src, Alpha255To256(255 - da));
    case HIPR_BLEND_XOR:
      return AlphaMulQ(src, Alpha255To256(255 - da)) + AlphaMulQ(dst, Alpha255To256(255 - sa));
    case HIPR_BLEND_PLUS:
      return SetARGB(std::min(sa + da, 255u), std::min(GetR(src) + GetR(dst), 255u),
                     std::min(GetG(src) + GetG(dst), 255u),
                     std::min(GetB(src) + GetB(dst), 255u));
    case HIPR_BLEND_MODULATE: return PMColorMul(src, dst);
    case HIPR_BLEND_SCREEN: return src + dst - PMColorMul(src, dst);
    case HIPR_BLEND_SOFT_LIGHT: return SoftLight(src, dst);
    default: return dst;
  }
}

struct Target {
  uint8_t* px;
  int w, h;
  int alpha_type = HIPR_ALPHA_PREMUL;
  int color_type = HIPR_COLOR_RGBA;
  uint32_t Get(int x, int y) const {
    const uint8_t* p = px + ((size_t)y * w + x) * 4;
    return color_type == HIPR_COLOR_BGRA ? SetARGB(p[3], p[2], p[1], p[0])
                                         : SetARGB(p[3], p[0], p[1], p[2]);
  }
  void Set(int x, int y, uint32_t c) {
    uint8_t* p = px + ((size_t)y * w + x) * 4;
    bool bgra = color_type == HIPR_COLOR_BGRA;
    p[0] = (uint8_t)(bgra ? GetB(c) : GetR(c));
    p[1] = (uint8_t)GetG(c);
    p[2] = (uint8_t)(bgra ? GetR(c) : GetB(c));
    p[3] = (uint8_t)GetA(c);
  }
  void Blend(int x, int y, uint32_t src, int mode) {
    if ((unsigned)x >= (unsigned)w || (unsigned)y >= (unsigned)h) return;
    // FastBlend (sw_render_target.cc:88-141)
    const uint32_t a = GetA(src);
    const uint8_t dst_a = px[((size_t)y * w + x) * 4 + 3];
    switch (mode) {
      case HIPR_BLEND_CLEAR: Set(x, y, 0); return;
      case HIPR_BLEND_SRC: Set(x, y, src); return;
      case HIPR_BLEND_DST: return;
      case HIPR_BLEND_SRC_OVER:
        if (a == 0) return;
        if (a == 255) {
          Set(x, y, src);
          return;
        }
        break;
      case HIPR_BLEND_DST_OVER:
        if (dst_a == 255) return;
        break;
      case HIPR_BLEND_DST_IN:
        if (a == 255) return;
        if (a == 0) {
          Set(x, y, 0);
          return;
        }
        break;
      case HIPR_BLEND_DST_OUT:
        if (a == 0) return;
        if (a == 255) {
          Set(x, y, 0);
          return;
        }
        break;
      default:
        break;
    }
    const bool pm = alpha_type == HIPR_ALPHA_PREMUL;
    uint32_t dst = pm ? Get(x, y) : ColorToPMColor(Get(x, y));
    uint32_t r = PorterDuffBlend(src, dst, mode);
    Set(x, y, pm ? r : PMColorToColor(r));
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

// PointsToUnit(p0, p1) of sw_span_brush.cc:200-219: rotate . scale .
// translate taking p0 to the origin and p1 to (1, 0).
Aff TwoPointsToUnit(float x0, float y0, float x1, float y1) {
  float ssx = x1 - x0, ssy = y1 - y0;
  float length = std::sqrt(ssx * ssx + ssy * ssy);
  float scale = length > 0 ? 1.0f / length : 0;
  float ux = ssx * scale, uy = ssy * scale;
  float sine = -uy, cosine = ux;
  Aff rot;
  rot.sx = cosine;
  rot.kx = -sine;
  rot.ky = sine;
  rot.sy = cosine;
  return Mul(Mul(rot, Scale(scale, scale)), Translate(-x0, -y0));
}

// ConicalGradientColorBrush (sw_span_brush.cc:385-506): the per-gradient
// set-up of OnPreBrush and the three cases of CalculateConical.
struct Conical {
  enum Kind { kTransparent, kRadial, kStrip, kFocal } kind = kTransparent;
  float c0x = 0, c0y = 0, scale = 0, sign = 1, bias = 0;  // kRadial
  float r2 = 0;                                          // kStrip
  Aff to_unit;                                           // kStrip, kFocal
  float f = 0, r1 = 0, r1sq = 0;                         // kFocal
  bool swap = false;

  void Setup(const hipr_gradient& g) {
    const float kNear = 1.f / 4096;
    float c0[2] = {g.points[0], g.points[1]}, c1[2] = {g.points[2], g.points[3]};
    float r0 = g.radius[0], r_1 = g.radius[1];
    float dcx = c1[0] - c0[0], dcy = c1[1] - c0[1];
    float delta_center = std::sqrt(dcx * dcx + dcy * dcy);
    float delta_radius = std::abs(r_1 - r0);
    kind = kTransparent;
    if (r0 < 0 || r_1 < 0) return;
    bool radial = delta_center < kNear, strip = delta_radius < kNear;
    if (radial) {
      if (strip) return;
      kind = kRadial;
      scale = (float)(1.0 / delta_radius);
      sign = delta_radius < 0 ? -1.f : 1.f;
      bias = r0 / delta_radius;
      c0x = c0[0];
      c0y = c0[1];
    } else if (strip) {
      kind = kStrip;
      float r = r0 / delta_center;
      r2 = r * r;
      to_unit = TwoPointsToUnit(c0[0], c0[1], c1[0], c1[1]);
    } else {
      kind = kFocal;
      swap = r_1 < kNear;
      if (swap) {
        std::swap(c0[0], c1[0]);
        std::swap(c0[1], c1[1]);
        std::swap(r0, r_1);
      }
      f = r0 / (r0 - r_1);
      // points are (x, y, 0, 1): the focal point's w is (1 - f) + f
      float cfx = c0[0] * (1.f - f) + c1[0] * f;
      float cfy = c0[1] * (1.f - f) + c1[1] * f;
      float cfw = 1.f * (1.f - f) + 1.f * f;
      float ex = c1[0] - cfx, ey = c1[1] - cfy, ew = 1.f - cfw;
      float len = std::sqrt(ex * ex + ey * ey + 0.f * 0.f + ew * ew);
      r1 = r_1 / len;
      r1sq = r1 * r1;
      to_unit = TwoPointsToUnit(cfx, cfy, c1[0], c1[1]);
    }
  }

  // false: the pixel is transparent
  bool T(float px, float py, float* t) const {
    const float kNear = 1.f / 4096;
    if (kind == kTransparent) return false;
    if (kind == kRadial) {
      float ax = (px - c0x) * scale, ay = (py - c0y) * scale;
      *t = std::sqrt(ax * ax + ay * ay) * sign - bias;
      return true;
    }
    float qx, qy;
    MapPoint(to_unit, px, py, &qx, &qy);
    if (kind == kStrip) {
      float v = r2 - qy * qy;
      if (v < 0.0) return false;
      *t = qx + std::sqrt(v);
      return true;
    }
    float xt = -1.f;
    if (std::abs(r1 - 1.f) < kNear) {
      xt = (qx * qx + qy * qy) / 2;
    } else if (r1 > 1.f) {
      float m = r1sq - 1.f;
      float delta = m * qy * qy + r1sq * qx * qx;
      xt = (std::sqrt(delta) - qx) / m;
    } else {
      float m = r1sq - 1.f;
      float delta = m * qy * qy + r1sq * qx * qx;
      if (delta > 0) {
        float xt1 = (std::sqrt(delta) - qx) / m;
        float xt2 = (-std::sqrt(delta) - qx) / m;
        xt = 1.f - f < 0 ? std::min(xt1, xt2) : std::max(xt1, xt2);
      }
    }
    if (xt < 0) return false;
    float tt = f + (1.f - f) * xt;
    if (swap) tt = 1.0 - tt;
    *t = tt;
    return true;
  }
};

struct GradientBrush {
  hipr_gradient g;
  std::vector<float> offsets;  // after EnsureColorOffsetsNonDescending
  Aff points_to_unit;
  Conical conical;

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
    } else if (g.type == HIPR_GRADIENT_CONICAL) {
      if (!conical.T(mx, my, &t)) return 0;
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
    ptu = TwoPointsToUnit(g.points[0], g.points[1], g.points[2], g.points[3]);
  } else if (g.type == HIPR_GRADIENT_RADIAL) {
    float r = g.radius[0];
    float scale = r > 0 ? 1.0f / r : 0;
    ptu = Mul(Scale(scale, scale), Translate(-g.points[0], -g.points[1]));
  } else if (g.type == HIPR_GRADIENT_CONICAL) {
    ptu = Aff();  // the conical brush maps device -> local itself
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
  if (g.type == HIPR_GRADIENT_CONICAL) b->conical.Setup(g);
}

// A brush: solid colour, gradient, or the premultiplied image of a blurred
// draw sampled nearest through a decal pixmap shader (PixmapBrush,
// sw_span_brush.cc:555-579; BitmapSampler, bitmap_sampler.cc:26-40,85-108).
// BitmapSampler (bitmap_sampler.cc:26-108) over an RGBA8 image.  Float ->
// uint32 conversions are x86-64's (through a 64-bit integer), so the -1
// column of a bilinear footprint at the left edge wraps to 2^32 - 1 and is
// clamped to the last column, as the reference's x86-64 build does.
struct Sampler {
  const uint8_t* px = nullptr;
  uint32_t w = 0, h = 0;
  int filter = HIPR_FILTER_NEAREST, tx = HIPR_TILE_DECAL, ty = HIPR_TILE_DECAL;

  static uint32_t ToU32(float f) { return (uint32_t)(int64_t)f; }
  static float Mod(float x, float y) { return x - y * std::floor(x / y); }
  static float Fract(float x) { return x - std::floor(x); }
  void SampleXY(float fx, float fy, float out[4]) const {
    uint32_t ix = std::min(ToU32(fx), w - 1), iy = std::min(ToU32(fy), h - 1);
    const uint8_t* p = px + ((size_t)iy * w + ix) * 4;
    for (int c = 0; c < 4; c++) out[c] = p[c] / 255.f;  // Color4fFromColor
  }
  uint32_t Color(float u, float v) const {  // BitmapSampler::GetColor, ARGB
    if ((tx == HIPR_TILE_DECAL && (u < 0.0 || u >= 1.0)) ||
        (ty == HIPR_TILE_DECAL && (v < 0.0 || v >= 1.0)))
      return 0;
    u = RemapFloatTile(u, tx);
    v = RemapFloatTile(v, ty);
    float c[4];
    if (filter == HIPR_FILTER_NEAREST) {
      SampleXY(u * w, v * h, c);
      return Color4fToColor(c);
    }
    float x = u * w, y = v * h;
    float i0 = std::floor(x - 0.5f), j0 = std::floor(y - 0.5f);
    if (tx == HIPR_TILE_REPEAT) i0 = Mod(i0, (float)w);
    if (ty == HIPR_TILE_REPEAT) j0 = Mod(j0, (float)h);
    float i1 = i0 + 1.0f, j1 = j0 + 1.0f;
    if (tx == HIPR_TILE_REPEAT) i1 = Mod(i1, (float)w);
    if (ty == HIPR_TILE_REPEAT) j1 = Mod(j1, (float)h);
    float a = Fract(x - 0.5f), b = Fract(y - 0.5f);
    float t00[4], t10[4], t01[4], t11[4];
    SampleXY(i0, j0, t00);
    SampleXY(i1, j0, t10);
    SampleXY(i0, j1, t01);
    SampleXY(i1, j1, t11);
    float w00 = (1 - a) * (1 - b), w10 = a * (1 - b), w01 = (1 - a) * b, w11 = a * b;
    for (int k = 0; k < 4; k++)
      c[k] = ((w00 * t00[k] + w10 * t10[k]) + w01 * t01[k]) + w11 * t11[k];
    return Color4fToColor(c);
  }
};

struct Brush {
  int kind = 0;  // 0 solid, 1 gradient, 2 image (blurred layer), 3 pixmap shader
  Sampler pix;           // kind 3
  bool pix_unpremul = false;
  Aff pix_ptu;
  uint32_t solid = 0;
  GradientBrush grad;
  const uint8_t* img = nullptr;
  int img_w = 0, img_h = 0;
  Aff img_ptu;
  uint8_t global_alpha = 255;
  // the paint's ColorFilter (color_filter.cc:113-185)
  int cf_type = HIPR_COLOR_FILTER_NONE;
  uint32_t cf_pm = 0;  // BLEND: ColorToPMColor of the filter colour
  int cf_mode = 0;
  int32_t cf_m[4][5] = {};  // MATRIX: (int16_t)(m * 255)

  uint32_t Filter(uint32_t c) const;

  uint32_t Color(int x, int y) const {
    if (kind == 0) return solid;
    if (kind == 1) return grad.Color(x, y);
    if (kind == 3) {  // PixmapBrush::CalculateColor (sw_span_brush.cc:569-579)
      float u, v;
      MapPoint(pix_ptu, x + 0.5f, y + 0.5f, &u, &v);
      uint32_t c = pix.Color(u, v);
      return pix_unpremul ? ColorToPMColor(c) : c;
    }
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
template <class It>
void ApplySpans(It begin, It end, const Brush& br, int blend, Target* t) {
  for (It it = begin; it != end; ++it) {
    const Span& s = *it;
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
      if (br.cf_type != HIPR_COLOR_FILTER_NONE) c = br.Filter(c);
      t->Blend(x + l, y, c, blend);
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

// The sRGB transfer tables of SRGBGammaColorFilter (color_filter.cc:128-166),
// from their definition: entry i = floor(255 f(i / 255)) with f the sRGB
// encoding (linear -> sRGB) or decoding, in double precision (a 1e-6 nudge
// keeps f(1) = 1 from rounding below 255).
const uint8_t* SrgbTable(bool linear_to_srgb) {
  static uint8_t tab[2][256];
  static bool init = [] {
    for (int i = 0; i < 256; i++) {
      double x = i / 255.0;
      double e = x <= 0.0031308 ? 12.92 * x : 1.055 * std::pow(x, 1 / 2.4) - 0.055;
      double d = x <= 0.04045 ? x / 12.92 : std::pow((x + 0.055) / 1.055, 2.4);
      tab[0][i] = (uint8_t)std::floor(255 * e + 1e-6);
      tab[1][i] = (uint8_t)std::floor(255 * d + 1e-6);
    }
    return true;
  }();
  (void)init;
  return tab[linear_to_srgb ? 0 : 1];
}

// ColorFilter::FilterColor of a premultiplied ARGB colour.
uint32_t FilterPM(const Brush& br, uint32_t c) {
  switch (br.cf_type) {
    case HIPR_COLOR_FILTER_BLEND:
      return PorterDuffBlend(br.cf_pm, c, br.cf_mode);
    case HIPR_COLOR_FILTER_MATRIX: {
      uint32_t u = PMColorToColor(c);
      int32_t v[4] = {(int32_t)GetR(u), (int32_t)GetG(u), (int32_t)GetB(u), (int32_t)GetA(u)};
      uint32_t o[4];
      for (int i = 0; i < 4; i++) {
        int32_t mul = v[0] * br.cf_m[i][0] + v[1] * br.cf_m[i][1] + v[2] * br.cf_m[i][2] +
                      v[3] * br.cf_m[i][3];
        o[i] = (uint32_t)std::clamp(mul / 255 + br.cf_m[i][4], 0, 255);
      }
      return ColorToPMColor(SetARGB(o[3], o[0], o[1], o[2]));
    }
    case HIPR_COLOR_FILTER_LINEAR_TO_SRGB:
    case HIPR_COLOR_FILTER_SRGB_TO_LINEAR: {
      const uint8_t* t = SrgbTable(br.cf_type == HIPR_COLOR_FILTER_LINEAR_TO_SRGB);
      uint32_t u = PMColorToColor(c);
      return ColorToPMColor(SetARGB(GetA(u), t[GetR(u)], t[GetG(u)], t[GetB(u)]));
    }
    default:  // NONE; COMPOSE passes colours through (color_filter.cc:182-185)
      return c;
  }
}

uint32_t Brush::Filter(uint32_t c) const { return FilterPM(*this, c); }

// GenerateBrush's pixmap branch (sw_canvas.cc:755-788): the image shader of
// draw p replaces the brush colour; points to unit = Scale(1/w, 1/h) .
// inverse(local) . inverse(CTM); global alpha Paint::GetAlphaF (the stroke
// colour's alpha for kStroke_Style, else the fill colour's, paint.cc:96-98).
void SetImageShader(const hipr_scene* s, int p, const Aff& ctm, Brush* br) {
  if (!s->image_shaders || s->image_shaders[p].image < 0) return;
  const hipr_image_shader& sh = s->image_shaders[p];
  const hipr_image& im = s->images[sh.image];
  br->kind = 3;
  br->pix.px = im.pixels;
  br->pix.w = (uint32_t)im.width;
  br->pix.h = (uint32_t)im.height;
  br->pix.filter = sh.filter;
  br->pix.tx = sh.tile_x;
  br->pix.ty = sh.tile_y;
  br->pix_unpremul = im.alpha_type == HIPR_ALPHA_UNPREMUL;
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
  br->pix_ptu = Mul(Mul(Scale(1.f / (uint32_t)im.width, 1.f / (uint32_t)im.height), inv), l2l);
  const hipr_paint& pt = s->paints[p];
  float alpha = pt.style == HIPR_STYLE_STROKE ? pt.stroke_color[3] : pt.color[3];
  br->global_alpha = (uint8_t)(255 * alpha);
}

void SetColorFilter(const hipr_scene* s, int p, Brush* br) {
  if (!s->color_filters) return;
  const hipr_color_filter& f = s->color_filters[p];
  br->cf_type = f.type;
  br->cf_mode = f.blend;
  br->cf_pm = ColorToPMColor(f.color);
  for (int i = 0; i < 20; i++) br->cf_m[i / 5][i % 5] = (int16_t)(f.matrix[i] * 255);
}

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
void AffToArray(const Aff& m, float f[6]) {
  f[0] = m.sx;
  f[1] = m.kx;
  f[2] = m.tx;
  f[3] = m.ky;
  f[4] = m.sy;
  f[5] = m.ty;
}

// One rasterisation and its brush: the spans sorted by row (emission order
// within a row, which keeps each pixel's blend sequence), the first span of
// each row, and the pixels of an image brush.
struct SpanOp {
  std::vector<Span> spans;
  std::vector<size_t> row_start;  // first span of row y0 + i
  int y0 = 0;
  Brush br;
  int blend = HIPR_BLEND_SRC_OVER;
  std::vector<uint8_t> layer;
};
using Ops = std::vector<std::unique_ptr<SpanOp>>;

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

// Rows [y0, y1) of an op onto a target.
void ApplyOp(const SpanOp& op, Target* t, int y0, int y1) {
  if (op.spans.empty()) return;
  const int oy1 = op.y0 + (int)op.row_start.size() - 1;
  const int a = std::max(y0, op.y0), e = std::min(y1, oy1);
  if (a >= e) return;
  ApplySpans(op.spans.begin() + op.row_start[a - op.y0],
             op.spans.begin() + op.row_start[e - op.y0], op.br, op.blend, t);
}

// The fill and/or stroke rasterisations of a path under canvas matrix `ctm`,
// in paint order (sw_canvas.cc:357-411, paint_order.hpp:12-31).
void ShapeOps(const hipr_scene* s, const PathData& path, const hipr_paint& paint,
              const Aff& ctm, Ops* ops, const float* clip = nullptr) {
  float m6[6];
  AffToArray(ctm, m6);
  auto shape = [&](bool stroke) {
    auto op = std::make_unique<SpanOp>();
    op->br = MakeBrush(s, paint, ctm, stroke);
    op->blend = paint.blend;
    float bounds[4];
    if (!stroke) {
      RastePath(path, &op->spans, bounds, nullptr, m6, clip);
    } else {
      PathData outline = StrokeOutline(QuadPath(path), paint);
      RastePath(outline, &op->spans, bounds, nullptr, m6, clip);
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

void DrawOps(const Ops& ops, Target* t) {
  for (const auto& op : ops) ApplyOp(*op, t, 0, t->h);
}

// The rectangle path of Path::AddRect(l, t, r, b) (path.cc:815-835).
PathData RectPath(float L, float T, float R, float B) {
  PathData rect;
  rect.MoveTo(Pt{L, T});
  rect.LineTo(Pt{R, T});
  rect.LineTo(Pt{R, B});
  rect.LineTo(Pt{L, B});
  rect.Close();
  return rect;
}

// MaskFilterOnFilter's post-passes for the other blur styles
// (mask_filter.cc:60-103): raw is the layer before the blur.
void BlurStylePass(int style, const std::vector<uint8_t>& raw, std::vector<uint8_t>* blurred) {
  if (style == HIPR_BLUR_NORMAL) return;
  const size_t n = raw.size() / 4;
  for (size_t i = 0; i < n; i++) {
    const uint8_t* r = &raw[4 * i];
    uint8_t* b = &(*blurred)[4 * i];
    uint32_t raw_a = r[3], blur_a = b[3];
    if (style == HIPR_BLUR_SOLID) {
      if (raw_a > 0) std::memcpy(b, r, 4);
    } else if (style == HIPR_BLUR_OUTER) {
      if (raw_a > 0 && raw_a >= blur_a) std::memset(b, 0, 4);
    } else if (style == HIPR_BLUR_INNER) {
      if (raw_a > 0) {
        constexpr float a_factor = 1.f / 255.f;
        uint32_t scale = (uint32_t)(a_factor * raw_a * blur_a);
        uint32_t c = AlphaMulQ(SetARGB(b[3], b[0], b[1], b[2]), scale);
        b[0] = (uint8_t)GetR(c);
        b[1] = (uint8_t)GetG(c);
        b[2] = (uint8_t)GetB(c);
        b[3] = (uint8_t)GetA(c);
      } else {
        std::memset(b, 0, 4);
      }
    }
  }
}

}  // namespace

// SWCanvas::HandleFilter + MaskFilterOnFilter + OnDrawImageRect
// (sw_canvas.cc:135-144,641-677,797-826; mask_filter.cc:51-103;
// image_filter.cc:33-41,184-194).  The layer is a fresh premultiplied RGBA
// canvas translated by -filter_bounds.LT (the draw's canvas matrix does not
// reach it); returns the filter bounds and leaves the blurred (and styled)
// premultiplied image in *blurred.
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
  blurred->clear();
  if (W <= 0 || H <= 0) return;
  std::vector<uint8_t> layer((size_t)W * H * 4, 0);
  Target lt{layer.data(), W, H};
  hipr_paint work = paint;
  work.blur_radius = 0;
  Ops ops;
  ShapeOps(s, path, work, Translate(-l, -tp), &ops);
  // the layer is drawn with the paint's colour filter and shader
  // (HandleFilter keeps them in work_paint, sw_canvas.cc:797-826)
  for (auto& op : ops) {
    SetColorFilter(s, p, &op->br);
    SetImageShader(s, p, Translate(-l, -tp), &op->br);
  }
  DrawOps(ops, &lt);
  blurred->assign((size_t)W * H * 4, 0);
  int radius = (int)std::round(r);
  StackBlur(layer.data(), blurred->data(), W, H, radius);
  BlurStylePass(paint.blur_style, layer, blurred);
}

namespace {

// The rasterisations of draw p onto a tw x th target.
void DrawEntryOpsUnfiltered(const hipr_scene* s, int p, int tw, int th, Ops* ops);

// Every brush of draw p carries the paint's colour filter: fills, strokes,
// DrawPaint and the composite of a blurred layer (DrawImage with the work
// paint, image_filter.cc:184-194).
void DrawEntryOps(const hipr_scene* s, int p, int tw, int th, Ops* ops) {
  const size_t first = ops->size();
  DrawEntryOpsUnfiltered(s, p, tw, th, ops);
  for (size_t i = first; i < ops->size(); i++) {
    Brush* br = &(*ops)[i]->br;
    SetColorFilter(s, p, br);
    // the blurred layer's composite samples the layer instead of the shader
    if (br->kind != 2) SetImageShader(s, p, DrawMatrix(s, p), br);
  }
}

void DrawEntryOpsUnfiltered(const hipr_scene* s, int p, int tw, int th, Ops* ops) {
  const hipr_paint& paint = s->paints[p];
  const Aff ctm = DrawMatrix(s, p);
  // the canvas's global clip bounds at this draw (GetScanClipBounds,
  // sw_canvas.hpp:157-161): the scan clip of its path and stroke rasters
  // and of a blurred layer's image draw; DrawPaint rasters without it
  const float* clip = s->clip_rects ? s->clip_rects + 4 * (size_t)p : nullptr;
  const int kind = s->draw_kinds ? s->draw_kinds[p] : HIPR_DRAW_PATH;
  if (kind == HIPR_DRAW_PAINT) {
    // OnDrawPaint: the target's rectangle, identity raster matrix, the
    // paint's fill brush (its shader through the canvas matrix).
    auto op = std::make_unique<SpanOp>();
    float bounds[4];
    RastePath(RectPath(0.f, 0.f, (float)tw, (float)th), &op->spans, bounds, nullptr);
    op->br = MakeBrush(s, paint, ctm, false);
    op->blend = paint.blend;
    FinishOp(op.get());
    ops->push_back(std::move(op));
    return;
  }
  if (paint.blur_radius > 0) {
    auto op = std::make_unique<SpanOp>();
    int fb[4];
    BlurredLayer(s, p, &op->layer, fb);
    int W = fb[2] - fb[0], H = fb[3] - fb[1];
    if (W <= 0 || H <= 0) return;
    // OnDrawImageRect onto this canvas: the filter-bounds rectangle through
    // the canvas matrix, a decal nearest pixmap shader whose local matrix
    // places the image on the rectangle, and the paint (style set to fill,
    // so its alpha is the fill colour's, Paint::GetAlphaF, paint.cc:96-98).
    float L = (float)fb[0], T = (float)fb[1], R = (float)fb[2], B = (float)fb[3];
    float m6[6];
    AffToArray(ctm, m6);
    float bounds[4];
    RastePath(RectPath(L, T, R, B), &op->spans, bounds, nullptr, m6, clip);
    op->br.kind = 2;
    op->br.img = op->layer.data();
    op->br.img_w = W;
    op->br.img_h = H;
    Aff inv, l2l;
    Invert(Translate(L, T), &inv);
    if (!Invert(ctm, &l2l)) l2l = Aff();
    op->br.img_ptu = Mul(Mul(Scale(1.f / W, 1.f / H), inv), l2l);
    op->br.global_alpha = (uint8_t)(255 * paint.color[3]);
    op->blend = paint.blend;
    FinishOp(op.get());
    ops->push_back(std::move(op));
    return;
  }
  ShapeOps(s, ScenePath(s, p), paint, ctm, ops, clip);
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

void RenderSceneEx(const hipr_scene* s, uint8_t* px, int w, int h, int alpha_type,
                   int color_type, int threads) {
  Target t{px, w, h, alpha_type, color_type};
  if (threads <= 1) {
    for (int p = 0; p < s->n_paths; p++) {
      Ops ops;
      DrawEntryOps(s, p, w, h, &ops);
      DrawOps(ops, &t);
    }
    return;
  }
  // Rasterising a draw reads nothing but its own path and paint, so the
  // draws of a chunk are rasterised concurrently; their spans are then
  // applied in draw order, each thread owning a band of rows (rows are
  // independent): the same image as the serial loop.
  const int chunk = 8 * threads;
  constexpr int kBand = 16;
  const int bands = (h + kBand - 1) / kBand;
  for (int p0 = 0; p0 < s->n_paths; p0 += chunk) {
    const int n = std::min(chunk, s->n_paths - p0);
    std::vector<Ops> ops(n);
    ParallelFor(n, threads, [&](int i) { DrawEntryOps(s, p0 + i, w, h, &ops[i]); });
    ParallelFor(bands, threads, [&](int b) {
      const int y0 = b * kBand, y1 = std::min(h, y0 + kBand);
      for (auto& draw : ops)
        for (auto& op : draw) ApplyOp(*op, &t, y0, y1);
    });
  }
}

void RenderScene(const hipr_scene* s, uint8_t* px, int w, int h) {
  RenderSceneEx(s, px, w, h, HIPR_ALPHA_PREMUL, HIPR_COLOR_RGBA, 1);
}

void RenderSceneParallel(const hipr_scene* s, uint8_t* px, int w, int h, int threads) {
  RenderSceneEx(s, px, w, h, HIPR_ALPHA_PREMUL, HIPR_COLOR_RGBA, threads);
}

void PathSpans(const hipr_scene* s, int p, int stroke,
               std::vector<Span>* spans, float bounds[4],
               std::vector<Edge>* edges) {
  PathData path = ScenePath(s, p);
  float m6[6];
  AffToArray(DrawMatrix(s, p), m6);
  if (!stroke) {
    RastePath(path, spans, bounds, edges, m6);
  } else {
    PathData quad = QuadPath(path);
    PathData outline = StrokeOutline(quad, s->paints[p]);
    RastePath(outline, spans, bounds, edges, m6);
  }
}

uint32_t GradientColorAt(const hipr_gradient& g, int x, int y) {
  GradientBrush b;
  PrepareGradient(g, Aff(), &b);
  return b.Color(x, y);
}

}  // namespace orc
