// hipr_composite.hpp -- stage (b)/(c) output: shade and blend coverage tiles.
//
// One wavefront per 16x16 canvas tile; lane l owns the 4 pixels of tile row
// l / 4, columns 4*(l % 4) .. +3 (one 16-byte RGBA8 load and store, coalesced
// across the wave).  The tile walks the canvas's draws in order, 64 at a time:
// each lane tests one draw (tile inside its visible tile range, map entry
// present), a ballot collects the draws that touch the tile, and they are
// applied in order.  Per pixel and cover this is SWSpanBrush::Brush/BrushH
// (src/render/sw/sw_span_brush.cc:66-138): cover & global alpha, AlphaMulQ
// when the cover is not 255, then SWRenderTarget::BlendPixel for kSrcOver
// (sw_render_target.cc:12-141, blend_mode.cc:125-191).  The target keeps
// skity's kRGBA byte order; every channel is computed independently, so the
// packed u32 arithmetic is the reference's AlphaMulQ/PMSrcOver.
#ifndef SKITY_HIP_COMPOSITE_HPP_
#define SKITY_HIP_COMPOSITE_HPP_

#include "hipr_walk.hpp"

namespace hipr {

enum PaintKind : int32_t {
  kPaintSolid = 0,
  kPaintGradient = 1,
  kPaintImage = 2,   // a blurred layer's composite (nearest, decal)
  kPaintPixmap = 3   // a pixmap shader
};

// An image of a pixmap shader: RGBA8 rows at `offset` in the scene's image
// buffer.
struct ImageDev {
  int32_t width, height;
  int64_t offset;
};

// Two-point conical gradient set-up (ConicalGradientColorBrush::OnPreBrush,
// sw_span_brush.cc:407-440), computed on the host.
enum ConicalKind : int32_t { kConicalNone = 0, kConicalRadial, kConicalStrip, kConicalFocal };
struct ConicalDev {
  int32_t kind, swap;
  float c0x, c0y, scale, sign, bias;  // concentric circles
  float r2;                           // equal radii: (r0 / |c1 - c0|)^2
  float ux, ukx, utx, uky, uy, uty;   // strip / focal: points -> unit
  float f, r1, r1sq;                  // focal
};

struct GradDev {
  int32_t type, tile, count, has_offsets;
  float sx, kx, tx, ky, sy, ty;  // device -> unit space (conical: device -> local)
  float bias, scale;             // sweep parameters
  float offsets[16];
  float colors[16][4];
  ConicalDev con;
};

// CalculateConical (sw_span_brush.cc:443-506): t at the local point (px, py),
// false where the pixel is transparent.
__device__ __forceinline__ bool ConicalT(const ConicalDev& c, float px, float py, float* t) {
  const float kNear = 1.f / 4096;
  if (c.kind == kConicalNone) return false;
  if (c.kind == kConicalRadial) {
    float ax = (px - c.c0x) * c.scale, ay = (py - c.c0y) * c.scale;
    *t = sqrtf(ax * ax + ay * ay) * c.sign - c.bias;
    return true;
  }
  float qx = px * c.ux + py * c.ukx + c.utx;
  float qy = px * c.uky + py * c.uy + c.uty;
  if (c.kind == kConicalStrip) {
    float v = c.r2 - qy * qy;
    if (v < 0.0f) return false;
    *t = qx + sqrtf(v);
    return true;
  }
  float xt = -1.f;
  if (fabsf(c.r1 - 1.f) < kNear) {
    xt = (qx * qx + qy * qy) / 2;
  } else if (c.r1 > 1.f) {
    float m = c.r1sq - 1.f;
    float delta = m * qy * qy + c.r1sq * qx * qx;
    xt = (sqrtf(delta) - qx) / m;
  } else {
    float m = c.r1sq - 1.f;
    float delta = m * qy * qy + c.r1sq * qx * qx;
    if (delta > 0) {
      float xt1 = (sqrtf(delta) - qx) / m;
      float xt2 = (-sqrtf(delta) - qx) / m;
      // std::min / std::max
      xt = 1.f - c.f < 0 ? (xt2 < xt1 ? xt2 : xt1) : (xt1 < xt2 ? xt2 : xt1);
    }
  }
  if (xt < 0) return false;
  float tt = c.f + (1.f - c.f) * xt;
  if (c.swap) tt = 1.f - tt;
  *t = tt;
  return true;
}

struct PaintDev {
  int32_t kind;
  uint32_t color;   // premultiplied, RGBA byte order (solid)
  int32_t ga;       // global alpha byte (cover & ga)
  int32_t grad;     // GradDev index
  int32_t layer;    // image: layer canvas index
  int32_t blend;    // HIPR_BLEND_* of the paint
  int32_t L, T, R, B;  // kJobRect: the rectangle of full coverage
  float im[6];      // image: device point -> unit square of the layer
  // pixmap shader: image, HIPR_FILTER_*, HIPR_TILE_* per axis, whether the
  // image is unpremultiplied; im[] maps device points to its unit square
  int32_t pix_img, pix_filter, pix_tx, pix_ty, pix_unpremul;
  // the paint's colour filter (HIPR_COLOR_FILTER_*)
  int32_t cf_type;
  uint32_t cf_pm;   // BLEND: the filter colour, premultiplied, RGBA byte order
  int32_t cf_mode;  // BLEND: HIPR_BLEND_*
  int32_t cf_m[20]; // MATRIX: (int16_t)(m * 255), row-major 4 x 5
};

__device__ __forceinline__ uint32_t AlphaMulQ(uint32_t c, uint32_t scale) {
  const uint32_t mask = 0xFF00FF;
  uint32_t rb = ((c & mask) * scale) >> 8;
  uint32_t ag = ((c >> 8) & mask) * scale;
  return (rb & mask) | (ag & ~mask);
}
__device__ __forceinline__ uint32_t BlendSrcOver(uint32_t src, uint32_t dst) {
  uint32_t a = src >> 24;
  if (a == 0) return dst;
  if (a == 255) return src;
  return src + AlphaMulQ(dst, 256 - a);
}
__device__ __forceinline__ uint32_t MulDiv255Round(uint32_t a, uint32_t b) {
  uint32_t prod = a * b + 128;
  return (prod + (prod >> 8)) >> 8;
}
__device__ __forceinline__ uint32_t UnitToByte(float v) {
  float s = v * 255;
  s = fminf(fmaxf(s, 0.f), 255.f);
  return (uint32_t)(uint8_t)s;
}
// Color4fToColor then ColorToPMColor (color.cc:53-59, color_priv.cc:103-106),
// packed in RGBA byte order.
__device__ __forceinline__ uint32_t PremulRGBA(const float c[4]) {
  uint32_t a = UnitToByte(c[3]), r = UnitToByte(c[0]), g = UnitToByte(c[1]),
           b = UnitToByte(c[2]);
  if (a != 255) {
    r = MulDiv255Round(r, a);
    g = MulDiv255Round(g, a);
    b = MulDiv255Round(b, a);
  }
  return r | (g << 8) | (b << 16) | (a << 24);
}

// ---- SWRenderTarget::BlendPixel (sw_render_target.cc:12-141) ------------
// Words hold one pixel with red in bits 0-7 and alpha in bits 24-31; every
// operation is per channel, so the reference's ARGB register arithmetic
// applies unchanged.
__device__ __forceinline__ uint32_t Chan(uint32_t c, int k) { return (c >> (8 * k)) & 0xFF; }
__device__ __forceinline__ uint32_t Pack(uint32_t r, uint32_t g, uint32_t b, uint32_t a) {
  return r | (g << 8) | (b << 16) | (a << 24);
}
// ColorToPMColor (color_priv.cc:103-106)
__device__ __forceinline__ uint32_t PremulWord(uint32_t c) {
  uint32_t a = c >> 24;
  if (a == 255) return c;
  return Pack(MulDiv255Round(Chan(c, 0), a), MulDiv255Round(Chan(c, 1), a),
              MulDiv255Round(Chan(c, 2), a), a);
}
// PMColorToColor (color_priv.cc:97-103): scale = 255 * 2^24 / a rounded to
// nearest, channel = (scale * c + 2^23) >> 24 in 32-bit arithmetic.
__device__ __forceinline__ uint32_t UnpremulWord(uint32_t c) {
  uint32_t a = c >> 24;
  uint32_t sc = a == 0 ? 0u : ((255u << 24) + a / 2) / a;
  auto ap = [&](uint32_t v) { return (sc * v + (1u << 23)) >> 24; };
  return Pack(ap(Chan(c, 0)), ap(Chan(c, 1)), ap(Chan(c, 2)), a);
}
__device__ __forceinline__ uint32_t PMColorMulWord(uint32_t s, uint32_t d) {
  return Pack(MulDiv255Round(Chan(s, 0), Chan(d, 0)), MulDiv255Round(Chan(s, 1), Chan(d, 1)),
              MulDiv255Round(Chan(s, 2), Chan(d, 2)), MulDiv255Round(s >> 24, d >> 24));
}
// SoftLightComponent (blend_mode.cc:93-110), the reference's order of float
// operations (this file is built without contraction).
__device__ float SoftLightComponent(float sx, float sy, float dx, float dy) {
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
  }
  return dx * (sy - 2 * sx + 1) + sx - sqrtf(dy * dx) * (sy - 2 * sx) - dy * sx;
}
__device__ uint32_t SoftLightWord(uint32_t src, uint32_t dst) {
  if ((dst >> 24) == 0) return src;
  float sa = (float)(src >> 24) / 255.f, da = (float)(dst >> 24) / 255.f;
  float c[4];
#pragma unroll
  for (int k = 0; k < 3; k++)
    c[k] = SoftLightComponent((float)Chan(src, k) / 255.f, sa, (float)Chan(dst, k) / 255.f, da);
  c[3] = sa + (1 - sa) * da;
  return Pack(UnitToByte(c[0]), UnitToByte(c[1]), UnitToByte(c[2]), UnitToByte(c[3]));
}
// PorterDuffBlend (blend_mode.cc:125-191): modes past kScreen other than
// kSoftLight fall back to kSrcOver.
__device__ uint32_t PorterDuffWord(uint32_t src, uint32_t dst, int mode) {
  if (mode > HIPR_BLEND_SCREEN && mode != HIPR_BLEND_SOFT_LIGHT) mode = HIPR_BLEND_SRC_OVER;
  const uint32_t sa = src >> 24, da = dst >> 24;
  switch (mode) {
    case HIPR_BLEND_CLEAR: return 0;
    case HIPR_BLEND_SRC: return src;
    case HIPR_BLEND_DST: return dst;
    case HIPR_BLEND_SRC_OVER: return sa == 0 ? dst : src + AlphaMulQ(dst, 256 - sa);
    case HIPR_BLEND_DST_OVER: return da == 255 ? dst : dst + AlphaMulQ(src, 256 - da);
    case HIPR_BLEND_SRC_IN: return da == 255 ? src : AlphaMulQ(src, da + 1);
    case HIPR_BLEND_DST_IN: return sa == 255 ? dst : AlphaMulQ(dst, sa + 1);
    case HIPR_BLEND_SRC_OUT: return da == 0 ? src : AlphaMulQ(src, 256 - da);
    case HIPR_BLEND_DST_OUT: return sa == 0 ? dst : AlphaMulQ(dst, 256 - sa);
    case HIPR_BLEND_SRC_ATOP: return AlphaMulQ(src, da + 1) + AlphaMulQ(dst, 256 - sa);
    case HIPR_BLEND_DST_ATOP: return AlphaMulQ(dst, sa + 1) + AlphaMulQ(src, 256 - da);
    case HIPR_BLEND_XOR: return AlphaMulQ(src, 256 - da) + AlphaMulQ(dst, 256 - sa);
    case HIPR_BLEND_PLUS:
      return Pack(min(Chan(src, 0) + Chan(dst, 0), 255u), min(Chan(src, 1) + Chan(dst, 1), 255u),
                  min(Chan(src, 2) + Chan(dst, 2), 255u), min(sa + da, 255u));
    case HIPR_BLEND_MODULATE: return PMColorMulWord(src, dst);
    case HIPR_BLEND_SCREEN: return src + dst - PMColorMulWord(src, dst);
    case HIPR_BLEND_SOFT_LIGHT: return SoftLightWord(src, dst);
    default: return dst;
  }
}
// BlendPixel: FastBlend on the stored pixel, else PorterDuffBlend with a
// target that is not premultiplied converted to and from premultiplied form.
__device__ uint32_t BlendStored(uint32_t src, uint32_t dst, int mode, bool premul) {
  const uint32_t a = src >> 24;
  switch (mode) {
    case HIPR_BLEND_CLEAR: return 0;
    case HIPR_BLEND_SRC: return src;
    case HIPR_BLEND_DST: return dst;
    case HIPR_BLEND_SRC_OVER:
      if (a == 0) return dst;
      if (a == 255) return src;
      break;
    case HIPR_BLEND_DST_OVER:
      if ((dst >> 24) == 255) return dst;
      break;
    case HIPR_BLEND_DST_IN:
      if (a == 255) return dst;
      if (a == 0) return 0;
      break;
    case HIPR_BLEND_DST_OUT:
      if (a == 0) return dst;
      if (a == 255) return 0;
      break;
    default:
      break;
  }
  uint32_t d = premul ? dst : PremulWord(dst);
  uint32_t r = PorterDuffWord(src, d, mode);
  return premul ? r : UnpremulWord(r);
}
__device__ __forceinline__ uint32_t SwapRB(uint32_t v) {
  return (v & 0xFF00FF00u) | ((v & 0xFFu) << 16) | ((v >> 16) & 0xFFu);
}

// RemapFloatTile (sw_span_brush.cc:21-32).
__device__ __forceinline__ float RemapTile(float t, int tile) {
  if (tile == HIPR_TILE_CLAMP) {
    t = fminf(fmaxf(t, 0.0f), 1.0f);
  } else if (tile == HIPR_TILE_REPEAT) {
    t = t - floorf(t);
  } else if (tile == HIPR_TILE_MIRROR) {
    float t1 = t - 1;
    float t2 = (float)((double)t1 - 2 * floor((double)t1 * 0.5) - 1);
    t = fabsf(t2);
  }
  return t;
}

// The stop table of the gradient being drawn, staged in LDS by the wave
// (16 offsets, 16 unpremultiplied colours).
struct GradStops {
  float offsets[16];
  float4 colors[16];
};

// GradientColorBrush::LerpColor + CalculateColor for linear / radial / sweep
// (sw_span_brush.cc:239-383); the stop scan reads the LDS copy `sl` of the
// gradient's table.
__device__ uint32_t GradientColor(const GradDev& g, const GradStops& sl, int x, int y) {
  float sx = x + 0.5f, sy = y + 0.5f;
  float mx = sx * g.sx + sy * g.kx + g.tx;
  float my = sx * g.ky + sy * g.sy + g.ty;
  float t;
  if (g.type == HIPR_GRADIENT_LINEAR) {
    t = mx;
  } else if (g.type == HIPR_GRADIENT_RADIAL) {
    t = sqrtf(mx * mx + my * my);
  } else if (g.type == HIPR_GRADIENT_CONICAL) {
    if (!ConicalT(g.con, mx, my, &t)) return 0;
  } else {
    float angle = atan2f(-my, -mx);
    const float k1Over2Pi = 0.1591549430918f;
    t = (float)(((double)(angle * k1Over2Pi) + 0.5 + (double)g.bias) * (double)g.scale);
  }
  if (fabsf(t) <= 1.f / 4096) {
    t = 0.0f;
  } else if (fabsf(t - 1.0f) <= 1.f / 4096) {
    t = 1.0f;
  }
  float c[4];
  if (g.tile == HIPR_TILE_DECAL && (t < 0.0f || t >= 1.0f)) return 0;
  t = RemapTile(t, g.tile);
  int n = g.count;
  bool stops = g.has_offsets != 0;
  float step = 1.f / (n - 1);
  if (stops && t <= sl.offsets[0]) {
    const float4 c0 = sl.colors[0];
    c[0] = c0.x, c[1] = c0.y, c[2] = c0.z, c[3] = c0.w;
    return PremulRGBA(c);
  }
  int i = 0;
  float start = 0.f, end = 0.f;
  int si = 0, ei = 1;
  for (i = 0; i < n - 1; i++) {
    if (stops) {
      start = sl.offsets[i];
      end = sl.offsets[i + 1];
    } else {
      start = step * i;
      end = step * (i + 1);
    }
    if (t >= start && t <= end) {
      si = i;
      ei = i + 1;
      break;
    }
  }
  if (i == n - 1) {
    const float4 cl = sl.colors[n - 1];
    c[0] = cl.x, c[1] = cl.y, c[2] = cl.z, c[3] = cl.w;
    return PremulRGBA(c);
  }
  float total = end - start, value = t - start;
  float mix = 0.5f;
  if (total > 0) mix = value / total;
  const float4 cs = sl.colors[si], ce = sl.colors[ei];
  c[0] = cs.x * (1 - mix) + ce.x * mix;
  c[1] = cs.y * (1 - mix) + ce.y * mix;
  c[2] = cs.z * (1 - mix) + ce.z * mix;
  c[3] = cs.w * (1 - mix) + ce.w * mix;
  return PremulRGBA(c);
}

// ColorFilter::FilterColor (color_filter.cc:113-185) of a premultiplied
// word: the blend filter's Porter-Duff blend of its colour over the source;
// the matrix filter on the unpremultiplied channels in integers,
// clamp(sum / 255 + offset); the sRGB filters through their tables
// (`srgb`: 256 entries linear -> sRGB, then 256 sRGB -> linear); compose
// passes the colour through.
__device__ uint32_t FilterWord(uint32_t c, const PaintDev& pd, const uint8_t* __restrict__ srgb) {
  switch (pd.cf_type) {
    case HIPR_COLOR_FILTER_BLEND:
      return PorterDuffWord(pd.cf_pm, c, pd.cf_mode);
    case HIPR_COLOR_FILTER_MATRIX: {
      const uint32_t u = UnpremulWord(c);
      const int32_t v[4] = {(int32_t)Chan(u, 0), (int32_t)Chan(u, 1), (int32_t)Chan(u, 2),
                            (int32_t)(u >> 24)};
      uint32_t o[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int32_t* m = pd.cf_m + 5 * i;
        int32_t mul = v[0] * m[0] + v[1] * m[1] + v[2] * m[2] + v[3] * m[3];
        int32_t r = mul / 255 + m[4];
        o[i] = (uint32_t)(r < 0 ? 0 : (r > 255 ? 255 : r));
      }
      return PremulWord(Pack(o[0], o[1], o[2], o[3]));
    }
    case HIPR_COLOR_FILTER_LINEAR_TO_SRGB:
    case HIPR_COLOR_FILTER_SRGB_TO_LINEAR: {
      const uint8_t* t = srgb + (pd.cf_type == HIPR_COLOR_FILTER_SRGB_TO_LINEAR ? 256 : 0);
      const uint32_t u = UnpremulWord(c);
      return PremulWord(Pack(t[Chan(u, 0)], t[Chan(u, 1)], t[Chan(u, 2)], u >> 24));
    }
    default:
      return c;
  }
}

struct CompositeParams {
  const Canvas* canvases;
  int32_t n_canvases;
  int32_t total_tiles;
  const Job* jobs;
  const JobCov* cov;
  const int64_t* map_base;
  const int32_t* map0;
  const int32_t* map1;
  const uint8_t* planes;
  const PaintDev* paints;
  const GradDev* grads;
  const uint8_t* srgb;  // the sRGB filter tables (512 entries)
  const ImageDev* images;       // pixmap shader images
  const uint8_t* image_pixels;
  // per job, the canvas tiles it touches: {tc0, tr0, tc1, tr1} (half-open;
  // empty when tc0 >= tc1), written by JobTileRangeKernel
  const int4* job_tiles;
  // optional: [0] += pixel blends performed, [1] += target pixels touched
  unsigned long long* counters;
};

// The tile range of every job on its canvas, once per frame: the composite
// scans these 16 bytes per job for each tile instead of the job and coverage
// records.  A rectangle job covers the tiles its rectangle meets; a walked
// job the tiles of its coverage bounds (as the walk wrote its maps).
__global__ void __launch_bounds__(256)
    JobTileRangeKernel(const Job* __restrict__ jobs, const JobCov* __restrict__ cov,
                       const PaintDev* __restrict__ paints, const Canvas* __restrict__ canvases,
                       int n_jobs, int4* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_jobs) return;
  const Job job = jobs[j];
  const Canvas& cv = canvases[job.canvas];
  int4 r = make_int4(0, 0, 0, 0);
  if (job.kind == kJobRect) {
    const PaintDev& pd = paints[job.paint];
    const int x0 = pd.L > 0 ? pd.L : 0, x1 = pd.R < cv.width ? pd.R : cv.width;
    const int y0 = pd.T > 0 ? pd.T : 0, y1 = pd.B < cv.height ? pd.B : cv.height;
    if (x0 < x1 && y0 < y1) r = make_int4(x0 >> 4, y0 >> 4, ((x1 - 1) >> 4) + 1, ((y1 - 1) >> 4) + 1);
  } else {
    const JobCov c = cov[j];
    PixBounds pb{c.left, c.top, c.right, c.bottom, !(c.left < c.right && c.top < c.bottom)};
    TileRange tr = JobTiles(pb, cv.width, cv.height);
    if (tr.ntc > 0) r = make_int4(tr.tc0, tr.tr0, tr.tc0 + tr.ntc, tr.tr0 + tr.ntr);
  }
  out[j] = r;
}

// SWSpanBrush::BrushH for one pixel and cover (sw_span_brush.cc:108-138):
// AlphaMulQ by the cover unless it is 255, then the target blend.
template <bool kFilters>
__device__ __forceinline__ uint32_t ApplyCover(uint32_t dst, uint32_t color, uint32_t cover,
                                               int mode, bool premul, const PaintDev& pd,
                                               const uint8_t* srgb) {
  uint32_t src = cover != 255 ? AlphaMulQ(color, cover) : color;
  if (kFilters && pd.cf_type != HIPR_COLOR_FILTER_NONE) src = FilterWord(src, pd, srgb);
  if (mode == HIPR_BLEND_SRC_OVER && premul) return BlendSrcOver(src, dst);
  return BlendStored(src, dst, mode, premul);
}

// PixmapBrush colour of a device pixel (sw_span_brush.cc:555-579): the
// pixel centre mapped to the layer's unit square, decal outside, nearest
// sample (BitmapSampler, bitmap_sampler.cc:26-40,85-108), then the
// Color4fFromColor -> Color4fToColor round trip per channel.
__device__ uint32_t ImageColor(const PaintDev& pd, const Canvas& layer, int x, int y) {
  float cx = x + 0.5f, cy = y + 0.5f;
  float u = cx * pd.im[0] + cy * pd.im[1] + pd.im[2];
  float v = cx * pd.im[3] + cy * pd.im[4] + pd.im[5];
  if (u < 0.0f || u >= 1.0f || v < 0.0f || v >= 1.0f) return 0;
  uint32_t ix = (uint32_t)(u * (float)(uint32_t)layer.width);
  uint32_t iy = (uint32_t)(v * (float)(uint32_t)layer.height);
  ix = ix > (uint32_t)layer.width - 1 ? (uint32_t)layer.width - 1 : ix;
  iy = iy > (uint32_t)layer.height - 1 ? (uint32_t)layer.height - 1 : iy;
  uint32_t s = ((const uint32_t*)layer.pixels)[(int64_t)iy * layer.width + ix];
  uint32_t o = 0;
#pragma unroll
  for (int ch = 0; ch < 4; ch++) {
    float f = (float)((s >> (8 * ch)) & 0xFF) / 255.f;
    o |= UnitToByte(f) << (8 * ch);
  }
  return o;
}

// BitmapSampler::GetColor (bitmap_sampler.cc:26-108) then
// PixmapBrush::CalculateColor (sw_span_brush.cc:569-579).  Float -> uint32
// conversions go through a 64-bit integer as on x86-64, so texel -1 of a
// bilinear footprint becomes 2^32 - 1 and is clamped to the last column.
__device__ __forceinline__ void SampleTexel(const uint32_t* __restrict__ px, uint32_t w,
                                            uint32_t h, float fx, float fy, float out[4]) {
  uint32_t ix = (uint32_t)(int64_t)fx, iy = (uint32_t)(int64_t)fy;
  ix = ix > w - 1 ? w - 1 : ix;
  iy = iy > h - 1 ? h - 1 : iy;
  const uint32_t c = px[(int64_t)iy * w + ix];
#pragma unroll
  for (int k = 0; k < 4; k++) out[k] = (float)((c >> (8 * k)) & 0xFF) / 255.f;
}

__device__ uint32_t PixmapColor(const PaintDev& pd, const ImageDev& im,
                                const uint8_t* __restrict__ pixels, int x, int y) {
  const float cx = x + 0.5f, cy = y + 0.5f;
  float u = cx * pd.im[0] + cy * pd.im[1] + pd.im[2];
  float v = cx * pd.im[3] + cy * pd.im[4] + pd.im[5];
  if ((pd.pix_tx == HIPR_TILE_DECAL && (u < 0.0f || u >= 1.0f)) ||
      (pd.pix_ty == HIPR_TILE_DECAL && (v < 0.0f || v >= 1.0f)))
    return 0;
  u = RemapTile(u, pd.pix_tx);
  v = RemapTile(v, pd.pix_ty);
  const uint32_t w = (uint32_t)im.width, h = (uint32_t)im.height;
  const uint32_t* px = (const uint32_t*)(pixels + im.offset);
  float c[4];
  if (pd.pix_filter == HIPR_FILTER_NEAREST) {
    SampleTexel(px, w, h, u * w, v * h, c);
  } else {
    const float fx = u * w, fy = v * h;
    float i0 = floorf(fx - 0.5f), j0 = floorf(fy - 0.5f);
    auto mod = [](float a, float b) { return a - b * floorf(a / b); };
    if (pd.pix_tx == HIPR_TILE_REPEAT) i0 = mod(i0, (float)w);
    if (pd.pix_ty == HIPR_TILE_REPEAT) j0 = mod(j0, (float)h);
    float i1 = i0 + 1.0f, j1 = j0 + 1.0f;
    if (pd.pix_tx == HIPR_TILE_REPEAT) i1 = mod(i1, (float)w);
    if (pd.pix_ty == HIPR_TILE_REPEAT) j1 = mod(j1, (float)h);
    const float a = (fx - 0.5f) - floorf(fx - 0.5f), b = (fy - 0.5f) - floorf(fy - 0.5f);
    float t00[4], t10[4], t01[4], t11[4];
    SampleTexel(px, w, h, i0, j0, t00);
    SampleTexel(px, w, h, i1, j0, t10);
    SampleTexel(px, w, h, i0, j1, t01);
    SampleTexel(px, w, h, i1, j1, t11);
    const float w00 = (1 - a) * (1 - b), w10 = a * (1 - b), w01 = (1 - a) * b, w11 = a * b;
#pragma unroll
    for (int k = 0; k < 4; k++)
      c[k] = ((w00 * t00[k] + w10 * t10[k]) + w01 * t01[k]) + w11 * t11[k];
  }
  uint32_t o = Pack(UnitToByte(c[0]), UnitToByte(c[1]), UnitToByte(c[2]), UnitToByte(c[3]));
  return pd.pix_unpremul ? PremulWord(o) : o;
}

// kRich: some draw of the frame has a colour filter or a pixmap shader; the
// frames without them run a variant that carries neither.
template <bool kRich>
__global__ void __launch_bounds__(64) CompositeKernel(CompositeParams P) {
  __shared__ GradStops stops;
  const int lane = Lane();
  int t = blockIdx.x;
  if (t >= P.total_tiles) return;
  // canvas lookup (canvases are few or sorted by tile_begin)
  int lo = 0, hi = P.n_canvases - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (P.canvases[mid].tile_begin <= t) {
      lo = mid;
    } else {
      hi = mid - 1;
    }
  }
  const Canvas cv = P.canvases[lo];
  const bool premul = (cv.format & kFormatUnpremul) == 0;
  const bool bgra = (cv.format & kFormatBGRA) != 0;
  int lt = t - cv.tile_begin;
  int tx = lt % cv.tiles_x, ty = lt / cv.tiles_x;
  int px0 = tx * kTile + 4 * (lane & 3);
  int py = ty * kTile + (lane >> 2);
  bool row_ok = py < cv.height;
  uint32_t pix[4];
  uint32_t* rowp = (uint32_t*)(cv.pixels + ((int64_t)py * cv.width) * 4);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    int x = px0 + q;
    pix[q] = (row_ok && x < cv.width) ? rowp[x] : 0;
    if (bgra) pix[q] = SwapRB(pix[q]);
  }
  int blends = 0;
  uint32_t touched = 0;  // bit q: pixel q blended at least once
  for (int jb = cv.job_begin; jb < cv.job_end; jb += kWave) {
    int j = jb + lane;
    int s0 = -1, s1 = -1;
    bool present = false;
    if (j < cv.job_end) {
      const int4 tb = P.job_tiles[j];
      if (tx >= tb.x && tx < tb.z && ty >= tb.y && ty < tb.w) {
        if (P.jobs[j].kind == kJobRect) {
          present = true;
        } else {
          int64_t mi = P.map_base[j] + (int64_t)(ty - tb.y) * (tb.z - tb.x) + (tx - tb.x);
          s0 = P.map0[mi];
          if (s0 >= 0) {
            s1 = P.map1[mi];
            present = true;
          }
        }
      }
    }
    uint64_t mask = __ballot(present);
    while (mask) {
      int b = __ffsll((unsigned long long)mask) - 1;
      mask &= mask - 1;
      int jj = jb + b;
      int q0 = __shfl(s0, b, kWave), q1 = __shfl(s1, b, kWave);
      const Job job = P.jobs[jj];
      // fields are read where used (the paint index is wave-uniform)
      const PaintDev& pd = P.paints[job.paint];
      const int mode = pd.blend;
      // covers of the four pixels: plane 0 and plane 1 of a walked job, or
      // 255 inside a rectangle job's rectangle
      // three-plane tiles: the first cover is blended where a direct span
      // lies (cover 0 included); otherwise where it is non-zero
      const bool three = job.kind != kJobRect && (job.flags & kJobKeepZero) != 0;
      uint32_t cov0, cov1 = 0, direct = 0;
      if (job.kind == kJobRect) {
        cov0 = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          int x = px0 + q;
          bool in = row_ok && x < cv.width && x >= pd.L && x < pd.R && py >= pd.T && py < pd.B;
          cov0 |= (in ? 0xFFu : 0u) << (8 * q);
        }
      } else {
        // the maps hold absolute plane numbers
        const uint32_t* pl0 = (const uint32_t*)(P.planes + (int64_t)q0 * kTilePlaneBytes) + lane;
        cov0 = pl0[0];
        if (three) {
          // direct covers (0 included), accumulated covers, and which
          // pixels have a direct span
          cov1 = pl0[kTilePlaneBytes / 4];
          direct = pl0[2 * kTilePlaneBytes / 4];
        } else if (q1 >= 0) {
          cov1 = *((const uint32_t*)(P.planes + (int64_t)q1 * kTilePlaneBytes) + lane);
        }
      }
      const uint32_t ga = (uint32_t)pd.ga;
      if (pd.kind == kPaintSolid && ga == 255) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
          uint32_t c0 = (cov0 >> (8 * q)) & 0xFF, c1 = (cov1 >> (8 * q)) & 0xFF;
          bool d0 = three ? ((direct >> (8 * q)) & 0xFF) != 0 : c0 != 0;
          if (d0) pix[q] = ApplyCover<kRich>(pix[q], pd.color, c0, mode, premul, pd, P.srgb);
          if (c1) pix[q] = ApplyCover<kRich>(pix[q], pd.color, c1, mode, premul, pd, P.srgb);
          blends += (int)d0 + (c1 != 0);
          touched |= (uint32_t)(d0 || c1 != 0) << q;
        }
      } else {
        const GradDev* g = pd.kind == kPaintGradient ? &P.grads[pd.grad] : nullptr;
        if (g) {
          // stage the stop table in LDS: one load per stop instead of a
          // global load per scan step of every pixel
          WaveSync();  // the previous gradient's reads are done
          if (lane < 16) {
            stops.offsets[lane] = g->offsets[lane];
            stops.colors[lane] = make_float4(g->colors[lane][0], g->colors[lane][1],
                                             g->colors[lane][2], g->colors[lane][3]);
          }
          WaveSync();
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
          uint32_t c0 = (cov0 >> (8 * q)) & 0xFF, c1 = (cov1 >> (8 * q)) & 0xFF;
          bool d0 = three ? ((direct >> (8 * q)) & 0xFF) != 0 : c0 != 0;
          if (d0 || c1) {
            uint32_t color =
                pd.kind == kPaintSolid    ? pd.color
                : g                       ? GradientColor(*g, stops, px0 + q, py)
                : kRich && pd.kind == kPaintPixmap
                    ? PixmapColor(pd, P.images[pd.pix_img], P.image_pixels, px0 + q, py)
                                          : ImageColor(pd, P.canvases[pd.layer], px0 + q, py);
            // Brush: cover & global alpha (sw_span_brush.cc:101)
            if (d0) pix[q] = ApplyCover<kRich>(pix[q], color, c0 & ga, mode, premul, pd, P.srgb);
            if (c1) pix[q] = ApplyCover<kRich>(pix[q], color, c1 & ga, mode, premul, pd, P.srgb);
            blends += (int)d0 + (c1 != 0);
            touched |= 1u << q;
          }
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 4; q++) {
    int x = px0 + q;
    if (row_ok && x < cv.width) rowp[x] = bgra ? SwapRB(pix[q]) : pix[q];
  }
  if (P.counters) {
    int nb = WaveSum(blends);
    int nt = WaveSum(__popc(touched));
    if (lane == 0 && (nb | nt)) {
      atomicAdd(&P.counters[0], (unsigned long long)nb);
      atomicAdd(&P.counters[1], (unsigned long long)nt);
    }
  }
}

}  // namespace hipr

#endif  // SKITY_HIP_COMPOSITE_HPP_
