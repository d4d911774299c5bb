// hipr_geom.hpp -- device-side geometry of the DrawPath path for gfx950.
//
// Fixed-point edge setup (SWEdge::SetLine, SWQuadEdge::SetQuad/UpdateQuad,
// src/render/sw/sw_edge.cc:18-292), curve-to-quad conversion
// (Cubic::ToQuads, src/geometry/cubic.cc:29-61; Conic::ChopIntoQuadsPOW2,
// src/geometry/conic.cc:20-67,169-199,300-327) and ChopQuadAtYExtrema
// (src/geometry/geometry.cc:311-349).  The float expressions are written in
// the reference's evaluation order and the library is compiled with
// -ffp-contract=off and IEEE division/sqrt, so every value rounds exactly as
// the reference's x86-64 build does.
#ifndef SKITY_HIP_GEOM_HPP_
#define SKITY_HIP_GEOM_HPP_

#include <hip/hip_runtime.h>
#include <stdint.h>

#define HD __host__ __device__ __forceinline__

namespace hipr {

typedef int32_t Fixed;
typedef int32_t FDot6;

constexpr Fixed kFixed1 = 1 << 16;
constexpr Fixed kFixedHalf = 1 << 15;
constexpr int32_t kMaxS32 = 0x7FFFFFFF;
constexpr int32_t kMinS32 = -0x7FFFFFFF;
constexpr int kAcc = 2;

// ---- edge record (16 x int32, 64 bytes) -----------------------------------
// 0 x, 1 snapped_x, 2 dx, 3 dy, 4 upper_x, 5 upper_y, 6 lower_y,
// 7 winding:8 | curve_count:8 | curve_shift:8 | is_quad:8,
// 8 qx, 9 qy, 10 qdx, 11 qdy, 12 qddx, 13 qddy, 14 q_last_x, 15 q_last_y.
// snapped_y of a quad edge always equals lower_y after UpdateQuad.
struct EdgeRec {
  Fixed x, snapped_x, dx, dy, upper_x, upper_y, lower_y;
  int32_t packed;
  Fixed qx, qy, qdx, qdy, qddx, qddy, q_last_x, q_last_y;
};
static_assert(sizeof(EdgeRec) == 64, "edge record is 64 bytes");

HD int EdgeWinding(int32_t packed) { return (int8_t)(packed & 0xFF); }
HD int EdgeCurveCount(int32_t packed) { return (int8_t)((packed >> 8) & 0xFF); }
HD int EdgeCurveShift(int32_t packed) { return (packed >> 16) & 0xFF; }
HD bool EdgeIsQuad(int32_t packed) { return ((packed >> 24) & 0xFF) != 0; }
HD int32_t PackEdge(int winding, int count, int shift, bool quad) {
  return (int32_t)((uint32_t)(uint8_t)(int8_t)winding |
                   ((uint32_t)(uint8_t)(int8_t)count << 8) |
                   ((uint32_t)(shift & 0xFF) << 16) | ((uint32_t)(quad ? 1 : 0) << 24));
}

// ---- 16.16 arithmetic (src/render/sw/sw_subpixel.hpp:38-92) ---------------
HD Fixed FixedMul(Fixed a, Fixed b) { return (Fixed)(((int64_t)a * b) >> 16); }
HD Fixed LShift32(int32_t v, int s) { return (int32_t)((uint32_t)v << s); }
HD Fixed FDot6ToFixed(FDot6 x) { return LShift32(x, 10); }
HD Fixed FDot6ToFixedDiv2(FDot6 x) { return LShift32(x, 16 - 6 - 1); }
HD Fixed FixedDiv(int32_t numer, int32_t denom) {
  int64_t q = (int64_t)((uint64_t)(int64_t)numer << 16) / denom;
  q = q < kMinS32 ? kMinS32 : q;
  q = q > kMaxS32 ? kMaxS32 : q;
  return (Fixed)q;
}
HD FDot6 FixedToFDot6(Fixed x) { return x >> 10; }
HD Fixed FixedCeilToFixed(Fixed x) {
  return (Fixed)((uint32_t)(x + kFixed1 - 1) & 0xFFFF0000u);
}
HD Fixed FixedFloorToFixed(Fixed x) { return (Fixed)((uint32_t)x & 0xFFFF0000u); }
HD Fixed FixedRoundToFixed(Fixed x) {
  return (Fixed)((uint32_t)(x + kFixedHalf) & 0xFFFF0000u);
}
HD Fixed IntToFixed(int n) { return (Fixed)((uint32_t)n << 16); }
HD int FixedRoundToInt(Fixed x) { return (x + kFixedHalf) >> 16; }
HD int FixedCeilToInt(Fixed x) { return (x + kFixed1 - 1) >> 16; }
HD int FixedFloorToInt(Fixed x) { return x >> 16; }
HD Fixed SnapY(Fixed y) {
  return (Fixed)((((uint32_t)y + (uint32_t)(kFixed1 >> (kAcc + 1))) >> (16 - kAcc))
                 << (16 - kAcc));
}
// float -> 16.16 through FDot6 of 4x the value (sw_edge.cc:24-29).
HD Fixed ScalarToFixedAA(float v) {
  FDot6 d = (FDot6)((v * 4.0f) * 64.0f);
  return FDot6ToFixed(d) >> kAcc;
}
HD int Clz32(uint32_t x) { return x == 0 ? 32 : __builtin_clz(x); }
HD int Abs32(int32_t v) { return v < 0 ? -v : v; }
HD int32_t Min32(int32_t a, int32_t b) { return a < b ? a : b; }
HD int32_t Max32(int32_t a, int32_t b) { return a > b ? a : b; }

// ---- SWEdge::UpdateLine / SetLine (sw_edge.cc:18-70) -----------------------
HD bool UpdateLine(EdgeRec* e, int* winding, Fixed x0, Fixed y0, Fixed x1,
                   Fixed y1, Fixed slope) {
  if (y0 > y1) {
    Fixed t = x0;
    x0 = x1;
    x1 = t;
    t = y0;
    y0 = y1;
    y1 = t;
    *winding = -*winding;
  }
  FDot6 x0x1 = FixedToFDot6(x1 - x0);
  FDot6 y0y1 = FixedToFDot6(y1 - y0);
  if (y0y1 == 0) return false;
  e->x = x0;
  e->dx = slope;
  e->dy = (x0x1 == 0 || slope == 0) ? kMaxS32 : Abs32(FixedDiv(y0y1, x0x1));
  e->upper_x = x0;
  e->upper_y = y0;
  e->lower_y = y1;
  return true;
}

struct P2 {
  float x, y;
};

HD bool SetLine(EdgeRec* e, P2 p0, P2 p1) {
  Fixed x0 = ScalarToFixedAA(p0.x);
  Fixed y0 = SnapY(ScalarToFixedAA(p0.y));
  Fixed x1 = ScalarToFixedAA(p1.x);
  Fixed y1 = SnapY(ScalarToFixedAA(p1.y));
  FDot6 y0y1 = FixedToFDot6(y1 - y0);
  if (y0y1 == 0) return false;
  FDot6 x0x1 = FixedToFDot6(x1 - x0);
  Fixed slope = FixedDiv(x0x1, y0y1);
  int w = 1;
  if (!UpdateLine(e, &w, x0, y0, x1, y1, slope)) return false;
  e->snapped_x = 0;
  e->packed = PackEdge(w, 0, 0, false);
  e->qx = e->qy = e->qdx = e->qdy = e->qddx = e->qddy = e->q_last_x = e->q_last_y = 0;
  return true;
}

// SWQuadEdge::UpdateQuad (sw_edge.cc:233-292).  `snapped_y` is passed in
// (it is KeepContinuous's y) and returned through lower_y.
HD bool UpdateQuad(EdgeRec* e, Fixed snapped_y) {
  int success = 0;
  int winding = EdgeWinding(e->packed);
  int count = EdgeCurveCount(e->packed);
  int shift = EdgeCurveShift(e->packed);
  Fixed oldx = e->qx, oldy = e->qy;
  Fixed dx = e->qdx, dy = e->qdy;
  Fixed newx = 0, newy = 0, nsx = 0, nsy = 0;
  Fixed snapped_x = e->snapped_x;
  do {
    Fixed slope;
    if (--count > 0) {
      newx = oldx + (dx >> shift);
      newy = oldy + (dy >> shift);
      if (Abs32(dy >> shift) >= kFixed1 * 2) {
        FDot6 diff_y = FixedToFDot6(newy - snapped_y);
        slope = diff_y ? FixedDiv(FixedToFDot6(newx - snapped_x), diff_y) : kMaxS32;
        nsy = Min32(e->q_last_y, FixedRoundToFixed(newy));
        nsx = newx - FixedMul(slope, newy - nsy);
      } else {
        nsy = Min32(e->q_last_y, SnapY(newy));
        nsx = newx;
        FDot6 diff_y = FixedToFDot6(nsy - snapped_y);
        slope = diff_y ? FixedDiv(FixedToFDot6(newx - snapped_x), diff_y) : kMaxS32;
      }
      dx += e->qddx;
      dy += e->qddy;
    } else {
      newx = e->q_last_x;
      newy = e->q_last_y;
      nsy = newy;
      nsx = newx;
      FDot6 diff_y = (newy - snapped_y) >> 10;
      slope = diff_y ? FixedDiv((newx - snapped_x) >> 10, diff_y) : kMaxS32;
    }
    if (slope < kMaxS32) {
      success = UpdateLine(e, &winding, snapped_x, snapped_y, nsx, nsy, slope);
    }
    oldx = newx;
    oldy = newy;
  } while (count > 0 && !success);
  e->qx = newx;
  e->qy = newy;
  e->qdx = dx;
  e->qdy = dy;
  e->snapped_x = nsx;
  // The reference also stores snapped_y = nsy here; the walk overwrites it
  // with the current y (KeepContinuous) before every later UpdateQuad, so it
  // is not kept.
  (void)nsy;
  e->packed = PackEdge(winding, count, shift, true);
  return success != 0;
}

HD int DiffToShift(FDot6 dx, FDot6 dy) {
  dx = Abs32(dx);
  dy = Abs32(dy);
  FDot6 dist = Max32(dx, dy) + (Min32(dx, dy) >> 1);
  dist = (dist + (1 << 4)) >> 5;
  return (32 - Clz32((uint32_t)dist)) >> 1;
}

// SWQuadEdge::SetQuad (sw_edge.cc:121-231).  Returns false for a quad whose
// ends round to the same quarter row.
HD bool SetQuad(EdgeRec* e, const P2 pts[3], Fixed* first_y_out = nullptr) {
  float scale = (float)(1 << (kAcc + 6));
  FDot6 x0 = (int)(pts[0].x * scale), y0 = (int)(pts[0].y * scale);
  FDot6 x1 = (int)(pts[1].x * scale), y1 = (int)(pts[1].y * scale);
  FDot6 x2 = (int)(pts[2].x * scale), y2 = (int)(pts[2].y * scale);
  int w = 1;
  if (y0 > y2) {
    FDot6 t = x0;
    x0 = x2;
    x2 = t;
    t = y0;
    y0 = y2;
    y2 = t;
    w = -1;
  }
  int top = (y0 + 32) >> 6, bottom = (y2 + 32) >> 6;
  if (top == bottom) return false;
  FDot6 dx = (LShift32(x1, 1) - x0 - x2) >> 2;
  FDot6 dy = (LShift32(y1, 1) - y0 - y2) >> 2;
  int shift = DiffToShift(dx, dy);
  if (shift == 0) {
    shift = 1;
  } else if (shift > 6) {
    shift = 6;
  }
  Fixed A = FDot6ToFixedDiv2(x0 - x1 - x1 + x2);
  Fixed B = FDot6ToFixed(x1 - x0);
  e->qx = FDot6ToFixed(x0) >> kAcc;
  e->qdx = (B + (A >> shift)) >> kAcc;
  e->qddx = (A >> (shift - 1)) >> kAcc;
  A = FDot6ToFixedDiv2(y0 - y1 - y1 + y2);
  B = FDot6ToFixed(y1 - y0);
  e->qy = SnapY(FDot6ToFixed(y0) >> kAcc);
  e->qdy = (B + (A >> shift)) >> kAcc;
  e->qddy = (A >> (shift - 1)) >> kAcc;
  e->q_last_x = FDot6ToFixed(x2) >> kAcc;
  e->q_last_y = SnapY(FDot6ToFixed(y2) >> kAcc);
  e->snapped_x = e->qx;
  // A freshly made SWQuadEdge is value-initialised (all zero) before
  // UpdateQuad fills the line part.
  e->x = e->dx = e->dy = e->upper_x = e->upper_y = e->lower_y = 0;
  e->packed = PackEdge(w, 1 << shift, shift - 1, true);
  Fixed first_y = e->qy;
  if (first_y_out) *first_y_out = first_y;  // SWQuadEdge::q_first_y
  UpdateQuad(e, first_y);
  return true;
}

// ---- Conic -> 2 quads ------------------------------------------------------
HD bool Between(float a, float b, float c) { return (a - b) * (c - b) <= 0; }
HD bool Finite(float v) { return !(__builtin_isinf(v) || __builtin_isnan(v)); }

HD void ConicToQuads(const P2 src[3], float w, P2 out[5]) {
  float scale = 1.0f / (w + 1.0f);
  float wp1x = w * src[1].x, wp1y = w * src[1].y;
  float mx = ((src[0].x + (wp1x + wp1x)) + src[2].x) * scale * 0.5f;
  float my = ((src[0].y + (wp1y + wp1y)) + src[2].y) * scale * 0.5f;
  if (!Finite(mx) || !Finite(my)) {
    double wd = w, w2 = wd * 2, sh = 1 / (1 + wd) * 0.5;
    mx = (float)((src[0].x + w2 * src[1].x + src[2].x) * sh);
    my = (float)((src[0].y + w2 * src[1].y + src[2].y) * sh);
  }
  P2 d01{(src[0].x + wp1x) * scale, (src[0].y + wp1y) * scale};
  P2 d02{mx, my};
  P2 d11{(wp1x + src[2].x) * scale, (wp1y + src[2].y) * scale};
  float start_y = src[0].y, end_y = src[2].y;
  if (Between(start_y, src[1].y, end_y)) {
    float mid_y = d02.y;
    if (!Between(start_y, mid_y, end_y)) {
      float closer = fabsf(mid_y - start_y) < fabsf(mid_y - end_y) ? start_y : end_y;
      d02.y = closer;
    }
    if (!Between(start_y, d01.y, d02.y)) d01.y = start_y;
    if (!Between(d02.y, d11.y, end_y)) d11.y = end_y;
  }
  out[0] = src[0];
  out[1] = d01;
  out[2] = d02;
  out[3] = d11;
  out[4] = src[2];
  bool fin = true;
  for (int i = 0; i < 5; i++) fin = fin && Finite(out[i].x) && Finite(out[i].y);
  if (!fin) {
    for (int i = 1; i < 4; i++) out[i] = src[1];
  }
}

// ---- Cubic -> quads ----------------------------------------------------------
struct CubicQ {
  P2 A, B, C, D;     // CubicCoeff
  P2 qA, qB, qC;     // QuadCoeff of the tangents
  int n;             // number of quads
};

HD int CubicQuadCount(const P2 p[4]) {
  float accuracy = 0.1f;
  double max_hypot2 = 432.0 * accuracy * accuracy;
  float ax = p[1].x * 3.f - p[0].x, ay = p[1].y * 3.f - p[0].y;
  float bx = p[2].x * 3.f - p[3].x, by = p[2].y * 3.f - p[3].y;
  float dx = bx - ax, dy = by - ay;
  float err = dx * dx + dy * dy;
  double q = ceil(pow((double)err / max_hypot2, 1. / 6.0));
  return q < 1. ? 1 : (int)q;
}

HD void CubicSetup(const P2 p[4], CubicQ* c) {
  P2 P0 = p[0], P1 = p[1], P2_ = p[2], P3 = p[3];
  c->A = P2{(P3.x + 3.f * (P1.x - P2_.x)) - P0.x, (P3.y + 3.f * (P1.y - P2_.y)) - P0.y};
  c->B = P2{3.f * ((P2_.x - (P1.x + P1.x)) + P0.x), 3.f * ((P2_.y - (P1.y + P1.y)) + P0.y)};
  c->C = P2{3.f * (P1.x - P0.x), 3.f * (P1.y - P0.y)};
  c->D = P0;
  P2 T0{3.f * (P1.x - P0.x), 3.f * (P1.y - P0.y)};
  P2 T1{3.f * (P2_.x - P1.x), 3.f * (P2_.y - P1.y)};
  P2 T2{3.f * (P3.x - P2_.x), 3.f * (P3.y - P2_.y)};
  c->qC = T0;
  c->qB = P2{(T1.x - T0.x) + (T1.x - T0.x), (T1.y - T0.y) + (T1.y - T0.y)};
  c->qA = P2{(T2.x - (T1.x + T1.x)) + T0.x, (T2.y - (T1.y + T1.y)) + T0.y};
  c->n = CubicQuadCount(p);
}

HD P2 CubicEval(const CubicQ& c, float t) {
  return P2{((c.A.x * t + c.B.x) * t + c.C.x) * t + c.D.x,
            ((c.A.y * t + c.B.y) * t + c.C.y) * t + c.D.y};
}
HD P2 CubicTangentEval(const CubicQ& c, float t) {
  return P2{(c.qA.x * t + c.qB.x) * t + c.qC.x, (c.qA.y * t + c.qB.y) * t + c.qC.y};
}
// Quad i of the cubic: control point and end point (Subsegment, cubic.cc:55-61).
HD void CubicQuad(const CubicQ& c, int i, P2* ctrl, P2* end) {
  double qc = (double)c.n;
  float t1 = (float)(i / qc);
  float t2 = (float)((i + 1) / qc);
  P2 s0 = CubicEval(c, t1);
  P2 s3 = CubicEval(c, t2);
  float scale = (t2 - t1) * (1.f / 3.f);
  P2 d1 = CubicTangentEval(c, t1), d2 = CubicTangentEval(c, t2);
  P2 s1{s0.x + d1.x * scale, s0.y + d1.y * scale};
  P2 s2{s3.x - d2.x * scale, s3.y - d2.y * scale};
  float ax = s1.x * 3.f - s0.x, ay = s1.y * 3.f - s0.y;
  float bx = s2.x * 3.f - s3.x, by = s2.y * 3.f - s3.y;
  *ctrl = P2{(ax + bx) / 4.f, (ay + by) / 4.f};
  *end = s3;
}

// ---- ChopQuadAtYExtrema ------------------------------------------------------
HD int ValidUnitDivide(float number, float denom, float* ratio) {
  if (number < 0) {
    number = -number;
    denom = -denom;
  }
  if (denom == 0 || number == 0 || number >= denom) return 0;
  float r = number / denom;
  if (__builtin_isnan(r)) return 0;
  if (r == 0) return 0;
  *ratio = r;
  return 1;
}

HD int ChopQuadAtYExtrema(const P2 src[3], P2 dst[5]) {
  float a = src[0].y, b = src[1].y, c = src[2].y;
  float ab = a - b, bc = b - c;
  if (ab < 0) bc = -bc;
  if (ab == 0 || bc < 0) {
    float t;
    if (ValidUnitDivide(a - b, (a - b - b) + c, &t)) {
      P2 p01{src[0].x + (src[1].x - src[0].x) * t, src[0].y + (src[1].y - src[0].y) * t};
      P2 p12{src[1].x + (src[2].x - src[1].x) * t, src[1].y + (src[2].y - src[1].y) * t};
      P2 m{p01.x + (p12.x - p01.x) * t, p01.y + (p12.y - p01.y) * t};
      dst[0] = src[0];
      dst[1] = P2{p01.x, m.y};
      dst[2] = m;
      dst[3] = P2{p12.x, m.y};
      dst[4] = src[2];
      return 1;
    }
    b = fabsf(a - b) < fabsf(b - c) ? a : c;
  }
  dst[0] = src[0];
  dst[1] = P2{src[1].x, b};
  dst[2] = src[2];
  return 0;
}

}  // namespace hipr

#endif  // SKITY_HIP_GEOM_HPP_
