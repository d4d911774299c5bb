// oracle_raster.cpp -- TEST INFRASTRUCTURE ONLY (parity checker).
//
// CPU restatement of skity's software fill rasteriser: Path iteration, the
// quad conversion of Stroke::QuadPath, fixed-point edge building and the
// analytic-AA ("AAA") scanline walk with its span builders.  Each function
// names the reference lines it follows.  Only the parity tests, smoke() and
// the CPU baseline of bench.py use this code.
#include "oracle_raster.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace orc {

namespace {

constexpr Fixed kFixed1 = 1 << 16;
constexpr Fixed kFixedHalf = 1 << 15;
constexpr int32_t kMaxS32 = INT32_MAX;
constexpr int32_t kMinS32 = -INT32_MAX;  // SW_MinS32 (sw_subpixel.hpp:36)
constexpr int32_t kNaN32 = INT32_MIN;
constexpr int kAccuracy = 2;  // SWEdge::kDefaultAccuracy (sw_edge.hpp:33)

// ---- 16.16 helpers: src/render/sw/sw_subpixel.hpp:38-92 -------------------
inline Fixed FixedMul(Fixed a, Fixed b) {
  return (Fixed)(((int64_t)a * b) >> 16);
}
inline Fixed LShift32(int32_t v, int s) { return (int32_t)((uint32_t)v << s); }
inline Fixed FDot6ToFixed(FDot6 x) { return LShift32(x, 10); }
inline Fixed FDot6ToFixedDiv2(FDot6 x) { return LShift32(x, 16 - 6 - 1); }
inline Fixed FixedDiv(int32_t numer, int32_t denom) {
  int64_t q = (int64_t)((uint64_t)(int64_t)numer << 16) / denom;
  if (q < kMinS32) q = kMinS32;
  if (q > kMaxS32) q = kMaxS32;
  return (Fixed)q;
}
inline Fixed FDot6Div(FDot6 a, FDot6 b) { return FixedDiv(a, b); }
inline FDot6 FixedToFDot6(Fixed x) { return x >> 10; }
inline Fixed FixedCeilToFixed(Fixed x) {
  return (Fixed)((uint32_t)(x + kFixed1 - 1) & 0xFFFF0000u);
}
inline Fixed FixedFloorToFixed(Fixed x) {
  return (Fixed)((uint32_t)x & 0xFFFF0000u);
}
inline Fixed FixedRoundToFixed(Fixed x) {
  return (Fixed)((uint32_t)(x + kFixedHalf) & 0xFFFF0000u);
}
inline Fixed IntToFixed(int n) { return (Fixed)((uint32_t)n << 16); }
inline int FixedRoundToInt(Fixed x) { return (x + kFixedHalf) >> 16; }
inline int FixedCeilToInt(Fixed x) { return (x + kFixed1 - 1) >> 16; }
inline int FixedFloorToInt(Fixed x) { return x >> 16; }

// SWEdge::SnapY (sw_edge.hpp:36-41): round to the nearest 1/4 pixel.
inline Fixed SnapY(Fixed y) {
  return (Fixed)((((uint32_t)y + (uint32_t)(kFixed1 >> (kAccuracy + 1))) >>
                  (16 - kAccuracy))
                 << (16 - kAccuracy));
}

// Float -> 16.16 as SetLine/CanBeIgnored do it (sw_edge.cc:24-29,76-81):
// times 4, to FDot6 (truncating), to Fixed, then back down by 2 bits.
inline Fixed ScalarToFixedAA(float v) {
  FDot6 d = (FDot6)((v * 4.0f) * 64.0f);
  return FDot6ToFixed(d) >> kAccuracy;
}

// CLZ (src/geometry/math.hpp:61-95), CLZ(0) == 32.
inline int Clz(uint32_t x) { return x == 0 ? 32 : __builtin_clz(x); }

typedef uint8_t Alpha;

// ---- Path::Iter with forceClose = false (src/graphic/path.cc:246-389) -----
struct Iter {
  const PathData& p;
  size_t vi = 0, pi = 0, wi = 0;
  Pt move_to{0, 0}, last_pt{0, 0};
  enum { kEmptyContour, kAfterMove, kAfterPrimitive } state = kEmptyContour;
  float weight = 1.f;
  explicit Iter(const PathData& path) : p(path) {}

  Pt ConsMoveTo() {
    if (state == kAfterMove) {
      state = kAfterPrimitive;
      return move_to;
    }
    return p.pts[pi - 1];
  }
  // AutoClose (path.cc:372-389).
  int AutoClose(Pt out[4]) {
    if (last_pt.x != move_to.x || last_pt.y != move_to.y) {
      if (std::isnan(last_pt.x) || std::isnan(last_pt.y) ||
          std::isnan(move_to.x) || std::isnan(move_to.y)) {
        return HIPR_VERB_CLOSE;
      }
      out[0] = last_pt;
      out[1] = move_to;
      last_pt = move_to;
      return HIPR_VERB_LINE;
    }
    out[0] = move_to;
    return HIPR_VERB_CLOSE;
  }
  // Returns a verb, or -1 for kDone.
  int Next(Pt out[4]) {
    if (vi == p.verbs.size()) return -1;
    int verb = p.verbs[vi++];
    switch (verb) {
      case HIPR_VERB_MOVE:
        if (vi == p.verbs.size()) return -1;
        move_to = p.pts[pi];
        out[0] = move_to;
        pi += 1;
        state = kAfterMove;
        last_pt = move_to;
        break;
      case HIPR_VERB_LINE:
        out[0] = ConsMoveTo();
        out[1] = p.pts[pi];
        last_pt = p.pts[pi];
        pi += 1;
        break;
      case HIPR_VERB_CONIC:
        weight = p.weights[wi++];
        [[fallthrough]];
      case HIPR_VERB_QUAD:
        out[0] = ConsMoveTo();
        out[1] = p.pts[pi];
        out[2] = p.pts[pi + 1];
        last_pt = p.pts[pi + 1];
        pi += 2;
        break;
      case HIPR_VERB_CUBIC:
        out[0] = ConsMoveTo();
        out[1] = p.pts[pi];
        out[2] = p.pts[pi + 1];
        out[3] = p.pts[pi + 2];
        last_pt = p.pts[pi + 2];
        pi += 3;
        break;
      case HIPR_VERB_CLOSE:
        verb = AutoClose(out);
        if (verb == HIPR_VERB_LINE) {
          vi--;
        } else {
          state = kEmptyContour;
        }
        last_pt = move_to;
        break;
      default:
        break;
    }
    return verb;
  }
};

// ---- Conic -> 2 quads: Conic::Chop + subdivided + ChopIntoQuadsPOW2(.., 1)
// (src/geometry/conic.cc:20-67,169-199,300-327).
inline bool Between(float a, float b, float c) { return (a - b) * (c - b) <= 0; }

void ConicChop(const Pt src[3], float w, Pt d0[3], Pt d1[3]) {
  float scale = 1.0f / (w + 1.0f);
  float wp1x = w * src[1].x, wp1y = w * src[1].y;
  float mx = ((src[0].x + (wp1x + wp1x)) + src[2].x) * scale * 0.5f;
  float my = ((src[0].y + (wp1y + wp1y)) + src[2].y) * scale * 0.5f;
  if (!std::isfinite(mx) || !std::isfinite(my)) {
    double wd = w, w2 = wd * 2, sh = 1 / (1 + wd) * 0.5;
    mx = (float)((src[0].x + w2 * src[1].x + src[2].x) * sh);
    my = (float)((src[0].y + w2 * src[1].y + src[2].y) * sh);
  }
  d0[0] = src[0];
  d0[1] = Pt{(src[0].x + wp1x) * scale, (src[0].y + wp1y) * scale};
  d0[2] = Pt{mx, my};
  d1[0] = Pt{mx, my};
  d1[1] = Pt{(wp1x + src[2].x) * scale, (wp1y + src[2].y) * scale};
  d1[2] = src[2];
}

void ConicToQuads(const Pt src[3], float w, Pt out[5]) {
  Pt d0[3], d1[3];
  ConicChop(src, w, d0, d1);
  float start_y = src[0].y, end_y = src[2].y;
  if (Between(start_y, src[1].y, end_y)) {
    float mid_y = d0[2].y;
    if (!Between(start_y, mid_y, end_y)) {
      float closer =
          std::fabs(mid_y - start_y) < std::fabs(mid_y - end_y) ? start_y : end_y;
      d0[2].y = d1[0].y = closer;
    }
    if (!Between(start_y, d0[1].y, d0[2].y)) d0[1].y = start_y;
    if (!Between(d1[0].y, d1[1].y, end_y)) d1[1].y = end_y;
  }
  out[0] = src[0];
  out[1] = d0[1];
  out[2] = d0[2];
  out[3] = d1[1];
  out[4] = d1[2];
  bool finite = true;
  for (int i = 0; i < 5; i++)
    finite = finite && std::isfinite(out[i].x) && std::isfinite(out[i].y);
  if (!finite) {
    for (int i = 1; i < 4; i++) out[i] = src[1];
  }
}

// ---- Cubic -> quads: Cubic::ToQuads / Subsegment (src/geometry/cubic.cc:
// 21-61) with CubicCoeff / QuadCoeff evaluation (src/geometry/geometry.cc:
// 44-67,135-169).
struct V2 {
  float x, y;
};
inline V2 operator+(V2 a, V2 b) { return {a.x + b.x, a.y + b.y}; }
inline V2 operator-(V2 a, V2 b) { return {a.x - b.x, a.y - b.y}; }
inline V2 operator*(V2 a, V2 b) { return {a.x * b.x, a.y * b.y}; }
inline V2 operator*(V2 a, float s) { return {a.x * s, a.y * s}; }

int CubicQuadCount(const Pt p[4]) {
  float accuracy = 0.1f;
  double max_hypot2 = 432.0 * accuracy * accuracy;
  V2 p1x2 = V2{p[1].x, p[1].y} * 3.f - V2{p[0].x, p[0].y};
  V2 p2x2 = V2{p[2].x, p[2].y} * 3.f - V2{p[3].x, p[3].y};
  V2 d = p2x2 - p1x2;
  float err = d.x * d.x + d.y * d.y;
  double qc = std::max(1., std::ceil(std::pow(err / max_hypot2, 1. / 6.0)));
  return (int)qc;
}

void CubicToQuads(const Pt p[4], std::vector<Pt>* quads /* 3 per quad */) {
  V2 P0{p[0].x, p[0].y}, P1{p[1].x, p[1].y}, P2{p[2].x, p[2].y},
      P3{p[3].x, p[3].y};
  V2 three{3, 3};
  // CubicCoeff
  V2 A = (P3 + three * (P1 - P2)) - P0;
  V2 B = three * ((P2 - (P1 + P1)) + P0);
  V2 C = three * (P1 - P0);
  V2 D = P0;
  // QuadCoeff of the tangents 3*(c1-p1), 3*(c2-c1), 3*(p2-c2)
  V2 T0{3.f * (P1.x - P0.x), 3.f * (P1.y - P0.y)};
  V2 T1{3.f * (P2.x - P1.x), 3.f * (P2.y - P1.y)};
  V2 T2{3.f * (P3.x - P2.x), 3.f * (P3.y - P2.y)};
  V2 qC = T0;
  V2 qB = (T1 - qC) + (T1 - qC);
  V2 qA = (T2 - (T1 + T1)) + qC;
  auto cubic_eval = [&](float t) {
    V2 tt{t, t};
    return ((A * tt + B) * tt + C) * tt + D;
  };
  auto quad_eval = [&](float t) {
    V2 tt{t, t};
    return (qA * tt + qB) * tt + qC;
  };
  int n = CubicQuadCount(p);
  double qcount = (double)n;
  for (int i = 0; i < n; i++) {
    float t1 = (float)(i / qcount);
    float t2 = (float)((i + 1) / qcount);
    V2 s0 = cubic_eval(t1);
    V2 s3 = cubic_eval(t2);
    float scale = (t2 - t1) * (1.f / 3.f);
    V2 s1 = s0 + quad_eval(t1) * scale;
    V2 s2 = s3 - quad_eval(t2) * scale;
    V2 a = s1 * 3.f - s0;
    V2 b = s2 * 3.f - s3;
    V2 ctrl{(a.x + b.x) / 4.f, (a.y + b.y) / 4.f};
    quads->push_back(Pt{s0.x, s0.y});
    quads->push_back(Pt{ctrl.x, ctrl.y});
    quads->push_back(Pt{s3.x, s3.y});
  }
}

// ---- ChopQuadAtYExtrema (src/geometry/geometry.cc:311-349) with
// valid_unit_divide (src/geometry/geometry.hpp:134-155).
int ValidUnitDivide(float number, float denom, float* ratio) {
  if (number < 0) {
    number = -number;
    denom = -denom;
  }
  if (denom == 0 || number == 0 || number >= denom) return 0;
  float r = number / denom;
  if (std::isnan(r)) return 0;
  if (r == 0) return 0;
  *ratio = r;
  return 1;
}

int IsNotMonotonic(float a, float b, float c) {
  float ab = a - b;
  float bc = b - c;
  if (ab < 0) bc = -bc;
  return ab == 0 || bc < 0;
}

int ChopQuadAtYExtrema(const Pt src[3], Pt dst[5]) {
  float a = src[0].y, b = src[1].y, c = src[2].y;
  if (IsNotMonotonic(a, b, c)) {
    float t;
    if (ValidUnitDivide(a - b, (a - b - b) + c, &t)) {
      // QuadCoeff::ChopQuadAt (geometry.cc:117-133), Interp = v0+(v1-v0)*t.
      auto interp = [](Pt p0, Pt p1, float tt) {
        return Pt{p0.x + (p1.x - p0.x) * tt, p0.y + (p1.y - p0.y) * tt};
      };
      Pt p01 = interp(src[0], src[1], t);
      Pt p12 = interp(src[1], src[2], t);
      dst[0] = src[0];
      dst[1] = p01;
      dst[2] = interp(p01, p12, t);
      dst[3] = p12;
      dst[4] = src[2];
      dst[1].y = dst[3].y = dst[2].y;
      return 1;
    }
    b = std::fabs(a - b) < std::fabs(b - c) ? a : c;
  }
  dst[0] = src[0];
  dst[1] = src[1];
  dst[2] = src[2];
  dst[1].y = b;
  return 0;
}

// ---- SWEdge / SWQuadEdge (src/render/sw/sw_edge.cc:18-292) -----------------
bool UpdateLine(Edge* e, Fixed x0, Fixed y0, Fixed x1, Fixed y1, Fixed slope) {
  if (y0 > y1) {
    std::swap(x0, x1);
    std::swap(y0, y1);
    e->winding = (int8_t)-e->winding;
  }
  FDot6 x0x1 = FixedToFDot6(x1 - x0);
  FDot6 y0y1 = FixedToFDot6(y1 - y0);
  if (y0y1 == 0) return false;
  e->x = x0;
  e->y = y0;
  e->dx = slope;
  e->dy = (x0x1 == 0 || slope == 0) ? kMaxS32 : std::abs(FDot6Div(y0y1, x0x1));
  e->upper_x = x0;
  e->upper_y = y0;
  e->lower_y = y1;
  return true;
}

bool SetLine(Edge* e, Pt p0, Pt p1) {
  Fixed x0 = ScalarToFixedAA(p0.x);
  Fixed y0 = SnapY(ScalarToFixedAA(p0.y));
  Fixed x1 = ScalarToFixedAA(p1.x);
  Fixed y1 = SnapY(ScalarToFixedAA(p1.y));
  e->winding = 1;
  FDot6 y0y1 = FixedToFDot6(y1 - y0);
  if (y0y1 == 0) return false;
  FDot6 x0x1 = FixedToFDot6(x1 - x0);
  Fixed slope = FDot6Div(x0x1, y0y1);
  e->curve_count = 0;
  e->curve_shift = 0;
  return UpdateLine(e, x0, y0, x1, y1, slope);
}

bool UpdateQuad(Edge* e) {
  int success = 0;
  int count = e->curve_count;
  Fixed oldx = e->qx, oldy = e->qy;
  Fixed dx = e->qdx, dy = e->qdy;
  Fixed newx = 0, newy = 0, nsx = 0, nsy = 0;
  int shift = e->curve_shift;
  do {
    Fixed slope;
    if (--count > 0) {
      newx = oldx + (dx >> shift);
      newy = oldy + (dy >> shift);
      if (std::abs(dy >> shift) >= kFixed1 * 2) {
        FDot6 diff_y = FixedToFDot6(newy - e->snapped_y);
        slope = diff_y ? FDot6Div(FixedToFDot6(newx - e->snapped_x), diff_y)
                       : kMaxS32;
        nsy = std::min<Fixed>(e->q_last_y, FixedRoundToFixed(newy));
        nsx = newx - FixedMul(slope, newy - nsy);
      } else {
        nsy = std::min(e->q_last_y, SnapY(newy));
        nsx = newx;
        FDot6 diff_y = FixedToFDot6(nsy - e->snapped_y);
        slope = diff_y ? FDot6Div(FixedToFDot6(newx - e->snapped_x), diff_y)
                       : kMaxS32;
      }
      dx += e->qddx;
      dy += e->qddy;
    } else {
      newx = e->q_last_x;
      newy = e->q_last_y;
      nsy = newy;
      nsx = newx;
      FDot6 diff_y = (newy - e->snapped_y) >> 10;
      slope = diff_y ? FDot6Div((newx - e->snapped_x) >> 10, diff_y) : kMaxS32;
    }
    if (slope < kMaxS32) {
      success = UpdateLine(e, e->snapped_x, e->snapped_y, nsx, nsy, slope);
    }
    oldx = newx;
    oldy = newy;
  } while (count > 0 && !success);
  e->qx = newx;
  e->qy = newy;
  e->qdx = dx;
  e->qdy = dy;
  e->snapped_x = nsx;
  e->snapped_y = nsy;
  e->curve_count = (int8_t)count;
  return success != 0;
}

int DiffToShift(FDot6 dx, FDot6 dy) {
  dx = std::abs(dx);
  dy = std::abs(dy);
  FDot6 dist = std::max(dx, dy) + (std::min(dx, dy) >> 1);
  dist = (dist + (1 << 4)) >> 5;
  return (32 - Clz((uint32_t)dist)) >> 1;
}

bool SetQuad(Edge* e, const Pt pts[3]) {
  float scale = (float)(1 << (kAccuracy + 6));
  FDot6 x0 = (int)(pts[0].x * scale), y0 = (int)(pts[0].y * scale);
  FDot6 x1 = (int)(pts[1].x * scale), y1 = (int)(pts[1].y * scale);
  FDot6 x2 = (int)(pts[2].x * scale), y2 = (int)(pts[2].y * scale);
  int w = 1;
  if (y0 > y2) {
    std::swap(x0, x2);
    std::swap(y0, y2);
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
  e->winding = (int8_t)w;
  e->curve_count = (int8_t)(1 << shift);
  e->curve_shift = (uint8_t)(shift - 1);
  Fixed A = FDot6ToFixedDiv2(x0 - x1 - x1 + x2);
  Fixed B = FDot6ToFixed(x1 - x0);
  e->qx = FDot6ToFixed(x0);
  e->qdx = B + (A >> shift);
  e->qddx = A >> (shift - 1);
  A = FDot6ToFixedDiv2(y0 - y1 - y1 + y2);
  B = FDot6ToFixed(y1 - y0);
  e->qy = FDot6ToFixed(y0);
  e->qdy = B + (A >> shift);
  e->qddy = A >> (shift - 1);
  e->q_last_x = FDot6ToFixed(x2);
  e->q_last_y = FDot6ToFixed(y2);
  e->qx >>= kAccuracy;
  e->qy >>= kAccuracy;
  e->qdx >>= kAccuracy;
  e->qdy >>= kAccuracy;
  e->qddx >>= kAccuracy;
  e->qddy >>= kAccuracy;
  e->q_last_x >>= kAccuracy;
  e->q_last_y >>= kAccuracy;
  e->qy = SnapY(e->qy);
  e->q_last_y = SnapY(e->q_last_y);
  e->q_first_y = e->qy;
  e->snapped_x = e->qx;
  e->snapped_y = e->qy;
  UpdateQuad(e);
  return true;
}

// CanBeIgnored (sw_edge.cc:72-88).
bool CanBeIgnored(float top, float bottom, Fixed y0, Fixed y1) {
  Fixed start_y = SnapY(ScalarToFixedAA(top));
  Fixed stop_y = SnapY(ScalarToFixedAA(bottom));
  return y0 >= stop_y || y1 <= start_y;
}

// ---- span builders (src/render/sw/sw_raster.cc:19-136) ---------------------
struct SpanSink {
  float scan_top;
  std::vector<Span> spans;
  // RealSpanBuilder::BuildSpan (sw_raster.cc:26-42)
  void Real(int x, int y, int width, int alpha, int accumulated = 0) {
    if (y < scan_top) return;
    spans.push_back(Span{x, y, width, alpha, accumulated});
  }
  void RealSpans(int x, int y, const Alpha* aa, int len) {
    if (y < scan_top) return;
    for (int i = 0; i < len; i++) Real(x + i, y, 1, aa[i]);
  }
  // SpanBuilder (sw_raster.cc:54-136)
  int left = 0;
  std::vector<Alpha> alphas;
  int curr_y = kNaN32;
  void Flush() {
    if (curr_y == kNaN32) return;
    size_t curr = 0, n = alphas.size();
    while (curr < n) {
      if (alphas[curr] > 0) {
        size_t start = curr, width = 1;
        Alpha a = alphas[curr];
        do {
          curr++;
          if (curr >= n) break;
          if (alphas[curr] == a) {
            width++;
          } else {
            break;
          }
        } while (alphas[curr]);
        Real(left + (int)start, curr_y, (int)width, a, 1);
      } else {
        curr++;
      }
    }
  }
  void FlushYIfNeeded(int y) {
    if (curr_y == y) return;
    if (curr_y == kNaN32) {
      curr_y = y;
      return;
    }
    Flush();
    std::fill(alphas.begin(), alphas.end(), 0);
    curr_y = y;
  }
  static Alpha SafeAdd(Alpha a, Alpha d) {
    return (uint16_t)a + (uint16_t)d > 255 ? 255 : (Alpha)(a + d);
  }
  void Acc(int x, int y, int width, Alpha a) {
    if (y < scan_top) return;
    FlushYIfNeeded(y);
    int off = x - left;
    for (int i = 0; i < width; i++) alphas[off + i] = SafeAdd(alphas[off + i], a);
  }
  void AccSpans(int x, int y, const Alpha* aa, int len) {
    if (y < scan_top) return;
    FlushYIfNeeded(y);
    int off = x - left;
    for (int i = 0; i < len; i++) alphas[off + i] = SafeAdd(alphas[off + i], aa[i]);
  }
};

// ---- trapezoid alpha helpers (sw_raster.cc:138-544) -------------------------
bool EdgesTooCloseE(const Edge* prev, const Edge* next, Fixed lower_y) {
  return next && prev && next->upper_y < lower_y &&
         prev->x + kFixed1 >= next->x - std::abs(next->dx);
}
bool EdgesTooCloseI(int prev_rite, Fixed ul, Fixed ll) {
  return prev_rite > FixedFloorToInt(ul) || prev_rite > FixedFloorToInt(ll);
}
inline Alpha PartialAlphaH(Alpha alpha, Fixed partial_height) {
  return (Alpha)FixedRoundToInt(alpha * partial_height);
}
inline Alpha PartialAlphaA(Alpha alpha, Alpha full) {
  return (Alpha)((alpha * full) >> 8);
}
inline Fixed ApproxIntersection(Fixed l1, Fixed r1, Fixed l2, Fixed r2) {
  if (l1 > r1) std::swap(l1, r1);
  if (l2 > r2) std::swap(l2, r2);
  return (std::max(l1, l2) + std::min(r1, r2)) / 2;
}
inline Alpha TrapezoidToAlpha(Fixed l1, Fixed l2) {
  Fixed area = (l1 + l2) / 2;
  return (Alpha)(area >> 8);
}
inline Alpha PartialTriangleToAlpha(Fixed a, Fixed b) {
  Fixed area = (a >> 11) * (a >> 11) * (b >> 11);
  return (Alpha)((area >> 8) & 0xFF);
}

struct Blitter {
  SpanSink* sink;
  void Single(int y, int x, Alpha alpha, Alpha full, bool no_real) {
    if (full == 0xFF && !no_real) {
      sink->Real(x, y, 1, alpha);
    } else {
      sink->Acc(x, y, 1, PartialAlphaA(alpha, full));
    }
  }
  void Two(int y, int x, Alpha a1, Alpha a2, Alpha full, bool no_real) {
    if (full == 0xFF && !no_real) {
      sink->Real(x, y, 1, a1);
      sink->Real(x + 1, y, 1, a2);
    } else {
      sink->Acc(x, y, 1, a1);
      sink->Acc(x + 1, y, 1, a2);
    }
  }
  void Full(int y, int x, int len, Alpha full, bool no_real) {
    if (full == 0xFF && !no_real) {
      sink->Real(x, y, len, full);
    } else {
      sink->Acc(x, y, len, full);
    }
  }
};

// compute_alpha_above_line / below_line (sw_raster.cc:312-368).
void AlphaAboveLine(Alpha* alphas, Fixed l, Fixed r, Fixed dY, Alpha full) {
  int R = FixedCeilToInt(r);
  if (R == 0) return;
  if (R == 1) {
    alphas[0] = PartialAlphaA((Alpha)(((R << 17) - l - r) >> 9), full);
    return;
  }
  Fixed first = kFixed1 - l;
  Fixed last = r - ((R - 1) << 16);
  Fixed first_h = FixedMul(first, dY);
  alphas[0] = (Alpha)(FixedMul(first, first_h) >> 9);
  Fixed alpha16 = first_h + (dY >> 1);
  for (int i = 1; i < R - 1; ++i) {
    alphas[i] = (Alpha)(alpha16 >> 8);
    alpha16 += dY;
  }
  alphas[R - 1] = (Alpha)(full - PartialTriangleToAlpha(last, dY));
}

void AlphaBelowLine(Alpha* alphas, Fixed l, Fixed r, Fixed dY, Alpha full) {
  int R = FixedCeilToInt(r);
  if (R == 0) return;
  if (R == 1) {
    alphas[0] = PartialAlphaA(TrapezoidToAlpha(l, r), full);
    return;
  }
  Fixed first = kFixed1 - l;
  Fixed last = r - ((R - 1) << 16);
  Fixed last_h = FixedMul(last, dY);
  alphas[R - 1] = (Alpha)(FixedMul(last, last_h) >> 9);
  Fixed alpha16 = last_h + (dY >> 1);
  for (int i = R - 2; i > 0; i--) {
    alphas[i] = (Alpha)((alpha16 >> 8) & 0xFF);
    alpha16 += dY;
  }
  alphas[0] = (Alpha)(full - PartialTriangleToAlpha(first, dY));
}

// blit_aaa_trapezoid_row (sw_raster.cc:370-455).
void BlitAAATrapezoidRow(Blitter* b, int y, Fixed ul, Fixed ur, Fixed ll,
                         Fixed lr, Fixed lDY, Fixed rDY, Alpha full,
                         bool no_real) {
  int L = FixedFloorToInt(ul), R = FixedCeilToInt(lr);
  int len = R - L;
  if (len == 1) {
    Alpha alpha = TrapezoidToAlpha(ur - ul, lr - ll);
    b->Single(y, L, alpha, full, no_real);
    return;
  }
  std::vector<Alpha> alphas(len + 1, full), temp(len + 1, 0);
  // The reference's scratch buffer is uninitialised; every temp entry read
  // below is written by AlphaBelow/AboveLine first.
  int uL = FixedFloorToInt(ul);
  int lL = FixedCeilToInt(ll);
  if (uL + 2 == lL) {
    Fixed first = IntToFixed(uL) + kFixed1 - ul;
    Fixed second = ll - ul - first;
    Alpha a1 = (Alpha)(full - PartialTriangleToAlpha(first, lDY));
    Alpha a2 = PartialTriangleToAlpha(second, lDY);
    alphas[0] = alphas[0] > a1 ? (Alpha)(alphas[0] - a1) : 0;
    alphas[1] = alphas[1] > a2 ? (Alpha)(alphas[1] - a2) : 0;
  } else {
    AlphaBelowLine(temp.data() + uL - L, ul - IntToFixed(uL),
                   ll - IntToFixed(uL), lDY, full);
    for (int i = uL; i < lL; ++i) {
      if (alphas[i - L] > temp[i - L]) {
        alphas[i - L] = (Alpha)(alphas[i - L] - temp[i - L]);
      } else {
        alphas[i - L] = 0;
      }
    }
  }
  int uR = FixedFloorToInt(ur);
  int lR = FixedCeilToInt(lr);
  if (uR + 2 == lR) {
    Fixed first = IntToFixed(uR) + kFixed1 - ur;
    Fixed second = lr - ur - first;
    Alpha a1 = PartialTriangleToAlpha(first, rDY);
    Alpha a2 = (Alpha)(full - PartialTriangleToAlpha(second, rDY));
    alphas[len - 2] = alphas[len - 2] > a1 ? (Alpha)(alphas[len - 2] - a1) : 0;
    alphas[len - 1] = alphas[len - 1] > a2 ? (Alpha)(alphas[len - 1] - a2) : 0;
  } else {
    AlphaAboveLine(temp.data() + uR - L, ur - IntToFixed(uR),
                   lr - IntToFixed(uR), rDY, full);
    for (int i = uR; i < lR; ++i) {
      if (alphas[i - L] > temp[i - L]) {
        alphas[i - L] = (Alpha)(alphas[i - L] - temp[i - L]);
      } else {
        alphas[i - L] = 0;
      }
    }
  }
  if (full == 0xFF && !no_real) {
    b->sink->RealSpans(L, y, alphas.data(), len);
  } else {
    b->sink->AccSpans(L, y, alphas.data(), len);
  }
}

// blit_trapezoid_row (sw_raster.cc:457-544).
void BlitTrapezoidRow(Blitter* b, int y, Fixed ul, Fixed ur, Fixed ll,
                      Fixed lr, Fixed lDY, Fixed rDY, Alpha full,
                      bool no_real) {
  if (ul > ur) return;
  if (ll > lr) ll = lr = ApproxIntersection(ul, ll, ur, lr);
  if (ul == ur && ll == lr) return;
  if (ul > ll) std::swap(ul, ll);
  if (ur > lr) std::swap(ur, lr);
  Fixed join_left = FixedCeilToFixed(ll);
  Fixed join_rite = FixedFloorToFixed(ur);
  if (join_left <= join_rite) {
    if (ul < join_left) {
      int len = FixedCeilToInt(join_left - ul);
      if (len == 1) {
        Alpha alpha = TrapezoidToAlpha(join_left - ul, join_left - ll);
        b->Single(y, ul >> 16, alpha, full, no_real);
      } else if (len == 2) {
        Fixed first = join_left - kFixed1 - ul;
        Fixed second = ll - ul - first;
        Alpha a1 = PartialTriangleToAlpha(first, lDY);
        Alpha a2 = (Alpha)(full - PartialTriangleToAlpha(second, lDY));
        b->Two(y, ul >> 16, a1, a2, full, no_real);
      } else {
        BlitAAATrapezoidRow(b, y, ul, join_left, ll, join_left, lDY, kMaxS32,
                            full, no_real);
      }
    }
    if (join_left < join_rite) {
      b->Full(y, FixedFloorToInt(join_left),
              FixedFloorToInt(join_rite - join_left), full, no_real);
    }
    if (lr > join_rite) {
      int len = FixedCeilToInt(lr - join_rite);
      if (len == 1) {
        Alpha alpha = TrapezoidToAlpha(ur - join_rite, lr - join_rite);
        b->Single(y, join_rite >> 16, alpha, full, no_real);
      } else if (len == 2) {
        Fixed first = join_rite + kFixed1 - ur;
        Fixed second = lr - ur - first;
        Alpha a1 = (Alpha)(full - PartialTriangleToAlpha(first, rDY));
        Alpha a2 = PartialTriangleToAlpha(second, rDY);
        b->Two(y, join_rite >> 16, a1, a2, full, no_real);
      } else {
        BlitAAATrapezoidRow(b, y, join_rite, ur, join_rite, lr, kMaxS32, rDY,
                            full, no_real);
      }
    }
  } else {
    BlitAAATrapezoidRow(b, y, ul, ur, ll, lr, lDY, rDY, full, no_real);
  }
}

// ---- active edge list (sw_raster.cc:159-247,546-729) -----------------------
// Edges live in a vector; index n is the head sentinel and n+1 the tail.
struct EdgeList {
  std::vector<Edge>& e;
  int head, tail;
  explicit EdgeList(std::vector<Edge>& edges) : e(edges) {}
  Edge* P(int i) { return i < 0 ? nullptr : &e[i]; }
  void Remove(int i) {
    e[e[i].prev].next = e[i].next;
    e[e[i].next].prev = e[i].prev;
  }
  void InsertAfter(int i, int after) {
    e[i].prev = after;
    e[i].next = e[after].next;
    e[e[after].next].prev = i;
    e[after].next = i;
  }
};

inline void UpdateNextNextY(Fixed y, Fixed next_y, Fixed* nny) {
  *nny = y > next_y && y < *nny ? y : *nny;
}

inline void CheckIntersection(EdgeList& L, int i, Fixed next_y, Fixed* nny) {
  const Edge& ed = L.e[i];
  const Edge& pv = L.e[ed.prev];
  if (pv.prev >= 0 && pv.x + pv.dx > ed.x + ed.dx) {
    *nny = next_y + (kFixed1 >> kAccuracy);
  }
}

void BackwardInsertEdgeBasedOnX(EdgeList& L, int i) {
  Fixed x = L.e[i].x;
  int prev = L.e[i].prev;
  while (L.e[prev].prev >= 0 && L.e[prev].x > x) prev = L.e[prev].prev;
  if (L.e[prev].next != i) {
    L.Remove(i);
    L.InsertAfter(i, prev);
  }
}

int BackwardInsertStart(EdgeList& L, int prev, Fixed x) {
  while (L.e[prev].prev >= 0 && L.e[prev].x > x) prev = L.e[prev].prev;
  return prev;
}

void InsertNewEdges(EdgeList& L, int ne, Fixed y, Fixed* nny) {
  if (L.e[ne].upper_y > y) {
    UpdateNextNextY(L.e[ne].upper_y, y, nny);
    return;
  }
  int prev = L.e[ne].prev;
  if (L.e[prev].x <= L.e[ne].x) {
    while (L.e[ne].upper_y <= y) {
      CheckIntersection(L, ne, y, nny);
      UpdateNextNextY(L.e[ne].lower_y, y, nny);
      ne = L.e[ne].next;
    }
    UpdateNextNextY(L.e[ne].upper_y, y, nny);
    return;
  }
  int start = BackwardInsertStart(L, prev, L.e[ne].x);
  do {
    int next = L.e[ne].next;
    bool moved = true;
    for (;;) {
      if (L.e[start].next == ne) {
        moved = false;
        break;
      }
      int after = L.e[start].next;
      if (L.e[after].x >= L.e[ne].x) break;
      start = after;
    }
    if (moved) {
      L.Remove(ne);
      L.InsertAfter(ne, start);
    }
    CheckIntersection(L, ne, y, nny);
    UpdateNextNextY(L.e[ne].lower_y, y, nny);
    start = ne;
    ne = next;
  } while (L.e[ne].upper_y <= y);
  UpdateNextNextY(L.e[ne].upper_y, y, nny);
}

inline void GoY(Edge* e, Fixed dst_y) {
  if (dst_y == e->y + kFixed1) {
    e->x = e->x + e->dx;
    e->y = dst_y;
  } else if (e->y != dst_y) {
    e->x = e->upper_x + FixedMul(e->dx, e->y - e->upper_y);
    e->y = dst_y;
  }
}
inline void GoYShift(Edge* e, Fixed dst_y, int y_shift) {
  e->y = dst_y;
  e->x += e->dx >> y_shift;
}

// WalkEdges (sw_raster.cc:546-677).
void WalkEdges(EdgeList& L, int fill, Blitter* b, int start_y, int stop_y,
               Fixed left_clip, Fixed right_clip) {
  Edge& H = L.e[L.head];
  Edge& T = L.e[L.tail];
  H.x = H.upper_x = left_clip;
  T.x = T.upper_x = right_clip;
  Fixed y = std::max(L.e[H.next].upper_y, IntToFixed(start_y));
  Fixed nny = kMaxS32;
  {
    int i;
    for (i = H.next; L.e[i].upper_y <= y; i = L.e[i].next) {
      GoY(&L.e[i], y);
      UpdateNextNextY(L.e[i].lower_y, y, &nny);
    }
    UpdateNextNextY(L.e[i].upper_y, y, &nny);
  }
  int winding_mask = fill == HIPR_FILL_EVEN_ODD ? 1 : -1;
  for (;;) {
    int w = 0;
    bool in_interval = false;
    Fixed prev_x = H.x;
    Fixed next_y = std::min(nny, FixedCeilToFixed(y + 1));
    int curr = H.next;
    int left_edge = L.head;
    Fixed left = left_clip;
    Fixed left_dy = 0;
    int prev_right = FixedFloorToInt(left_clip);
    nny = kMaxS32;
    int y_shift = 0;
    if ((next_y - y) & (kFixed1 >> 2)) {
      y_shift = 2;
      next_y = y + (kFixed1 >> 2);
    } else if ((next_y - y) & (kFixed1 >> 1)) {
      y_shift = 1;
    }
    Alpha full_alpha = PartialAlphaH(0xFF, next_y - y);
    while (L.e[curr].upper_y <= y) {
      Edge& ce = L.e[curr];
      w += ce.winding;
      bool prev_in = in_interval;
      in_interval = (w & winding_mask) != 0;
      bool is_left = in_interval && !prev_in;
      bool is_right = !in_interval && prev_in;
      if (is_left) {
        left = std::max(ce.x, left_clip);
        left_dy = ce.dy;
        left_edge = curr;
        GoYShift(&ce, next_y, y_shift);
      } else if (is_right) {
        Fixed right = std::min(right_clip, ce.x);
        GoYShift(&ce, next_y, y_shift);
        Fixed next_left = std::max(left_clip, L.e[left_edge].x);
        Fixed next_right = std::min(right_clip, ce.x);
        Fixed right_dy = ce.dy;
        bool no_real =
            full_alpha == 0xFF &&
            (EdgesTooCloseI(prev_right, left, L.e[left_edge].x) ||
             EdgesTooCloseE(&ce, L.P(ce.next), next_y));
        BlitTrapezoidRow(b, y >> 16, left, right, next_left, next_right,
                         left_dy, right_dy, full_alpha, no_real);
        prev_right = FixedCeilToInt(std::max(right, ce.x));
      } else {
        GoYShift(&ce, next_y, y_shift);
      }
      int next = ce.next;
      while (ce.lower_y <= next_y) {
        if (ce.curve_count > 0) {
          ce.snapped_x = ce.x;  // KeepContinuous (sw_edge.cc:294-297)
          ce.snapped_y = ce.y;
          if (!UpdateQuad(&ce)) break;
        } else {
          break;
        }
      }
      if (ce.lower_y <= next_y) {
        L.Remove(curr);
      } else {
        UpdateNextNextY(ce.lower_y, next_y, &nny);
        Fixed new_x = ce.x;
        if (new_x < prev_x) {
          BackwardInsertEdgeBasedOnX(L, curr);
        } else {
          prev_x = new_x;
        }
        CheckIntersection(L, curr, next_y, &nny);
      }
      curr = next;
    }
    if (in_interval) {
      const Edge& le = L.e[left_edge];
      bool no_real = full_alpha == 0xFF &&
                     EdgesTooCloseE(L.P(le.prev), &le, next_y);
      BlitTrapezoidRow(b, y >> 16, left, right_clip,
                       std::max(left_clip, le.x), right_clip, left_dy, 0,
                       full_alpha, no_real);
    }
    y = next_y;
    if (y >= IntToFixed(stop_y)) break;
    InsertNewEdges(L, curr, y, &nny);
  }
}

}  // namespace

int CubicQuadCountPublic(const Pt p[4]) { return CubicQuadCount(p); }

// ---- Path builder (src/graphic/path.cc:494-531,837-859,1327-1339) ----------
void PathData::MoveTo(Pt p) {
  if (!verbs.empty() && verbs.back() == HIPR_VERB_MOVE) {
    pts.back() = p;
  } else {
    last_move_index = (int)pts.size();
    has_last_move = true;
    verbs.push_back(HIPR_VERB_MOVE);
    pts.push_back(p);
  }
}
void PathData::InjectMoveToIfNeeded() {
  if (!has_last_move || last_move_index < 0) {
    Pt p{0, 0};
    if (!verbs.empty() && has_last_move) p = pts[~last_move_index];
    MoveTo(p);
  }
}
void PathData::LineTo(Pt p) {
  InjectMoveToIfNeeded();
  verbs.push_back(HIPR_VERB_LINE);
  pts.push_back(p);
}
void PathData::QuadTo(Pt p1, Pt p2) {
  InjectMoveToIfNeeded();
  verbs.push_back(HIPR_VERB_QUAD);
  pts.push_back(p1);
  pts.push_back(p2);
}
void PathData::Close() {
  if (!verbs.empty() && verbs.back() != HIPR_VERB_CLOSE) {
    verbs.push_back(HIPR_VERB_CLOSE);
  }
  if (has_last_move && last_move_index >= 0) last_move_index = ~last_move_index;
}

PathData ScenePath(const hipr_scene* s, int p) {
  PathData d;
  for (int i = s->verb_offsets[p]; i < s->verb_offsets[p + 1]; i++)
    d.verbs.push_back(s->verbs[i]);
  for (int i = s->point_offsets[p]; i < s->point_offsets[p + 1]; i++)
    d.pts.push_back(Pt{s->points[2 * i], s->points[2 * i + 1]});
  if (s->weights) {
    for (int i = s->weight_offsets[p]; i < s->weight_offsets[p + 1]; i++)
      d.weights.push_back(s->weights[i]);
  }
  d.fill = s->fill_types ? s->fill_types[p] : HIPR_FILL_WINDING;
  return d;
}

// Stroke::QuadPath (src/geometry/stroke.cc:914-962).
PathData QuadPath(const PathData& src) {
  PathData dst;
  dst.fill = src.fill;
  Iter it(src);
  Pt pts[4];
  std::vector<Pt> quads;
  for (;;) {
    int verb = it.Next(pts);
    if (verb < 0) break;
    switch (verb) {
      case HIPR_VERB_MOVE:
        dst.MoveTo(pts[0]);
        break;
      case HIPR_VERB_LINE:
        dst.LineTo(pts[1]);
        break;
      case HIPR_VERB_QUAD:
        dst.QuadTo(pts[1], pts[2]);
        break;
      case HIPR_VERB_CONIC: {
        Pt q[5];
        ConicToQuads(pts, it.weight, q);
        dst.QuadTo(q[1], q[2]);
        dst.QuadTo(q[3], q[4]);
      } break;
      case HIPR_VERB_CUBIC: {
        quads.clear();
        CubicToQuads(pts, &quads);
        for (size_t i = 0; i < quads.size(); i += 3)
          dst.QuadTo(quads[i + 1], quads[i + 2]);
      } break;
      case HIPR_VERB_CLOSE:
        dst.Close();
        break;
      default:
        break;
    }
  }
  return dst;
}

bool PathBounds(const PathData& p, float ltrb[4]) {
  size_t count = p.pts.size();
  if (p.verbs.size() > 1 && p.verbs.back() == HIPR_VERB_MOVE) count--;
  if (count == 0) {
    ltrb[0] = ltrb[1] = ltrb[2] = ltrb[3] = 0;
    return true;
  }
  float l = p.pts[0].x, t = p.pts[0].y, r = l, b = t;
  for (size_t i = 1; i < count; i++) {
    l = std::min(l, p.pts[i].x);
    r = std::max(r, p.pts[i].x);
    t = std::min(t, p.pts[i].y);
    b = std::max(b, p.pts[i].y);
  }
  ltrb[0] = l;
  ltrb[1] = t;
  ltrb[2] = r;
  ltrb[3] = b;
  return true;
}

// SWEdgeBuilder::BuildEdges with PathEdgeIter (src/render/sw/sw_edge.cc:
// 299-336; src/graphic/path_priv.hpp:75-167).
void BuildEdges(const PathData& q, float top, float bottom,
                std::vector<Edge>* edges) {
  auto add_line = [&](Pt a, Pt b) {
    Edge e{};
    if (SetLine(&e, a, b) && !CanBeIgnored(top, bottom, e.upper_y, e.lower_y)) {
      e.is_quad = false;
      edges->push_back(e);
    }
  };
  auto add_quad = [&](const Pt* p3) {
    Edge e{};
    if (SetQuad(&e, p3) &&
        !CanBeIgnored(top, bottom, e.q_first_y, e.q_last_y)) {
      e.is_quad = true;
      edges->push_back(e);
    }
  };
  size_t pi = 0, move_i = 0;
  bool needs_close = false;
  for (size_t vi = 0; vi < q.verbs.size(); vi++) {
    int v = q.verbs[vi];
    if (v == HIPR_VERB_MOVE) {
      if (needs_close) {
        add_line(q.pts[pi - 1], q.pts[move_i]);
        needs_close = false;
      }
      move_i = pi;
      pi++;
    } else if (v == HIPR_VERB_CLOSE) {
      if (needs_close) {
        add_line(q.pts[pi - 1], q.pts[move_i]);
        needs_close = false;
      }
    } else {
      int n = (v + 2) / 2;
      needs_close = true;
      const Pt* seg = &q.pts[pi - 1];
      pi += n;
      if (v == HIPR_VERB_LINE) {
        add_line(seg[0], seg[1]);
      } else if (v == HIPR_VERB_QUAD) {
        Pt mono[5];
        int k = ChopQuadAtYExtrema(seg, mono);
        for (int i = 0; i <= k; i++) add_quad(&mono[i * 2]);
      }
    }
  }
  if (needs_close) add_line(q.pts[pi - 1], q.pts[move_i]);
}

// SWRaster::RastePath (sw_raster.cc:731-786) with SortEdges / ProcessEdges
// (sw_raster.cc:679-729).  std::sort is not stable; equal (upper_y, x, dx)
// keys keep their emission order here.
void RastePath(const PathData& path, std::vector<Span>* spans,
               float out_bounds[4], std::vector<Edge>* edges_out, const float* m,
               const float* clip) {
  spans->clear();
  PathData q = QuadPath(path);
  if (m) {
    // Path::CopyWithMatrix: glm mat4 * vec4(x, y, 0, 1) sums the column
    // products pairwise, (m0 x + m1 y) + (m2 0 + m3 1), i.e.
    // (sx x + kx y) + (0 + tx) with every operation rounded separately.
    for (auto& pt : q.pts) {
      float x = pt.x, y = pt.y;
      pt.x = (m[0] * x + m[1] * y) + (0.f + m[2]);
      pt.y = (m[3] * x + m[4] * y) + (0.f + m[5]);
    }
  }
  float sb[4];
  PathBounds(q, sb);
  float bl = std::floor(sb[0]), bt = std::floor(sb[1]);
  float br = std::ceil(sb[2]), bb = std::ceil(sb[3]);
  if (out_bounds) {
    out_bounds[0] = bl;
    out_bounds[1] = bt;
    out_bounds[2] = br;
    out_bounds[3] = bb;
  }
  // Rect::Intersect with the clip bounds (rect.cc:160-172; by default the
  // cull rect +-1e9): on no overlap the scan bounds are SetEmpty.
  const float cull[4] = {-1e9f, -1e9f, 1e9f, 1e9f};
  const float* c = clip ? clip : cull;
  float il = std::max(sb[0], c[0]), ir = std::min(sb[2], c[2]);
  float it = std::max(sb[1], c[1]), ib = std::min(sb[3], c[3]);
  if (!(il < ir && it < ib)) il = it = ir = ib = 0;
  float sl = std::floor(il), st = std::floor(it), sr = std::ceil(ir),
        sbm = std::ceil(ib);
  if (!(sl < sr && st < sbm)) return;
  std::vector<Edge> edges;
  BuildEdges(q, st, sbm, &edges);
  if (edges_out) *edges_out = edges;
  if (edges.empty()) return;
  int n = (int)edges.size();
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b2) {
    const Edge& A = edges[a];
    const Edge& B = edges[b2];
    if (A.upper_y != B.upper_y) return A.upper_y < B.upper_y;
    if (A.x != B.x) return A.x < B.x;
    return A.dx < B.dx;
  });
  std::vector<Edge> list(n + 2);
  for (int i = 0; i < n; i++) list[i] = edges[order[i]];
  for (int i = 0; i < n; i++) {
    list[i].prev = i - 1;
    list[i].next = i + 1;
  }
  EdgeList L(list);
  L.head = n;
  L.tail = n + 1;
  Edge& H = list[n];
  Edge& T = list[n + 1];
  H = Edge{};
  T = Edge{};
  H.prev = -1;
  H.next = 0;
  H.upper_y = H.lower_y = kMinS32;
  H.x = kMinS32;
  H.dx = 0;
  H.dy = kMaxS32;
  H.upper_x = kMinS32;
  list[0].prev = n;
  T.prev = n - 1;
  T.next = -1;
  T.upper_y = T.lower_y = kMaxS32;
  T.x = kMaxS32;
  T.dx = 0;
  T.dy = kMaxS32;
  T.upper_x = kMaxS32;
  list[n - 1].next = n + 1;

  SpanSink sink;
  sink.scan_top = st;
  sink.left = (int)bl;
  sink.alphas.assign((size_t)(int)(br - bl), 0);
  Blitter blit{&sink};
  int start_y = (int)bt;
  int stop_y = (int)sbm;
  // static_cast<uint32_t>(float) << 16 (sw_raster.cc:779-780); x86-64
  // converts through a 64-bit integer, so negative bounds wrap like int.
  Fixed left_bound = (Fixed)((uint32_t)(int64_t)sl << 16);
  Fixed right_bound = (Fixed)((uint32_t)(int64_t)sr << 16);
  WalkEdges(L, path.fill, &blit, start_y, stop_y, left_bound, right_bound);
  sink.Flush();
  *spans = std::move(sink.spans);
}

}  // namespace orc
