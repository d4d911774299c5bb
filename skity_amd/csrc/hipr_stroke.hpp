// hipr_stroke.hpp -- stage (a): stroke outlines on the device.
//
// Stroke::StrokePath / PathStroker (src/geometry/stroke.cc:15-901) run by one
// wavefront per stroked draw over the quad path of its source
// (Stroke::QuadPath, stroke.cc:914-962, produced on the fly); the lanes
// stroke chunks of quad segments ahead of the stroker.  The recursion
// of PathStroker::QuadStroke (depth <= 33) is an explicit stack; caps and
// joins, the unit-arc conics of round joins (Conic::BuildUnitArc,
// conic.cc:72-167) and the reversal of the inner offset path
// (Path::ReversePathTo, path.cc:1042-1080) follow the reference's float
// expressions.  The outline is written as a skity Path (verbs, points, conic
// weights).  mode 0 only counts (exact sizes for the write pass); mode 1
// writes, with the inner offset path of each contour kept in a scratch
// region until it is reversed onto the outline.
#ifndef SKITY_HIP_STROKE_HPP_
#define SKITY_HIP_STROKE_HPP_

#include "hipr_flatten.hpp"

// The stroker is host-and-device code: the library runs it on the device;
// a host build of the same functions checks it against the oracle on the CPU.
#define HIPR_SD __host__ __device__ __attribute__((always_inline))
// Inlined into the stroke kernel so the stroker's state stays in registers
// (a called function would keep it in scratch memory).
#define HIPR_SDI __host__ __device__ __attribute__((always_inline)) inline

namespace hipr {

namespace stroke {

HIPR_SD inline int Max(int a, int b) { return a > b ? a : b; }

constexpr float kNearlyZero = 1.f / 4096;
constexpr float kRoot2Over2 = 0.707106781f;
constexpr float kSqrt2 = 1.41421356f;
constexpr int kRecursiveLimits = 11 * 3;

struct V {
  float x, y;
};
HIPR_SD inline V Add(V a, V b) { return V{a.x + b.x, a.y + b.y}; }
HIPR_SD inline V Sub(V a, V b) { return V{a.x - b.x, a.y - b.y}; }
HIPR_SD inline V Mul(V a, float s) { return V{a.x * s, a.y * s}; }
HIPR_SD inline bool Eq(V a, V b) { return a.x == b.x && a.y == b.y; }
HIPR_SD inline float Dot(V a, V b) { return a.x * b.x + a.y * b.y; }
HIPR_SD inline float Cross(V a, V b) { return a.x * b.y - a.y * b.x; }
HIPR_SD inline bool NearlyZero(float x, float tol = kNearlyZero) {
  return fabsf(x) <= tol;
}
HIPR_SD inline bool IsFin(float v) { return !(__builtin_isinf(v) || __builtin_isnan(v)); }

// PointSetLength<false> (src/geometry/point_priv.hpp)
HIPR_SD bool SetLength(V* pt, float x, float y, float length) {
  double xx = x, yy = y;
  double dmag = sqrt(xx * xx + yy * yy);
  double dscale = (double)length / dmag;
  x = (float)(x * dscale);
  y = (float)(y * dscale);
  if (!IsFin(x) || !IsFin(y) || (x == 0 && y == 0)) {
    pt->x = 0;
    pt->y = 0;
    return false;
  }
  pt->x = x;
  pt->y = y;
  return true;
}
HIPR_SD inline V RotCCW(V s) { return V{s.y, -s.x}; }
HIPR_SD inline V RotCW(V s) { return V{-s.y, s.x}; }
HIPR_SD inline float DistSqd(V a, V b) {
  float dx = a.x - b.x, dy = a.y - b.y;
  return dx * dx + dy * dy;
}
HIPR_SD inline bool CanNorm(float dx, float dy) {
  return IsFin(dx) && IsFin(dy) && (dx != 0 || dy != 0);
}

HIPR_SD V QuadEval(const V q[3], float t) {
  V C = q[0];
  V B = Add(Sub(q[1], C), Sub(q[1], C));
  V A = Add(Sub(q[2], Add(q[1], q[1])), C);
  return V{(A.x * t + B.x) * t + C.x, (A.y * t + B.y) * t + C.y};
}
HIPR_SD V QuadTangent(const V q[3], float t) {
  if ((t == 0 && Eq(q[0], q[1])) || (t == 1 && Eq(q[1], q[2]))) return Sub(q[2], q[0]);
  V B = Sub(q[1], q[0]);
  V A = Sub(Sub(q[2], q[1]), B);
  V T{A.x * t + B.x, A.y * t + B.y};
  return Add(T, T);
}
HIPR_SD float PtToLine(V pt, V a, V b) {
  V dxy = Sub(b, a), ab0 = Sub(pt, a);
  float number = Dot(dxy, ab0), denom = Dot(dxy, dxy);
  float t = number / denom;
  if (t >= 0 && t <= 1) {
    V hit{a.x * (1 - t) + b.x * t, a.y * (1 - t) + b.y * t};
    return DistSqd(hit, pt);
  }
  return DistSqd(pt, a);
}
HIPR_SD int FindUnitQuadRoots(float A, float B, float C, float roots[2]) {
  if (A == 0) return ValidUnitDivide(-C, B, roots);
  int n = 0;
  double dr = (double)B * B - 4 * (double)A * C;
  if (dr < 0) return 0;
  dr = sqrt(dr);
  float R = (float)dr;
  if (__builtin_isinf(R)) return 0;
  float Q = (B < 0) ? -(B - R) / 2 : -(B + R) / 2;
  n += ValidUnitDivide(Q, A, &roots[n]);
  n += ValidUnitDivide(C, Q, &roots[n]);
  if (n == 2) {
    if (roots[0] > roots[1]) {
      float t = roots[0];
      roots[0] = roots[1];
      roots[1] = t;
    } else if (roots[0] == roots[1]) {
      n -= 1;
    }
  }
  return n;
}

// A Path builder over a flat output region (counting past the capacity).
struct PathW {
  uint8_t* verbs;
  float* pts;
  float* w;
  int cap_v, cap_p, cap_w;
  int nv, np, nw;
  int last_verb;      // -1: empty
  V last_pt;
  V move_pt;          // last MoveTo point (for an injected move)
  bool move_open;     // a contour is open (Close not yet seen)
  bool any_move;
  HIPR_SD void Init(uint8_t* v, float* p, float* ww, int cv, int cp, int cw) {
    verbs = v;
    pts = p;
    w = ww;
    cap_v = cv;
    cap_p = cp;
    cap_w = cw;
    Reset();
  }
  HIPR_SD void Reset() {
    nv = np = nw = 0;
    last_verb = -1;
    last_pt = V{0, 0};
    move_pt = V{0, 0};
    move_open = false;
    any_move = false;
  }
  HIPR_SD void PutVerb(int v) {
    if (verbs && nv < cap_v) verbs[nv] = (uint8_t)v;
    nv++;
    last_verb = v;
  }
  HIPR_SD void PutPt(V p) {
    if (pts && np < cap_p) {
      pts[2 * np] = p.x;
      pts[2 * np + 1] = p.y;
    }
    np++;
    last_pt = p;
  }
  HIPR_SD void PutW(float x) {
    if (w && nw < cap_w) w[nw] = x;
    nw++;
  }
  HIPR_SD void MoveTo(V p) {
    if (last_verb == HIPR_VERB_MOVE) {
      if (pts && np - 1 < cap_p) {
        pts[2 * (np - 1)] = p.x;
        pts[2 * (np - 1) + 1] = p.y;
      }
      last_pt = p;
    } else {
      PutVerb(HIPR_VERB_MOVE);
      PutPt(p);
    }
    move_pt = p;
    move_open = true;
    any_move = true;
  }
  HIPR_SD void Inject() {
    if (!move_open) MoveTo(any_move ? move_pt : V{0, 0});
  }
  HIPR_SD void LineTo(V p) {
    Inject();
    PutVerb(HIPR_VERB_LINE);
    PutPt(p);
  }
  HIPR_SD void QuadTo(V c, V p) {
    Inject();
    PutVerb(HIPR_VERB_QUAD);
    PutPt(c);
    PutPt(p);
  }
  HIPR_SD void ConicTo(V c, V p, float wt) {
    if (!(wt > 0)) {
      LineTo(p);
    } else if (__builtin_isinf(wt)) {
      LineTo(c);
      LineTo(p);
    } else if (wt == 1.f) {
      QuadTo(c, p);
    } else {
      Inject();
      PutVerb(HIPR_VERB_CONIC);
      PutPt(c);
      PutPt(p);
      PutW(wt);
    }
  }
  HIPR_SD void Close() {
    if (nv > 0 && last_verb != HIPR_VERB_CLOSE) PutVerb(HIPR_VERB_CLOSE);
    move_open = false;
  }
};

struct M3 {
  float a, b, tx, c, d, ty;
};
HIPR_SD bool NearIdentity(const M3& m) {
  const float e = 1.1920929e-07f;
  return fabsf(m.a - 1) <= e && fabsf(m.d - 1) <= e && fabsf(m.b) <= e && fabsf(m.c) <= e &&
         fabsf(m.tx) <= e && fabsf(m.ty) <= e;
}
HIPR_SD M3 MMul(const M3& p, const M3& q) {
  if (NearIdentity(p)) return q;
  if (NearIdentity(q)) return p;
  M3 r;
  r.a = p.a * q.a + p.b * q.c;
  r.b = p.a * q.b + p.b * q.d;
  r.tx = (p.a * q.tx + p.b * q.ty) + p.tx;
  r.c = p.c * q.a + p.d * q.c;
  r.d = p.c * q.b + p.d * q.d;
  r.ty = (p.c * q.tx + p.d * q.ty) + p.ty;
  return r;
}
HIPR_SD inline V MApply(const M3& m, V p) {
  return V{(m.a * p.x + m.b * p.y) + m.tx, (m.c * p.x + m.d * p.y) + m.ty};
}

struct QC {
  V quad[3];
  V tan_start, tan_end;
  float start_t, mid_t, end_t;
  bool start_set, end_set;
  HIPR_SD bool Init(float s, float e) {
    start_t = s;
    mid_t = (s + e) * 0.5f;
    end_t = e;
    start_set = end_set = false;
    return start_t < mid_t && mid_t < end_t;
  }
  HIPR_SD void InitWithStart(const QC& p) {
    Init(p.start_t, p.mid_t);
    quad[0] = p.quad[0];
    tan_start = p.tan_start;
    start_set = true;
  }
  HIPR_SD void InitWithEnd(const QC& p) {
    Init(p.mid_t, p.end_t);
    quad[2] = p.quad[2];
    tan_end = p.tan_end;
    end_set = true;
  }
};

// One quad segment's outer and inner offset curves, stroked ahead of the
// stroker from a presumed start point: n[side] quads of 4 floats (control
// point, end point) at q[side].  ok = 0 when a side needed more recursion
// levels or quads than were kept.
struct SpecView {
  V start;
  int ok;
  int n[2];
  const float* q[2];
  // from the same start: CheckQuadLinear's result and reduction point, and
  // NormalUnit of the legs start -> control (ab) and control -> end (bc)
  int linear;
  V reduction;
  int ok_ab, ok_bc;
  V n_ab, u_ab, n_bc, u_bc;
};
HIPR_SD inline bool SameBits(V a, V b) {
  uint32_t u[4];
  __builtin_memcpy(&u[0], &a, 8);
  __builtin_memcpy(&u[2], &b, 8);
  return u[0] == u[2] && u[1] == u[3];
}

struct Stroker {
  float radius, miter_limit, inv_res, inv_res_sq;
  int cap, join;
  PathW* outer;
  PathW* inner;
  V first_normal, prev_normal, first_unit, prev_unit, first_pt, prev_pt, first_outer_pt;
  int seg_count;
  int max_inner_v, max_inner_p, max_inner_w;  // largest inner path reversed
  bool join_completed;
  int depth;
  int stroke_type;
  QC* stack;         // the quad-fit recursion (caller-provided: LDS in the
  int* state;        // stroke kernel); level i at [i * stack_stride]
  int stack_stride;
  int stack_cap;     // levels; a recursion that needs more sets `overflow`
  bool overflow;

  HIPR_SD void Setup(float r, float ml, int cp, int jn, PathW* o, PathW* in, QC* stack_mem,
                     int* state_mem) {
    stack = stack_mem;
    state = state_mem;
    stack_stride = 1;
    stack_cap = kRecursiveLimits + 3;  // never reached: depth <= kRecursiveLimits
    overflow = false;
    radius = r;
    miter_limit = ml;
    cap = cp;
    join = (jn == HIPR_JOIN_MITER && ml <= 1.f) ? HIPR_JOIN_BEVEL : jn;
    inv_res = 1.f / (1.f * 4);
    inv_res_sq = inv_res * inv_res;
    outer = o;
    inner = in;
    seg_count = -1;
    max_inner_v = max_inner_p = max_inner_w = 0;
    join_completed = false;
    depth = 0;
    stroke_type = 1;
    first_normal = prev_normal = first_unit = prev_unit = V{0, 0};
    first_pt = prev_pt = first_outer_pt = V{0, 0};
  }

  // ---- joins (stroke.cc:200-330) ----
  HIPR_SD void InnerJoin(PathW* in, V pivot, V after) {
    in->LineTo(pivot);
    in->LineTo(V{pivot.x - after.x, pivot.y - after.y});
  }
  HIPR_SD int AngleType(float dot) {  // 0 180, 1 sharp, 2 shallow, 3 line
    if (dot >= 0) return NearlyZero(1.f - dot) ? 3 : 2;
    return NearlyZero(1.f + dot) ? 0 : 1;
  }
  HIPR_SD int BuildUnitArc(V start, V stop, bool ccw, const M3& user, V pts[5][3],
                              float ws[5]) {
    float x = Dot(start, stop);
    float y = Cross(start, stop);
    float absY = fabsf(y);
    if (absY <= kNearlyZero && x > 0 && ((y >= 0 && !ccw) || (y <= 0 && ccw))) return 0;
    if (ccw) y = -y;
    int quadrant = 0;
    if (NearlyZero(y)) {
      quadrant = 2;
    } else if (NearlyZero(x)) {
      quadrant = y > 0 ? 1 : 3;
    } else {
      if (y < 0) quadrant += 2;
      if ((x < 0) != (y < 0)) quadrant += 1;
    }
    const V qp[8] = {{1, 0}, {1, 1}, {0, 1}, {-1, 1}, {-1, 0}, {-1, -1}, {0, -1}, {1, -1}};
    int count = quadrant;
    for (int i = 0; i < count; i++) {
      pts[i][0] = qp[i * 2];
      pts[i][1] = qp[i * 2 + 1];
      pts[i][2] = qp[i * 2 + 2];
      ws[i] = kRoot2Over2;
    }
    V finalP{x, y};
    V lastQ = qp[quadrant * 2];
    float dot = Dot(lastQ, finalP);
    if (dot < 1) {
      V off{lastQ.x + x, lastQ.y + y};
      float cos_half = sqrtf((1 + dot) / 2);
      SetLength(&off, off.x, off.y, 1.f / cos_half);
      if (CanNorm(lastQ.x - off.x, lastQ.y - off.y)) {
        pts[count][0] = lastQ;
        pts[count][1] = off;
        pts[count][2] = finalP;
        ws[count] = cos_half;
        count++;
      }
    }
    M3 rot{start.x, -start.y, 0.f, start.y, start.x, 0.f};
    if (ccw) {
      M3 flip{1.f, 0.f, 0.f, 0.f, -1.f, 0.f};
      rot = MMul(rot, flip);
    }
    M3 m = MMul(user, rot);
    for (int i = 0; i < count; i++)
      for (int jj = 0; jj < 3; jj++) pts[i][jj] = MApply(m, pts[i][jj]);
    return count;
  }
  // The two paths trade contents where the join's outer side is the inner
  // path (a runtime choice of path object would keep both in memory).
  HIPR_SDI void SwapPaths() {
    PathW t = *outer;
    *outer = *inner;
    *inner = t;
  }
  HIPR_SDI void Join(V before_un, V pivot, V after_un) {
    bool sw = false;
    JoinSides(before_un, pivot, after_un, &sw);
    if (sw) SwapPaths();
  }
  HIPR_SDI void JoinSides(V before_un, V pivot, V after_un, bool* sw) {
    PathW* o = outer;
    PathW* in = inner;
    if (join == HIPR_JOIN_BEVEL) {
      V after = Mul(after_un, radius);
      if (!(before_un.x * after_un.y > before_un.y * after_un.x)) {
        SwapPaths();
        *sw = true;
        after = V{-after.x, -after.y};
      }
      o->LineTo(V{pivot.x + after.x, pivot.y + after.y});
      InnerJoin(in, pivot, after);
      return;
    }
    float dot = Dot(before_un, after_un);
    int at = AngleType(dot);
    if (join == HIPR_JOIN_ROUND) {
      if (at == 3) return;
      V before = before_un, after = after_un;
      bool ccw = false;
      if (!(before.x * after.y > before.y * after.x)) {
        SwapPaths();
        *sw = true;
        before = V{-before.x, -before.y};
        after = V{-after.x, -after.y};
        ccw = true;
      }
      M3 T{1.f, 0.f, pivot.x, 0.f, 1.f, pivot.y};
      M3 S{radius, 0.f, 0.f, 0.f, radius, 0.f};
      M3 user = MMul(T, S);
      V cp[5][3];
      float ws[5];
      int count = BuildUnitArc(before, after, ccw, user, cp, ws);
      if (count > 0) {
        for (int i = 0; i < count; i++) o->ConicTo(cp[i][1], cp[i][2], ws[i]);
        after = Mul(after, radius);
        InnerJoin(in, pivot, after);
      }
      return;
    }
    // miter
    V before = before_un, after = after_un, mid;
    if (at == 3) return;
    if (at == 0) {
      after = Mul(after, radius);
      o->LineTo(V{pivot.x + after.x, pivot.y + after.y});
      InnerJoin(in, pivot, after);
      return;
    }
    if (!(before.x * after.y > before.y * after.x)) {
      SwapPaths();
      *sw = true;
      before = V{-before.x, -before.y};
      after = V{-after.x, -after.y};
    }
    if (0 == dot && miter_limit >= kSqrt2) {
      mid = Mul(Add(before, after), radius);
    } else {
      V od = Add(before, after);
      float k = 2.f / (od.x * od.x + od.y * od.y);
      mid = Mul(V{k * od.x, k * od.y}, radius);
      if (sqrtf(mid.x * mid.x + mid.y * mid.y) >= miter_limit * radius) {
        after = Mul(after, radius);
        o->LineTo(V{pivot.x + after.x, pivot.y + after.y});
        InnerJoin(in, pivot, after);
        return;
      }
    }
    o->LineTo(V{pivot.x + mid.x, pivot.y + mid.y});
    after = Mul(after, radius);
    o->LineTo(V{pivot.x + after.x, pivot.y + after.y});
    InnerJoin(in, pivot, after);
  }
  // ---- caps (stroke.cc:163-194) ----
  HIPR_SDI void Cap(V pivot, V normal, V stop) {
    if (cap == HIPR_CAP_BUTT) {
      outer->LineTo(stop);
    } else if (cap == HIPR_CAP_SQUARE) {
      V par = RotCW(normal);
      outer->LineTo(V{pivot.x + normal.x + par.x, pivot.y + normal.y + par.y});
      outer->LineTo(V{pivot.x - normal.x + par.x, pivot.y - normal.y + par.y});
      outer->LineTo(stop);
    } else {
      V par = RotCW(normal);
      V pc = Add(pivot, par);
      outer->ConicTo(Add(pc, normal), pc, kRoot2Over2);
      outer->ConicTo(Sub(pc, normal), stop, kRoot2Over2);
    }
  }
  HIPR_SD bool NormalUnit(V before, V after, V* normal, V* unit) {
    if (!SetLength(unit, (after.x - before.x) * 1.f, (after.y - before.y) * 1.f, 1.0f))
      return false;
    *unit = RotCCW(*unit);
    *normal = Mul(*unit, radius);
    return true;
  }
  // `ahead`: the normal and unit of prev_pt -> curr computed ahead (ok_ab,
  // n_ab, u_ab of a SpecView whose start is prev_pt), or null.
  HIPR_SD bool PreJoinTo(V curr, V* normal, V* unit, const SpecView* ahead = nullptr) {
    float px = prev_pt.x, py = prev_pt.y;
    bool ok;
    if (ahead) {
      ok = ahead->ok_ab != 0;
      *normal = ahead->n_ab;
      *unit = ahead->u_ab;
    } else {
      ok = NormalUnit(prev_pt, curr, normal, unit);
    }
    if (!ok) {
      if (cap == HIPR_CAP_BUTT) return false;
      *normal = V{radius, 0};
      *unit = V{1, 0};
    }
    if (seg_count == 0) {
      first_normal = *normal;
      first_unit = *unit;
      first_outer_pt = V{px + normal->x, py + normal->y};
      outer->MoveTo(first_outer_pt);
      inner->MoveTo(V{px - normal->x, py - normal->y});
    } else {
      Join(prev_unit, prev_pt, *unit);
    }
    return true;
  }
  HIPR_SD void PostJoinTo(V curr, V normal, V unit) {
    join_completed = true;
    prev_pt = curr;
    prev_unit = unit;
    prev_normal = normal;
    seg_count += 1;
  }
  // Path::ReversePathTo of the inner contour onto the outer path.
  HIPR_SDI void ReverseInner() {
    // the inner path is largest here (after the closing join or the end cap)
    max_inner_v = Max(max_inner_v, inner->nv);
    max_inner_p = Max(max_inner_p, inner->np);
    max_inner_w = Max(max_inner_w, inner->nw);
    if (inner->nv == 0) return;
    if (!outer->verbs || inner->nv > inner->cap_v || inner->np > inner->cap_p ||
        inner->nw > inner->cap_w) {
      // count mode, or an inner contour that outgrew its range (the job is
      // written again into exact ranges, and its stored contour is
      // incomplete): the reversal appends every non-move verb and point
      outer->nv += inner->nv - 1;
      outer->np += inner->np - 1;
      outer->nw += inner->nw;
      outer->last_verb = HIPR_VERB_LINE;
      return;
    }
    int vi = inner->nv, pi = inner->np - 1, wi = inner->nw;
    while (vi > 0) {
      int v = inner->verbs[--vi];
      int n = v == HIPR_VERB_MOVE || v == HIPR_VERB_LINE ? 1
              : v == HIPR_VERB_QUAD || v == HIPR_VERB_CONIC ? 2 : v == HIPR_VERB_CUBIC ? 3 : 0;
      pi -= n;
      auto P = [&](int i) { return V{inner->pts[2 * i], inner->pts[2 * i + 1]}; };
      if (v == HIPR_VERB_MOVE) return;
      if (v == HIPR_VERB_LINE) {
        outer->LineTo(P(pi));
      } else if (v == HIPR_VERB_QUAD) {
        outer->QuadTo(P(pi + 1), P(pi));
      } else if (v == HIPR_VERB_CONIC) {
        outer->ConicTo(P(pi + 1), P(pi), inner->w[--wi]);
      }
    }
  }
  HIPR_SDI void FinishContour(bool close) {
    if (seg_count > 0) {
      if (close) {
        Join(prev_unit, prev_pt, first_unit);
        outer->Close();
        V pt = inner->last_pt;
        outer->MoveTo(pt);
        ReverseInner();
        outer->Close();
      } else {
        V pt = inner->last_pt;
        Cap(prev_pt, prev_normal, pt);
        ReverseInner();
        Cap(first_pt, V{-first_normal.x, -first_normal.y}, first_outer_pt);
        outer->Close();
      }
    }
    inner->Reset();
    seg_count = -1;
  }
  HIPR_SDI void MoveTo(V pt) {
    if (seg_count > 0) FinishContour(false);
    seg_count = 0;
    first_pt = prev_pt = pt;
    join_completed = false;
  }
  HIPR_SDI void LineTo(V curr) {
    bool teeny = NearlyZero(prev_pt.x - curr.x, kNearlyZero * inv_res) &&
                 NearlyZero(prev_pt.y - curr.y, kNearlyZero * inv_res);
    if (teeny && cap == HIPR_CAP_BUTT) return;
    if (teeny && join_completed) return;
    V normal, unit;
    if (!PreJoinTo(curr, &normal, &unit)) return;
    outer->LineTo(V{curr.x + normal.x, curr.y + normal.y});
    inner->LineTo(V{curr.x - normal.x, curr.y - normal.y});
    PostJoinTo(curr, normal, unit);
  }
  // ---- quad stroking (stroke.cc:486-712) ----
  HIPR_SD void SetRayPts(V tpt, V* dxy, V* on, V* tangent) {
    if (!SetLength(dxy, dxy->x, dxy->y, radius)) {
      dxy->x = radius;
      dxy->y = 0;
    }
    float flip = (float)stroke_type;
    on->x = tpt.x + flip * dxy->y;
    on->y = tpt.y - flip * dxy->x;
    if (tangent) {
      tangent->x = on->x + dxy->x;
      tangent->y = on->y + dxy->y;
    }
  }
  HIPR_SD void QuadPerpRay(const V q[3], float t, V* tpt, V* on, V* tangent) {
    float tc = t < 0 ? 0 : (t > 1.f ? 1.f : t);
    *tpt = QuadEval(q, tc);
    V dxy = QuadTangent(q, tc);
    if (dxy.x == 0 && dxy.y == 0) dxy = Sub(q[2], q[0]);
    SetRayPts(*tpt, &dxy, on, tangent);
  }
  HIPR_SD int IntersectRay(QC* qp) {  // 0 split, 1 degenerate, 2 quad
    V start = qp->quad[0], end = qp->quad[2];
    V aLen = Sub(qp->tan_start, start), bLen = Sub(qp->tan_end, end);
    float denom = Cross(aLen, bLen);
    if (denom == 0 || denom == INFINITY) return 1;
    V ab0 = Sub(start, end);
    float numerA = Cross(bLen, ab0), numerB = Cross(aLen, ab0);
    if ((numerA >= 0) == (numerB >= 0)) {
      float d1 = PtToLine(start, end, qp->tan_end);
      float d2 = PtToLine(end, start, qp->tan_start);
      if (fmaxf(d1, d2) <= inv_res_sq) return 1;
      return 0;
    }
    numerA /= denom;
    bool valid = numerA > numerA - 1;
    if (valid) {
      qp->quad[1].x = start.x * (1 - numerA) + qp->tan_start.x * numerA;
      qp->quad[1].y = start.y * (1 - numerA) + qp->tan_start.y * numerA;
      return 2;
    }
    return 1;
  }
  HIPR_SD bool SharpAngle(const V q[3]) {
    V smaller = Sub(q[1], q[0]), larger = Sub(q[1], q[2]);
    float sl = Dot(smaller, smaller), ll = Dot(larger, larger);
    if (sl > ll) {
      V t = smaller;
      smaller = larger;
      larger = t;
      ll = sl;
    }
    if (!SetLength(&smaller, smaller.x, smaller.y, ll)) return false;
    return Dot(smaller, larger) > 0;
  }
  HIPR_SD int CloseEnough(const V s[3], const V ray[2], QC* qp) {
    V mid = QuadEval(s, 0.5f);
    if (DistSqd(ray[0], mid) <= inv_res * inv_res) return SharpAngle(qp->quad) ? 0 : 2;
    float xmin = fminf(fminf(s[0].x, s[1].x), s[2].x);
    float xmax = fmaxf(fmaxf(s[0].x, s[1].x), s[2].x);
    float ymin = fminf(fminf(s[0].y, s[1].y), s[2].y);
    float ymax = fmaxf(fmaxf(s[0].y, s[1].y), s[2].y);
    V p = ray[0];
    if (p.x + inv_res < xmin || p.x - inv_res > xmax || p.y + inv_res < ymin ||
        p.y - inv_res > ymax)
      return 0;
    V vec = Sub(ray[1], ray[0]);
    float r[3];
    for (int n = 0; n < 3; ++n) r[n] = (s[n].y - ray[0].y) * vec.x - (s[n].x - ray[0].x) * vec.y;
    float A = r[2], B = r[1], C = r[0];
    A += C - 2 * B;
    B -= C;
    float roots[2];
    int rc = FindUnitQuadRoots(A, 2 * B, C, roots);
    if (rc != 1) return 0;
    V qpt = QuadEval(s, roots[0]);
    float error = inv_res * (1.f - fabsf(roots[0] - 0.5f) * 2);
    if (DistSqd(ray[0], qpt) <= error * error) return SharpAngle(qp->quad) ? 0 : 2;
    return 0;
  }
  HIPR_SD int Compare(const V q[3], QC* qp) {
    if (!qp->start_set) {
      V tpt;
      QuadPerpRay(q, qp->start_t, &tpt, &qp->quad[0], &qp->tan_start);
      qp->start_set = true;
    }
    if (!qp->end_set) {
      V tpt;
      QuadPerpRay(q, qp->end_t, &tpt, &qp->quad[2], &qp->tan_end);
      qp->end_set = true;
    }
    int rt = IntersectRay(qp);
    if (rt != 2) return rt;
    V ray[2];
    QuadPerpRay(q, qp->mid_t, &ray[1], &ray[0], nullptr);
    return CloseEnough(qp->quad, ray, qp);
  }
  // PathStroker::QuadStroke with its recursion as an explicit stack.
  HIPR_SDI void QuadStroke(const V q[3], PathW* path) {
    const int s = stack_stride;
    int sp = 0;
    stack[0].Init(0, 1);
    state[0] = 0;
    while (sp >= 0) {
      QC* node = &stack[sp * s];
      int* ns = &state[sp * s];
      if (*ns == 0) {
        int rt = Compare(q, node);
        if (rt == 2) {
          path->QuadTo(node->quad[1], node->quad[2]);
          sp--;
          continue;
        }
        if (rt == 1) {
          sp--;
          continue;
        }
        if (++depth > kRecursiveLimits) return;  // every level returns false
        if (sp + 1 >= stack_cap) {
          overflow = true;
          return;
        }
        *ns = 1;
        stack[(sp + 1) * s].InitWithStart(*node);
        state[(sp + 1) * s] = 0;
        sp++;
      } else if (*ns == 1) {
        *ns = 2;
        stack[(sp + 1) * s].InitWithEnd(*node);
        state[(sp + 1) * s] = 0;
        sp++;
      } else {
        --depth;
        sp--;
      }
    }
  }
  HIPR_SD int CheckQuadLinear(const V q[3], V* reduction) {
    V p1c = Sub(q[1], q[0]);
    bool d1 = !CanNorm(p1c.x, p1c.y);
    V cp2 = Sub(q[2], q[1]);
    bool d2 = !CanNorm(cp2.x, cp2.y);
    if (d1 && d2) return 0;
    if (d1 || d2) return 1;
    // quad_in_line
    float pt_max = -1;
    int o1 = 0, o2 = 0;
    for (int index = 0; index < 2; ++index) {
      for (int in = index + 1; in < 3; ++in) {
        V d = Sub(q[in], q[index]);
        float m = fmaxf(fabsf(d.x), fabsf(d.y));
        if (pt_max < m) {
          o1 = index;
          o2 = in;
          pt_max = m;
        }
      }
    }
    int mid = o1 ^ o2 ^ 3;
    float line_slop = pt_max * pt_max * 0.000005f;
    if (!(PtToLine(q[mid], q[o1], q[o2]) <= line_slop)) return 2;
    // find_quad_max_curvature
    float Ax = q[1].x - q[0].x, Ay = q[1].y - q[0].y;
    float Bx = q[0].x - q[1].x - q[1].x + q[2].x;
    float By = q[0].y - q[1].y - q[1].y + q[2].y;
    float numer = -(Ax * Bx + Ay * By);
    float denom = Bx * Bx + By * By;
    if (denom < 0) {
      numer = -numer;
      denom = -denom;
    }
    float t;
    if (numer <= 0) {
      t = 0;
    } else if (numer >= denom) {
      t = 1;
    } else {
      t = numer / denom;
    }
    if (t == 0 || t == 1) return 1;
    *reduction = QuadEval(q, t);
    return 3;
  }
  // `spec`, when given, holds this segment's two offset curves stroked
  // ahead from a presumed start point; they stand in for the two QuadStroke
  // calls when that start is the stroker's current point and the recursion
  // depth is zero (the state QuadStroke would start from).
  HIPR_SDI void QuadTo(V c, V p2, const SpecView* spec = nullptr) {
    V quad[3] = {prev_pt, c, p2};
    // values computed ahead from this very start point stand in for the
    // stroker's own calls
    const SpecView* ahead = spec && SameBits(prev_pt, spec->start) ? spec : nullptr;
    V reduction;
    int rt;
    if (ahead) {
      rt = ahead->linear;
      reduction = ahead->reduction;
    } else {
      rt = CheckQuadLinear(quad, &reduction);
    }
    if (rt == 0 || rt == 1) {
      LineTo(p2);
      return;
    }
    if (rt == 3) {
      LineTo(reduction);
      int save = join;
      join = HIPR_JOIN_ROUND;
      LineTo(p2);
      join = save;
      return;
    }
    V nAB, uAB, nBC, uBC;
    if (!PreJoinTo(c, &nAB, &uAB, ahead)) {
      LineTo(p2);
      return;
    }
    if (ahead && ahead->ok && depth == 0) {
      for (int k = 0; k < spec->n[0]; k++) {
        const float* f = spec->q[0] + 4 * k;
        outer->QuadTo(V{f[0], f[1]}, V{f[2], f[3]});
      }
      for (int k = 0; k < spec->n[1]; k++) {
        const float* f = spec->q[1] + 4 * k;
        inner->QuadTo(V{f[0], f[1]}, V{f[2], f[3]});
      }
    } else {
      stroke_type = 1;
      QuadStroke(quad, outer);
      stroke_type = -1;
      QuadStroke(quad, inner);
    }
    bool ok_bc;
    if (ahead) {
      ok_bc = ahead->ok_bc != 0;
      nBC = ahead->n_bc;
      uBC = ahead->u_bc;
    } else {
      ok_bc = NormalUnit(quad[1], quad[2], &nBC, &uBC);
    }
    if (!ok_bc) {
      nBC = nAB;
      uBC = uAB;
    }
    PostJoinTo(p2, nBC, uBC);
  }
};

}  // namespace stroke

namespace stroke {
// Storage of the quad-fit recursion of one stroke.
struct StrokeStack {
  QC quads[kRecursiveLimits + 3];
  int state[kRecursiveLimits + 3];
};

// The stroker calls of Path::Iter(quad path, false) over the quad path of a
// source path (Stroke::QuadPath, stroke.cc:914-962), one at a time: the quad
// path is produced on the fly and never stored.
enum OpKind { kOpMove, kOpLine, kOpQuad, kOpClose };
struct Op {
  int kind;
  V s, c, p;  // quad: start (the previous quad-path point), control, end; move, line: p
};
struct OpIter {
  const uint8_t* verbs;
  const float* pts;
  const float* w;
  int nv, k, pi, wi;
  V src_move, src_last;  // Iter over the source
  V q_move, q_last;      // Iter over the quad path
  int sub, nsub;         // quads of the conic or cubic in progress: emitted, total
  bool cubic;
  int close_step;        // 1, 2: closing lines still to check; 3: the close
  P2 q5[5];
  CubicQ cq;

  HIPR_SDI void Init(const Job& job, GeomSrc scene) {
    verbs = scene.verbs + job.verb_begin;
    pts = scene.pts + 2 * (int64_t)job.pt_begin;
    w = scene.weights + job.w_begin;
    nv = job.verb_end - job.verb_begin;
    k = pi = wi = 0;
    src_move = src_last = q_move = q_last = V{0, 0};
    sub = nsub = 0;
    cubic = false;
    close_step = 0;
  }
  HIPR_SDI V P(int i) const { return V{pts[2 * i], pts[2 * i + 1]}; }
  HIPR_SDI void Line(Op* op, V p) {
    op->kind = kOpLine;
    op->p = p;
    q_last = p;
  }
  HIPR_SDI void Quad(Op* op, V c, V p) {
    op->kind = kOpQuad;
    op->s = q_last;
    op->c = c;
    op->p = p;
    q_last = p;
  }
  // The next stroker call; false at the end of the path.
  HIPR_SDI bool Next(Op* op) {
    for (;;) {
      if (sub < nsub) {
        int i = sub++;
        if (cubic) {
          P2 c, e;
          CubicQuad(cq, i, &c, &e);
          Quad(op, V{c.x, c.y}, V{e.x, e.y});
        } else {
          Quad(op, V{q5[2 * i + 1].x, q5[2 * i + 1].y}, V{q5[2 * i + 2].x, q5[2 * i + 2].y});
        }
        return true;
      }
      if (close_step != 0) {
        // Iter(source).AutoClose emits the closing line into the quad path,
        // then Iter(quad path) closes with a line if its last point differs.
        if (close_step == 1) {
          close_step = 2;
          if (!Eq(src_last, src_move)) {
            src_last = src_move;
            Line(op, src_move);
            return true;
          }
        }
        if (close_step == 2) {
          close_step = 3;
          if (!Eq(q_last, q_move)) {
            Line(op, q_move);
            return true;
          }
        }
        close_step = 0;
        src_last = src_move;
        op->kind = kOpClose;
        return true;
      }
      if (k >= nv) return false;
      int v = verbs[k++];
      if (v == HIPR_VERB_MOVE) {
        if (k == nv) return false;  // trailing move: Iter returns kDone
        src_move = src_last = P(pi);
        pi++;
        q_move = q_last = src_move;
        op->kind = kOpMove;
        op->p = q_move;
        return true;
      } else if (v == HIPR_VERB_LINE) {
        V e = P(pi);
        pi++;
        src_last = e;
        Line(op, e);
        return true;
      } else if (v == HIPR_VERB_QUAD) {
        V c = P(pi), e = P(pi + 1);
        pi += 2;
        src_last = e;
        Quad(op, c, e);
        return true;
      } else if (v == HIPR_VERB_CONIC) {
        P2 src3[3] = {{pts[2 * (pi - 1)], pts[2 * (pi - 1) + 1]},
                      {pts[2 * pi], pts[2 * pi + 1]},
                      {pts[2 * (pi + 1)], pts[2 * (pi + 1) + 1]}};
        ConicToQuads(src3, w[wi], q5);
        wi++;
        pi += 2;
        src_last = V{src3[2].x, src3[2].y};
        cubic = false;
        sub = 0;
        nsub = 2;
      } else if (v == HIPR_VERB_CUBIC) {
        P2 src4[4] = {{pts[2 * (pi - 1)], pts[2 * (pi - 1) + 1]},
                      {pts[2 * pi], pts[2 * pi + 1]},
                      {pts[2 * (pi + 1)], pts[2 * (pi + 1) + 1]},
                      {pts[2 * (pi + 2)], pts[2 * (pi + 2) + 1]}};
        pi += 3;
        CubicSetup(src4, &cq);
        src_last = V{src4[3].x, src4[3].y};
        cubic = true;
        sub = 0;
        nsub = cq.n;
      } else if (v == HIPR_VERB_CLOSE) {
        close_step = 1;
      }
    }
  }
};

HIPR_SDI void Apply(Stroker* st, const Op& op) {
  if (op.kind == kOpMove) {
    st->MoveTo(op.p);
  } else if (op.kind == kOpLine) {
    st->LineTo(op.p);
  } else if (op.kind == kOpQuad) {
    st->QuadTo(op.c, op.p);
  } else {
    st->FinishContour(true);
  }
}
}  // namespace stroke

// Output region of one stroke job.
struct StrokeOut {
  int32_t verb_begin, pt_begin, w_begin;   // in the outline arrays
  int32_t nv, np, nw;                      // sizes (count mode result)
  int32_t scratch_begin;                   // inner-path scratch (points)
  int32_t inner_nv, inner_np, inner_nw;    // largest inner contour
};

// Outline of one stroke job.  mode 0 counts (fills so's sizes); mode 1
// writes into the regions `so` describes; mode 2 writes into them as far as
// they reach and reports the full sizes (a region too small shows as a size
// above its capacity, and nothing is written past it).  `feed` drives the
// stroker through the quad path.
template <class Feed>
HIPR_SDI void StrokeOutline(const Job& job, const hipr_paint& paint, GeomSrc scene,
                            StrokeOut* so_io, int mode, uint8_t* o_verbs, float* o_pts,
                            float* o_w, uint8_t* s_verbs, float* s_pts, float* s_w,
                            stroke::StrokeStack* stack_mem, Feed feed) {
  using namespace stroke;
  StrokeOut so = *so_io;
  PathW outer, inner;
  if (mode != 0) {
    outer.Init(o_verbs + so.verb_begin, o_pts + 2 * (int64_t)so.pt_begin, o_w + so.w_begin,
               so.nv, so.np, so.nw);
    inner.Init(s_verbs + so.scratch_begin, s_pts + 2 * (int64_t)so.scratch_begin,
               s_w + so.scratch_begin, so.inner_nv, so.inner_np, so.inner_nw);
  } else {
    outer.Init(nullptr, nullptr, nullptr, 0, 0, 0);
    inner.Init(nullptr, nullptr, nullptr, 0, 0, 0);
  }
  Stroker stk;
  stk.Setup(paint.stroke_width * 0.5f, paint.miter_limit, paint.cap, paint.join, &outer,
            &inner, stack_mem->quads, stack_mem->state);
  OpIter it;
  it.Init(job, scene);
  feed(&stk, &it);
  stk.FinishContour(false);
  if (mode != 1) {
    so.nv = outer.nv;
    so.np = outer.np;
    so.nw = outer.nw;
    so.inner_nv = Max(stk.max_inner_v, 1);
    so.inner_np = Max(stk.max_inner_p, 1);
    so.inner_nw = Max(stk.max_inner_w, 1);
    *so_io = so;
  }
}

// The stroker run call by call on one thread.
HIPR_SD inline void StrokeOne(const Job& job, const hipr_paint& paint, GeomSrc scene,
                              StrokeOut* so_io, int mode, uint8_t* o_verbs, float* o_pts,
                              float* o_w, uint8_t* s_verbs, float* s_pts, float* s_w,
                              stroke::StrokeStack* stack_mem) {
  StrokeOutline(job, paint, scene, so_io, mode, o_verbs, o_pts, o_w, s_verbs, s_pts, s_w,
                stack_mem, [](stroke::Stroker* st, stroke::OpIter* it) {
                  stroke::Op op;
                  while (it->Next(&op)) stroke::Apply(st, op);
                });
}

namespace stroke {
// One wavefront per stroke.  The quad path is taken in chunks of up to
// kSpecLanes quad segments; lane i strokes segment i's two offset curves
// (PathStroker::QuadStroke) from the segment's start point in the quad
// path, with a recursion stack of kSpecDepth levels in LDS; then the whole
// wave runs the stroker over the chunk call by call, taking a lane's curves
// in place of its own QuadStroke calls where QuadTo accepts them.  A segment
// whose curves needed deeper recursion or more quads than a lane keeps, or
// whose start differs from the stroker's point, is stroked there as before,
// so the outline is the one-thread stroker's.
constexpr int kSpecLanes = 64;
constexpr int kSpecDepth = 6;
constexpr int kSpecQuads = 8;
constexpr int kChunkOps = 128;
constexpr int kSpecStride = 2 * 4 * kSpecQuads + 1;  // floats per lane (odd: no bank conflicts)
struct SpecStack {
  QC quads[kSpecDepth * kSpecLanes];  // level i of lane l at [i * kSpecLanes + l]
  int state[kSpecDepth * kSpecLanes];
};
struct WaveLds {
  union {
    StrokeStack serial;  // the stroker's own recursion (call by call)
    SpecStack spec;      // the lanes' recursions (ahead of it)
  } stack;
  float out[kSpecLanes * kSpecStride];
  int ok[kSpecLanes];
  int n[2][kSpecLanes];
  int linear[kSpecLanes], ok_ab[kSpecLanes], ok_bc[kSpecLanes];
  V leg[5][kSpecLanes];  // reduction, n_ab, u_ab, n_bc, u_bc
  int quad_op[kSpecLanes];  // chunk position of each quad segment
  Op ops[kChunkOps];
};

// Lane `lane`'s segment: both offset curves into its LDS output.
HIPR_SDI void StrokeAhead(const Stroker& cfg, WaveLds* L, int lane) {
  const Op o = L->ops[L->quad_op[lane]];
  const V quad[3] = {o.s, o.c, o.p};
  Stroker sp;
  sp.radius = cfg.radius;
  sp.miter_limit = cfg.miter_limit;
  sp.inv_res = cfg.inv_res;
  sp.inv_res_sq = cfg.inv_res_sq;
  sp.cap = cfg.cap;
  sp.join = cfg.join;
  sp.outer = sp.inner = nullptr;
  sp.stack = &L->stack.spec.quads[lane];
  sp.state = &L->stack.spec.state[lane];
  sp.stack_stride = kSpecLanes;
  sp.stack_cap = kSpecDepth;
  float* out = &L->out[lane * kSpecStride];
  int ok = 0, n0 = 0, n1 = 0;
  V reduction;
  auto side = [&](int type, float* dst, int* n) __attribute__((always_inline)) {
    PathW pw;
    pw.Init(nullptr, dst, nullptr, 0, 2 * kSpecQuads, 0);
    pw.move_open = pw.any_move = true;  // no injected move
    sp.stroke_type = type;
    sp.depth = 0;
    sp.overflow = false;
    sp.QuadStroke(quad, &pw);
    *n = pw.nv;
    return !sp.overflow && sp.depth == 0 && pw.nv <= kSpecQuads;
  };
  reduction = V{0, 0};
  const int linear = sp.CheckQuadLinear(quad, &reduction);
  if (linear == 2) {
    ok = side(1, out, &n0);
    ok = side(-1, out + 4 * kSpecQuads, &n1) && ok;
  }
  V n_ab{0, 0}, u_ab{0, 0}, n_bc{0, 0}, u_bc{0, 0};
  const bool ok_ab = sp.NormalUnit(o.s, o.c, &n_ab, &u_ab);
  const bool ok_bc = sp.NormalUnit(o.c, o.p, &n_bc, &u_bc);
  L->ok[lane] = ok;
  L->n[0][lane] = n0;
  L->n[1][lane] = n1;
  L->linear[lane] = linear;
  L->ok_ab[lane] = ok_ab;
  L->ok_bc[lane] = ok_bc;
  L->leg[0][lane] = reduction;
  L->leg[1][lane] = n_ab;
  L->leg[2][lane] = u_ab;
  L->leg[3][lane] = n_bc;
  L->leg[4][lane] = u_bc;
}

// The chunked feed.  `writer` marks the one lane that stores a chunk's calls,
// `sync()` orders the LDS between the phases, and `ahead(nq)` runs
// StrokeAhead for lanes [0, nq) (the wave's lanes on the device; a loop in
// the host build that checks this feed).
template <class Sync, class Ahead>
HIPR_SDI void FeedChunks(Stroker* st, OpIter* it, WaveLds* L, bool writer, Sync sync,
                         Ahead ahead) {
  bool more = true;
  while (more) {
    int n = 0, nq = 0;
    while (n < kChunkOps && nq < kSpecLanes) {
      Op op;
      if (!it->Next(&op)) {
        more = false;
        break;
      }
      if (writer) {
        L->ops[n] = op;
        if (op.kind == kOpQuad) L->quad_op[nq] = n;
      }
      nq += op.kind == kOpQuad;
      n++;
    }
    sync();
    ahead(nq);
    sync();
    int q = 0;
    for (int i = 0; i < n; i++) {
      const Op op = L->ops[i];
      if (op.kind == kOpQuad) {
        SpecView sv;
        sv.start = op.s;
        sv.ok = L->ok[q];
        sv.n[0] = L->n[0][q];
        sv.n[1] = L->n[1][q];
        sv.q[0] = &L->out[q * kSpecStride];
        sv.q[1] = sv.q[0] + 4 * kSpecQuads;
        sv.linear = L->linear[q];
        sv.ok_ab = L->ok_ab[q];
        sv.ok_bc = L->ok_bc[q];
        sv.reduction = L->leg[0][q];
        sv.n_ab = L->leg[1][q];
        sv.u_ab = L->leg[2][q];
        sv.n_bc = L->leg[3][q];
        sv.u_bc = L->leg[4][q];
        st->QuadTo(op.c, op.p, &sv);
        q++;
      } else {
        Apply(st, op);
      }
    }
    sync();  // the next chunk reuses the LDS
  }
}
}  // namespace stroke

// One wavefront per stroked draw; the launch order puts the longest quad
// paths first.  Every lane runs the stroker on the same values, so the
// outline stores of the 64 lanes carry one value per address.
template <int kMode>
__global__ void __launch_bounds__(64)
    StrokeKernel(const Job* __restrict__ jobs, const int32_t* __restrict__ list, int n,
                 GeomSrc scene, const hipr_paint* __restrict__ paints, StrokeOut* outs,
                 uint8_t* o_verbs, float* o_pts, float* o_w, uint8_t* s_verbs, float* s_pts,
                 float* s_w) {
  __shared__ stroke::WaveLds lds;
  const int li = blockIdx.x;
  if (li >= n) return;
  const Job job = jobs[list[li]];
  StrokeOut so = outs[li];
  StrokeOutline(job, paints[job.path], scene, &so, kMode, o_verbs, o_pts, o_w, s_verbs, s_pts,
                s_w, &lds.stack.serial,
                [&](stroke::Stroker* st, stroke::OpIter* it) __attribute__((always_inline)) {
                  const int lane = threadIdx.x;
                  stroke::FeedChunks(
                      st, it, &lds, lane == 0, []() { __syncthreads(); },
                      [&](int nq) {
                        if (lane < nq) stroke::StrokeAhead(*st, &lds, lane);
                      });
                });
  if (kMode != 1 && threadIdx.x == 0) outs[li] = so;
}

}  // namespace hipr

#endif  // SKITY_HIP_STROKE_HPP_
