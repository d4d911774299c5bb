// oracle_stroke.cpp -- TEST INFRASTRUCTURE ONLY (parity checker).
//
// Restatement of Stroke::StrokePath / PathStroker (src/geometry/stroke.cc:
// 15-901) with the point helpers of src/geometry/point_priv.hpp, the quad
// helpers of src/geometry/geometry.cc/hpp (pt_to_line, FindUnitQuadRoots,
// QuadCoeff::EvalQuadAt / EvalQuadTangentAt) and Conic::BuildUnitArc
// (src/geometry/conic.cc:72-167).  Only quads, lines and conics reach the
// stroker (Stroke::QuadPath runs first).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "oracle_raster.hpp"

namespace orc {

namespace {

constexpr float kNearlyZero = 1.f / 4096;
constexpr float kRoot2Over2 = 0.707106781f;
constexpr float kSqrt2 = 1.41421356f;
constexpr int kRecursiveLimits = 11 * 3;

struct V {
  float x, y;
};
inline V operator+(V a, V b) { return {a.x + b.x, a.y + b.y}; }
inline V operator-(V a, V b) { return {a.x - b.x, a.y - b.y}; }
inline V operator*(V a, float s) { return {a.x * s, a.y * s}; }
inline bool operator==(V a, V b) { return a.x == b.x && a.y == b.y; }
inline float Dot(V a, V b) { return a.x * b.x + a.y * b.y; }
inline float Cross(V a, V b) { return a.x * b.y - a.y * b.x; }
inline bool NearlyZero(float x, float tol = kNearlyZero) { return std::fabs(x) <= tol; }
inline Pt P(V v) { return Pt{v.x, v.y}; }
inline V Vv(Pt p) { return V{p.x, p.y}; }

// point_priv.hpp helpers
bool SetLength(V* pt, float x, float y, float length) {
  double xx = x, yy = y;
  double dmag = std::sqrt(xx * xx + yy * yy);
  double dscale = length / dmag;
  x = (float)(x * dscale);
  y = (float)(y * dscale);
  if (!std::isfinite(x) || !std::isfinite(y) || (x == 0 && y == 0)) {
    pt->x = 0;
    pt->y = 0;
    return false;
  }
  pt->x = x;
  pt->y = y;
  return true;
}
bool SetNormal(V* v, float x, float y) { return SetLength(v, x, y, 1.0f); }
inline V RotateCCW(V s) { return V{s.y, -s.x}; }
inline V RotateCW(V s) { return V{-s.y, s.x}; }
inline float DistSqd(V a, V b) {
  float dx = a.x - b.x, dy = a.y - b.y;
  return dx * dx + dy * dy;
}
inline bool CanNormalize(float dx, float dy) {
  return std::isfinite(dx) && std::isfinite(dy) && (dx || dy);
}
inline bool EqualsWithinTol(V a, V b, float tol) {
  return NearlyZero(a.x - b.x, tol) && NearlyZero(a.y - b.y, tol);
}

// QuadCoeff (geometry.cc:44-106)
V QuadEval(const V q[3], float t) {
  V C = q[0], B = (q[1] - C) + (q[1] - C), A = (q[2] - (q[1] + q[1])) + C;
  return V{(A.x * t + B.x) * t + C.x, (A.y * t + B.y) * t + C.y};
}
V QuadTangent(const V q[3], float t) {
  if ((t == 0 && q[0] == q[1]) || (t == 1 && q[1] == q[2])) return q[2] - q[0];
  V B = q[1] - q[0];
  V A = (q[2] - q[1]) - B;
  V T{A.x * t + B.x, A.y * t + B.y};
  return T + T;
}
void QuadEvalAt(const V q[3], float t, V* pt, V* tangent) {
  if (t < 0) t = 0;
  if (t > 1.f) t = 1.f;
  if (pt) *pt = QuadEval(q, t);
  if (tangent) *tangent = QuadTangent(q, t);
}

// pt_to_line (geometry.cc:253-271)
float PtToLine(V pt, V a, V b) {
  V dxy = b - a;
  V ab0 = pt - a;
  float number = Dot(dxy, ab0);
  float denom = Dot(dxy, dxy);
  float t = number / denom;
  if (t >= 0 && t <= 1) {
    V hit{a.x * (1 - t) + b.x * t, a.y * (1 - t) + b.y * t};
    return DistSqd(hit, pt);
  }
  return DistSqd(pt, a);
}

// valid_unit_divide / FindUnitQuadRoots (geometry.hpp:134-188)
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
int FindUnitQuadRoots(float A, float B, float C, float roots[2]) {
  if (A == 0) return ValidUnitDivide(-C, B, roots);
  float* r = roots;
  double dr = (double)B * B - 4 * (double)A * C;
  if (dr < 0) return 0;
  dr = std::sqrt(dr);
  float R = (float)dr;
  if (std::isinf(R)) return 0;
  float Q = (B < 0) ? -(B - R) / 2 : -(B + R) / 2;
  r += ValidUnitDivide(Q, A, r);
  r += ValidUnitDivide(C, Q, r);
  if (r - roots == 2) {
    if (roots[0] > roots[1]) {
      std::swap(roots[0], roots[1]);
    } else if (roots[0] == roots[1]) {
      r -= 1;
    }
  }
  return (int)(r - roots);
}

// Path builder additions used by the stroker (path.cc:533-553,1042-1094).
struct PathB : PathData {
  void ConicTo(V p1, V p2, float w) {
    if (!(w > 0)) {
      LineTo(P(p2));
    } else if (std::isinf(w)) {
      LineTo(P(p1));
      LineTo(P(p2));
    } else if (w == 1.f) {
      QuadTo(P(p1), P(p2));
    } else {
      InjectMoveToIfNeeded();
      verbs.push_back(HIPR_VERB_CONIC);
      pts.push_back(P(p1));
      pts.push_back(P(p2));
      weights.push_back(w);
    }
  }
  void Reset() {
    verbs.clear();
    pts.clear();
    weights.clear();
    last_move_index = -1;
    has_last_move = false;
  }
  V LastPt() const { return pts.empty() ? V{0, 0} : Vv(pts.back()); }
  void ReversePathTo(const PathB& src) {
    if (src.verbs.empty()) return;
    int vi = (int)src.verbs.size();
    int pi = (int)src.pts.size() - 1;
    int wi = (int)src.weights.size();
    while (vi > 0) {
      int v = src.verbs[--vi];
      int n = v == HIPR_VERB_MOVE || v == HIPR_VERB_LINE ? 1
              : v == HIPR_VERB_QUAD || v == HIPR_VERB_CONIC ? 2
              : v == HIPR_VERB_CUBIC ? 3 : 0;
      pi -= n;
      switch (v) {
        case HIPR_VERB_MOVE:
          return;
        case HIPR_VERB_LINE:
          LineTo(src.pts[pi]);
          break;
        case HIPR_VERB_QUAD:
          QuadTo(src.pts[pi + 1], src.pts[pi]);
          break;
        case HIPR_VERB_CONIC:
          ConicTo(Vv(src.pts[pi + 1]), Vv(src.pts[pi]), src.weights[--wi]);
          break;
        default:
          break;
      }
    }
  }
};

// ---- skity's 4x4 Matrix restricted to the plane, with glm's
// isIdentity(epsilon) shortcut of Matrix operator* (matrix.cc:452-460).
struct M3 {
  float a = 1, b = 0, tx = 0;  // x' = a x + b y + tx
  float c = 0, d = 1, ty = 0;  // y' = c x + d y + ty
};
bool NearIdentity(const M3& m) {
  const float e = 1.1920929e-07f;
  return std::fabs(m.a - 1) <= e && std::fabs(m.d - 1) <= e && std::fabs(m.b) <= e &&
         std::fabs(m.c) <= e && std::fabs(m.tx) <= e && std::fabs(m.ty) <= e;
}
M3 MMul(const M3& p, const M3& q) {
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
inline V MApply(const M3& m, V p) {
  return V{(m.a * p.x + m.b * p.y) + m.tx, (m.c * p.x + m.d * p.y) + m.ty};
}

struct ConicV {
  V pts[3];
  float w;
};

// Conic::BuildUnitArc (conic.cc:72-167).
int BuildUnitArc(V start, V stop, bool ccw, const M3& user, ConicV dst[5]) {
  float x = Dot(start, stop);
  float y = Cross(start, stop);
  float absY = std::fabs(y);
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
  static const V qp[8] = {{1, 0}, {1, 1}, {0, 1}, {-1, 1}, {-1, 0}, {-1, -1}, {0, -1}, {1, -1}};
  int count = quadrant;
  for (int i = 0; i < count; i++) {
    dst[i].pts[0] = qp[i * 2];
    dst[i].pts[1] = qp[i * 2 + 1];
    dst[i].pts[2] = qp[i * 2 + 2];
    dst[i].w = kRoot2Over2;
  }
  V finalP{x, y};
  V lastQ = qp[quadrant * 2];
  float dot = Dot(lastQ, finalP);
  if (dot < 1) {
    V off{lastQ.x + x, lastQ.y + y};
    float cos_half = std::sqrt((1 + dot) / 2);
    SetLength(&off, off.x, off.y, 1.f / cos_half);
    if (CanNormalize(lastQ.x - off.x, lastQ.y - off.y)) {
      dst[count].pts[0] = lastQ;
      dst[count].pts[1] = off;
      dst[count].pts[2] = finalP;
      dst[count].w = cos_half;
      count++;
    }
  }
  M3 rot;
  rot.a = start.x;
  rot.c = start.y;
  rot.b = -start.y;
  rot.d = start.x;
  if (ccw) {
    M3 flip;
    flip.d = -1.f;
    rot = MMul(rot, flip);
  }
  M3 m = MMul(user, rot);
  for (int i = 0; i < count; i++)
    for (int j = 0; j < 3; j++) dst[i].pts[j] = MApply(m, dst[i].pts[j]);
  return count;
}

bool IsClockwise(V before, V after) { return before.x * after.y > before.y * after.x; }

enum AngleType { kNearly180, kSharp, kShallow, kNearlyLine };
AngleType Dot2AngleType(float dot) {
  if (dot >= 0) return NearlyZero(1.f - dot) ? kNearlyLine : kShallow;
  return NearlyZero(1.f + dot) ? kNearly180 : kSharp;
}

void HandleInnerJoin(PathB* inner, V pivot, V after) {
  inner->LineTo(Pt{pivot.x, pivot.y});
  inner->LineTo(Pt{pivot.x - after.x, pivot.y - after.y});
}

void Join(int join, PathB* outer, PathB* inner, V before_un, V pivot, V after_un,
          float radius, float miter_limit) {
  if (join == HIPR_JOIN_BEVEL) {
    V after = after_un * radius;
    if (!IsClockwise(before_un, after_un)) {
      std::swap(outer, inner);
      after = V{-after.x, -after.y};
    }
    outer->LineTo(Pt{pivot.x + after.x, pivot.y + after.y});
    HandleInnerJoin(inner, pivot, after);
    return;
  }
  if (join == HIPR_JOIN_ROUND) {
    float dot = Dot(before_un, after_un);
    AngleType at = Dot2AngleType(dot);
    if (at == kNearlyLine) return;
    V before = before_un, after = after_un;
    bool ccw = false;
    if (!IsClockwise(before, after)) {
      std::swap(outer, inner);
      before = V{-before.x, -before.y};
      after = V{-after.x, -after.y};
      ccw = true;
    }
    M3 T, S;
    T.tx = pivot.x;
    T.ty = pivot.y;
    S.a = radius;
    S.d = radius;
    M3 user = MMul(T, S);
    ConicV conics[5];
    int count = BuildUnitArc(before, after, ccw, user, conics);
    if (count > 0) {
      for (int i = 0; i < count; i++) outer->ConicTo(conics[i].pts[1], conics[i].pts[2], conics[i].w);
      after = after * radius;
      HandleInnerJoin(inner, pivot, after);
    }
    return;
  }
  // miter
  float dot = Dot(before_un, after_un);
  AngleType at = Dot2AngleType(dot);
  V before = before_un, after = after_un, mid;
  if (at == kNearlyLine) return;
  if (at == kNearly180) {
    after = after * radius;
    outer->LineTo(Pt{pivot.x + after.x, pivot.y + after.y});
    HandleInnerJoin(inner, pivot, after);
    return;
  }
  if (!IsClockwise(before, after)) {
    std::swap(outer, inner);
    before = V{-before.x, -before.y};
    after = V{-after.x, -after.y};
  }
  if (0 == dot && miter_limit >= kSqrt2) {
    mid = (before + after) * radius;
  } else {
    V out_dir = before + after;
    float k = 2.f / (out_dir.x * out_dir.x + out_dir.y * out_dir.y);
    mid = V{k * out_dir.x, k * out_dir.y} * radius;
    if (std::sqrt(mid.x * mid.x + mid.y * mid.y) >= miter_limit * radius) {
      after = after * radius;
      outer->LineTo(Pt{pivot.x + after.x, pivot.y + after.y});
      HandleInnerJoin(inner, pivot, after);
      return;
    }
  }
  outer->LineTo(Pt{pivot.x + mid.x, pivot.y + mid.y});
  after = after * radius;
  outer->LineTo(Pt{pivot.x + after.x, pivot.y + after.y});
  HandleInnerJoin(inner, pivot, after);
}

void Cap(int cap, PathB* path, V pivot, V normal, V stop) {
  if (cap == HIPR_CAP_BUTT) {
    path->LineTo(P(stop));
  } else if (cap == HIPR_CAP_SQUARE) {
    V par = RotateCW(normal);
    path->LineTo(Pt{pivot.x + normal.x + par.x, pivot.y + normal.y + par.y});
    path->LineTo(Pt{pivot.x - normal.x + par.x, pivot.y - normal.y + par.y});
    path->LineTo(P(stop));
  } else {
    V par = RotateCW(normal);
    V pc = pivot + par;
    path->ConicTo(pc + normal, pc, kRoot2Over2);
    path->ConicTo(pc - normal, stop, kRoot2Over2);
  }
}

bool QuadInLine(const V q[3]) {
  float pt_max = -1;
  int outer1 = 0, outer2 = 0;
  for (int index = 0; index < 2; ++index) {
    for (int inner = index + 1; inner < 3; ++inner) {
      V d = q[inner] - q[index];
      float m = std::max(std::fabs(d.x), std::fabs(d.y));
      if (pt_max < m) {
        outer1 = index;
        outer2 = inner;
        pt_max = m;
      }
    }
  }
  int mid = outer1 ^ outer2 ^ 3;
  const float slop = 0.000005f;
  float line_slop = pt_max * pt_max * slop;
  return PtToLine(q[mid], q[outer1], q[outer2]) <= line_slop;
}

float FindQuadMaxCurvature(const V s[3]) {
  float Ax = s[1].x - s[0].x, Ay = s[1].y - s[0].y;
  float Bx = s[0].x - s[1].x - s[1].x + s[2].x;
  float By = s[0].y - s[1].y - s[1].y + s[2].y;
  float numer = -(Ax * Bx + Ay * By);
  float denom = Bx * Bx + By * By;
  if (denom < 0) {
    numer = -numer;
    denom = -denom;
  }
  if (numer <= 0) return 0;
  if (numer >= denom) return 1;
  return numer / denom;
}

bool SharpAngle(const V q[3]) {
  V smaller = q[1] - q[0];
  V larger = q[1] - q[2];
  float sl = Dot(smaller, smaller), ll = Dot(larger, larger);
  if (sl > ll) {
    std::swap(smaller, larger);
    ll = sl;
  }
  if (!SetLength(&smaller, smaller.x, smaller.y, ll)) return false;
  return Dot(smaller, larger) > 0;
}

int IntersectQuadRay(const V line[2], const V quad[3], float roots[2]) {
  V vec = line[1] - line[0];
  float r[3];
  for (int n = 0; n < 3; ++n)
    r[n] = (quad[n].y - line[0].y) * vec.x - (quad[n].x - line[0].x) * vec.y;
  float A = r[2], B = r[1], C = r[0];
  A += C - 2 * B;
  B -= C;
  return FindUnitQuadRoots(A, 2 * B, C, roots);
}

struct QuadConstruct {
  V quad[3];
  V tan_start, tan_end;
  float start_t, mid_t, end_t;
  bool start_set, end_set;
  bool Init(float s, float e) {
    start_t = s;
    mid_t = (s + e) * 0.5f;
    end_t = e;
    start_set = end_set = false;
    return start_t < mid_t && mid_t < end_t;
  }
  bool InitWithStart(const QuadConstruct* p) {
    if (!Init(p->start_t, p->mid_t)) return false;
    quad[0] = p->quad[0];
    tan_start = p->tan_start;
    start_set = true;
    return true;
  }
  bool InitWithEnd(const QuadConstruct* p) {
    if (!Init(p->mid_t, p->end_t)) return false;
    quad[2] = p->quad[2];
    tan_end = p->tan_end;
    end_set = true;
    return true;
  }
};

class Stroker {
 public:
  Stroker(float radius, float miter_limit, int cap, int join)
      : radius_(radius), miter_limit_(miter_limit), cap_(cap), join_(join) {
    inv_res_ = 1.f / (1.f * 4);
    inv_res_sq_ = inv_res_ * inv_res_;
    if (join == HIPR_JOIN_MITER && miter_limit <= 1.f) join_ = HIPR_JOIN_BEVEL;
  }
  void MoveTo(V pt) {
    if (seg_count_ > 0) FinishContour(false);
    seg_count_ = 0;
    first_pt_ = prev_pt_ = pt;
    join_completed_ = false;
  }
  void LineTo(V curr) {
    bool teeny = EqualsWithinTol(prev_pt_, curr, kNearlyZero * inv_res_);
    if (teeny && cap_ == HIPR_CAP_BUTT) return;
    if (teeny && join_completed_) return;
    V normal, unit;
    if (!PreJoinTo(curr, &normal, &unit)) return;
    outer_.LineTo(Pt{curr.x + normal.x, curr.y + normal.y});
    inner_.LineTo(Pt{curr.x - normal.x, curr.y - normal.y});
    PostJoinTo(curr, normal, unit);
  }
  void QuadTo(V c, V p2) {
    V quad[3] = {prev_pt_, c, p2};
    V reduction;
    int rt = CheckQuadLinear(quad, &reduction);
    if (rt == 0 || rt == 1) {  // point or line
      LineTo(p2);
      return;
    }
    if (rt == 3) {  // degenerate
      LineTo(reduction);
      int save = join_;
      join_ = HIPR_JOIN_ROUND;
      LineTo(p2);
      join_ = save;
      return;
    }
    V nAB, uAB, nBC, uBC;
    if (!PreJoinTo(c, &nAB, &uAB)) {
      LineTo(p2);
      return;
    }
    QuadConstruct qp;
    stroke_type_ = 1;
    qp.Init(0, 1);
    QuadStroke(quad, &qp);
    stroke_type_ = -1;
    qp.Init(0, 1);
    QuadStroke(quad, &qp);
    if (!SetNormalUnit(quad[1], quad[2], &nBC, &uBC)) {
      nBC = nAB;
      uBC = uAB;
    }
    PostJoinTo(p2, nBC, uBC);
  }
  void Close() { FinishContour(true); }
  void Done(PathB* dst) {
    FinishContour(false);
    *dst = outer_;
  }

 private:
  float radius_, miter_limit_;
  int cap_, join_;
  float inv_res_, inv_res_sq_;
  PathB outer_, inner_;
  V first_normal_{0, 0}, prev_normal_{0, 0}, first_unit_{0, 0}, prev_unit_{0, 0};
  V first_pt_{0, 0}, prev_pt_{0, 0}, first_outer_pt_{0, 0};
  int seg_count_ = -1;
  bool join_completed_ = false;
  int depth_ = 0;
  int stroke_type_ = 1;

  bool SetNormalUnit(V before, V after, V* normal, V* unit) {
    if (!SetNormal(unit, (after.x - before.x) * 1.f, (after.y - before.y) * 1.f)) return false;
    *unit = RotateCCW(*unit);
    *normal = *unit * radius_;
    return true;
  }
  bool PreJoinTo(V curr, V* normal, V* unit) {
    float px = prev_pt_.x, py = prev_pt_.y;
    if (!SetNormalUnit(prev_pt_, curr, normal, unit)) {
      if (cap_ == HIPR_CAP_BUTT) return false;
      *normal = V{radius_, 0};
      *unit = V{1, 0};
    }
    if (seg_count_ == 0) {
      first_normal_ = *normal;
      first_unit_ = *unit;
      first_outer_pt_ = V{px + normal->x, py + normal->y};
      outer_.MoveTo(P(first_outer_pt_));
      inner_.MoveTo(Pt{px - normal->x, py - normal->y});
    } else {
      Join(join_, &outer_, &inner_, prev_unit_, prev_pt_, *unit, radius_, miter_limit_);
    }
    return true;
  }
  void PostJoinTo(V curr, V normal, V unit) {
    join_completed_ = true;
    prev_pt_ = curr;
    prev_unit_ = unit;
    prev_normal_ = normal;
    seg_count_ += 1;
  }
  void FinishContour(bool close) {
    if (seg_count_ > 0) {
      if (close) {
        Join(join_, &outer_, &inner_, prev_unit_, prev_pt_, first_unit_, radius_, miter_limit_);
        outer_.Close();
        V pt = inner_.LastPt();
        outer_.MoveTo(P(pt));
        outer_.ReversePathTo(inner_);
        outer_.Close();
      } else {
        V pt = inner_.LastPt();
        Cap(cap_, &outer_, prev_pt_, prev_normal_, pt);
        outer_.ReversePathTo(inner_);
        Cap(cap_, &outer_, first_pt_, V{-first_normal_.x, -first_normal_.y}, first_outer_pt_);
        outer_.Close();
      }
    }
    inner_.Reset();
    seg_count_ = -1;
  }
  // CheckQuadLinear: 0 point, 1 line, 2 quad, 3 degenerate (stroke.cc:815-841)
  int CheckQuadLinear(const V q[3], V* reduction) {
    V p1c = q[1] - q[0];
    bool d1 = !CanNormalize(p1c.x, p1c.y);
    V cp2 = q[2] - q[1];
    bool d2 = !CanNormalize(cp2.x, cp2.y);
    if (d1 & d2) return 0;
    if (d1 | d2) return 1;
    if (!QuadInLine(q)) return 2;
    float t = FindQuadMaxCurvature(q);
    if (t == 0 || t == 1) return 1;
    QuadEvalAt(q, t, reduction, nullptr);
    return 3;
  }
  void SetRayPts(V tpt, V* dxy, V* on, V* tangent) const {
    if (!SetLength(dxy, dxy->x, dxy->y, radius_)) {
      dxy->x = radius_;
      dxy->y = 0;
    }
    float flip = (float)stroke_type_;
    on->x = tpt.x + flip * dxy->y;
    on->y = tpt.y - flip * dxy->x;
    if (tangent) {
      tangent->x = on->x + dxy->x;
      tangent->y = on->y + dxy->y;
    }
  }
  void QuadPerpRay(const V q[3], float t, V* tpt, V* on, V* tangent) const {
    V dxy;
    QuadEvalAt(q, t, tpt, &dxy);
    if (dxy.x == 0 && dxy.y == 0) dxy = q[2] - q[0];
    SetRayPts(*tpt, &dxy, on, tangent);
  }
  // 0 split, 1 degenerate, 2 quad
  int IntersectRay(QuadConstruct* qp) const {
    V start = qp->quad[0], end = qp->quad[2];
    V aLen = qp->tan_start - start, bLen = qp->tan_end - end;
    float denom = Cross(aLen, bLen);
    if (denom == 0 || denom == INFINITY) return 1;
    V ab0 = start - end;
    float numerA = Cross(bLen, ab0);
    float numerB = Cross(aLen, ab0);
    if ((numerA >= 0) == (numerB >= 0)) {
      float d1 = PtToLine(start, end, qp->tan_end);
      float d2 = PtToLine(end, start, qp->tan_start);
      if (std::max(d1, d2) <= inv_res_sq_) return 1;
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
  bool PtInQuadBounds(const V q[3], V pt) const {
    float xmin = std::min(std::min(q[0].x, q[1].x), q[2].x);
    if (pt.x + inv_res_ < xmin) return false;
    float xmax = std::max(std::max(q[0].x, q[1].x), q[2].x);
    if (pt.x - inv_res_ > xmax) return false;
    float ymin = std::min(std::min(q[0].y, q[1].y), q[2].y);
    if (pt.y + inv_res_ < ymin) return false;
    float ymax = std::max(std::max(q[0].y, q[1].y), q[2].y);
    if (pt.y - inv_res_ > ymax) return false;
    return true;
  }
  int StrokeCloseEnough(const V stroke[3], const V ray[2], QuadConstruct* qp) const {
    V mid = QuadEval(stroke, 0.5f);
    if (DistSqd(ray[0], mid) <= inv_res_ * inv_res_) {
      return SharpAngle(qp->quad) ? 0 : 2;
    }
    if (!PtInQuadBounds(stroke, ray[0])) return 0;
    float roots[2];
    int rc = IntersectQuadRay(ray, stroke, roots);
    if (rc != 1) return 0;
    V qpt = QuadEval(stroke, roots[0]);
    float error = inv_res_ * (1.f - std::fabs(roots[0] - 0.5f) * 2);
    if (DistSqd(ray[0], qpt) <= error * error) {
      return SharpAngle(qp->quad) ? 0 : 2;
    }
    return 0;
  }
  int CompareQuadQuad(const V q[3], QuadConstruct* qp) {
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
    return StrokeCloseEnough(qp->quad, ray, qp);
  }
  bool QuadStroke(const V q[3], QuadConstruct* qp) {
    int rt = CompareQuadQuad(q, qp);
    if (rt == 2) {
      PathB* path = stroke_type_ == 1 ? &outer_ : &inner_;
      path->QuadTo(P(qp->quad[1]), P(qp->quad[2]));
      return true;
    }
    if (rt == 1) return true;
    if (++depth_ > kRecursiveLimits) return false;
    QuadConstruct half;
    (void)half.InitWithStart(qp);
    if (!QuadStroke(q, &half)) return false;
    (void)half.InitWithEnd(qp);
    if (!QuadStroke(q, &half)) return false;
    --depth_;
    return true;
  }
};

}  // namespace

// Stroke::StrokePath over a quad path (stroke.cc:866-901).  The outline is a
// plain Path (winding fill) holding lines, quads and conics.
PathData StrokeOutline(const PathData& q, const hipr_paint& paint) {
  Stroker st(paint.stroke_width * 0.5f, paint.miter_limit, paint.cap, paint.join);
  // Path::Iter(src, false) over the quad path.
  size_t pi = 0;
  V move_to{0, 0}, last{0, 0};
  bool after_move = false;
  int last_verb = HIPR_VERB_MOVE;
  for (size_t vi = 0; vi < q.verbs.size(); vi++) {
    int v = q.verbs[vi];
    if (v == HIPR_VERB_MOVE) {
      if (vi + 1 == q.verbs.size()) break;
      move_to = Vv(q.pts[pi]);
      last = move_to;
      after_move = true;
      pi++;
      st.MoveTo(move_to);
    } else if (v == HIPR_VERB_LINE) {
      after_move = false;
      last = Vv(q.pts[pi]);
      st.LineTo(last);
      pi++;
      last_verb = HIPR_VERB_LINE;
    } else if (v == HIPR_VERB_QUAD) {
      after_move = false;
      V c = Vv(q.pts[pi]);
      last = Vv(q.pts[pi + 1]);
      st.QuadTo(c, last);
      pi += 2;
      last_verb = HIPR_VERB_QUAD;
    } else if (v == HIPR_VERB_CLOSE) {
      if (!(last == move_to)) {
        // AutoClose emits the closing line first (path.cc:372-389)
        st.LineTo(move_to);
        last_verb = HIPR_VERB_LINE;
        last = move_to;
      }
      st.Close();
    } else {
      // conic / cubic verbs are skipped by StrokePath (stroke.cc:887-890)
      pi += v == HIPR_VERB_CONIC ? 2 : 3;
    }
  }
  (void)after_move;
  PathB out;
  st.Done(&out);
  PathData r = out;
  r.fill = HIPR_FILL_WINDING;
  return r;
}

}  // namespace orc
