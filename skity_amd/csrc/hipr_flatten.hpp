// hipr_flatten.hpp -- stage (a): curve flattening and fixed-point edge build.
//
// One wavefront per raster job.  Lane i of a 64-verb chunk owns verb i and
// produces the edges of that segment exactly as Stroke::QuadPath
// (src/geometry/stroke.cc:914-962) followed by SWEdgeBuilder::BuildEdges
// (src/render/sw/sw_edge.cc:299-336) would, in the same order:
//   * point / conic-weight offsets of every verb come from wave prefix sums;
//   * the quad path's start point of a segment is the previous segment's
//     quad-path end point (for a cubic that is its own evaluation at t = 1,
//     cubic.cc:37-51), or the contour's move point;
//   * every contour is closed by one line from its last quad-path point to
//     its move point (Path::Iter::AutoClose, path.cc:372-389, and
//     PathEdgeIter's close line, path_priv.hpp:121-166, give the same edge);
//   * SWEdge::CanBeIgnored (sw_edge.cc:72-88) rejects no edge of a job
//     without a clip: its scan bounds are the floored/ceiled bounds of the
//     same points, so every edge lies inside.  A job with a clip first
//     takes its quad path's bounds in a pass of its own, then drops the
//     edges whose rows lie wholly above or below its scan bounds (full
//     SetLine / SetQuad in both passes, so the counts agree).
// Edges are compacted with a wave exclusive scan of per-lane edge counts.
#ifndef SKITY_HIP_FLATTEN_HPP_
#define SKITY_HIP_FLATTEN_HPP_

#include "hipr_common.hpp"

namespace hipr {

__device__ __forceinline__ int VerbPoints(int v) {
  return v == HIPR_VERB_MOVE || v == HIPR_VERB_LINE ? 1
         : v == HIPR_VERB_QUAD || v == HIPR_VERB_CONIC ? 2
         : v == HIPR_VERB_CUBIC ? 3 : 0;
}
__device__ __forceinline__ bool IsSegment(int v) {
  return v >= HIPR_VERB_LINE && v <= HIPR_VERB_CUBIC;
}

// Visitor interface: Emit(quad?, points) for each quad-path piece in order.
struct EdgeSink {
  EdgeRec* out;      // job's edge array (write mode) or nullptr (count mode)
  int32_t cap;
  int32_t pos;       // next edge index within the job
  float minx, miny, maxx, maxy;
  bool clipped;      // CanBeIgnored against [scan_y0, scan_y1)
  Fixed scan_y0, scan_y1;
  __device__ bool Keep(Fixed y0, Fixed y1) const {
    return !clipped || !(y0 >= scan_y1 || y1 <= scan_y0);
  }
  __device__ void Bound(P2 p) {
    minx = fminf(minx, p.x);
    maxx = fmaxf(maxx, p.x);
    miny = fminf(miny, p.y);
    maxy = fmaxf(maxy, p.y);
  }
  __device__ void Line(P2 a, P2 b, bool write) {
    if (!write && !clipped) {
      Fixed y0 = SnapY(ScalarToFixedAA(a.y)), y1 = SnapY(ScalarToFixedAA(b.y));
      if (FixedToFDot6(y1 - y0) != 0) pos++;
      return;
    }
    EdgeRec e;
    if (SetLine(&e, a, b) && Keep(e.upper_y, e.lower_y)) {
      if (write && pos < cap) out[pos] = e;
      pos++;
    }
  }
  __device__ void MonoQuad(const P2 q[3], bool write) {
    if (!write && !clipped) {
      int y0 = (int)(q[0].y * 256.f), y2 = (int)(q[2].y * 256.f);
      if (((y0 + 32) >> 6) != ((y2 + 32) >> 6)) pos++;
      return;
    }
    EdgeRec e;
    Fixed first_y = 0;
    if (SetQuad(&e, q, &first_y) && Keep(first_y, e.q_last_y)) {
      if (write && pos < cap) out[pos] = e;
      pos++;
    }
  }
  __device__ void Quad(P2 a, P2 c, P2 b, bool write) {
    P2 src[3] = {a, c, b};
    P2 mono[5];
    int n = ChopQuadAtYExtrema(src, mono);
    MonoQuad(&mono[0], write);
    if (n == 1) MonoQuad(&mono[2], write);
  }
};

// Visits the quad-path pieces of segment `verb` whose source points start at
// pts[pi] (the source start point is pts[pi - 1]) and whose quad-path start
// point is qstart.  Returns the quad-path end point.
__device__ P2 VisitSegment(int verb, const float* pts, int pi, const float* w,
                           int wi, P2 qstart, const Affine& M, EdgeSink* sink,
                           bool write, bool bound) {
  auto P = [&](int i) { return P2{pts[2 * i], pts[2 * i + 1]}; };
  // the canvas matrix on quad-path points (Path::CopyWithMatrix)
  auto T = [&](P2 p) { return M.Map(p); };
  P2 qs = T(qstart);
  if (verb == HIPR_VERB_LINE) {
    P2 e = P(pi);
    P2 te = T(e);
    if (bound) sink->Bound(te);
    sink->Line(qs, te, write);
    return e;
  }
  if (verb == HIPR_VERB_QUAD) {
    P2 c = T(P(pi)), e = P(pi + 1);
    P2 te = T(e);
    if (bound) {
      sink->Bound(c);
      sink->Bound(te);
    }
    sink->Quad(qs, c, te, write);
    return e;
  }
  if (verb == HIPR_VERB_CONIC) {
    P2 src[3] = {P(pi - 1), P(pi), P(pi + 1)};
    P2 q[5];
    ConicToQuads(src, w[wi], q);
    P2 q1 = T(q[1]), q2 = T(q[2]), q3 = T(q[3]), q4 = T(q[4]);
    if (bound) {
      sink->Bound(q1);
      sink->Bound(q2);
      sink->Bound(q3);
      sink->Bound(q4);
    }
    sink->Quad(qs, q1, q2, write);
    sink->Quad(q2, q3, q4, write);
    return q[4];
  }
  // cubic
  P2 src[4] = {P(pi - 1), P(pi), P(pi + 1), P(pi + 2)};
  CubicQ cq;
  CubicSetup(src, &cq);
  P2 prev = qs;
  P2 end = qstart;
  for (int i = 0; i < cq.n; i++) {
    P2 ctrl, e;
    CubicQuad(cq, i, &ctrl, &e);
    P2 tc = T(ctrl), te = T(e);
    if (bound) {
      sink->Bound(tc);
      sink->Bound(te);
    }
    sink->Quad(prev, tc, te, write);
    prev = te;
    end = e;
  }
  return end;
}

// Quad-path end point of a segment, without emitting anything.
__device__ __forceinline__ P2 SegmentQEnd(int verb, const float* pts, int pi) {
  auto P = [&](int i) { return P2{pts[2 * i], pts[2 * i + 1]}; };
  if (verb == HIPR_VERB_CUBIC) {
    P2 src[4] = {P(pi - 1), P(pi), P(pi + 1), P(pi + 2)};
    CubicQ cq;
    CubicSetup(src, &cq);
    return CubicEval(cq, 1.0f);
  }
  return P(pi + VerbPoints(verb) - 1);
}

// mode 0: count edges and bounds; mode 1: also write edge records.  Work
// item i is job list[i], or job i without a list.
template <int kMode>
__global__ void __launch_bounds__(64)
    FlattenKernel(const Job* __restrict__ jobs, int n_items, GeomSrc scene,
                  GeomSrc outline, const int64_t* __restrict__ edge_base,
                  const int32_t* __restrict__ edge_cap, EdgeRec* __restrict__ edges,
                  JobGeo* __restrict__ geo, const int32_t* __restrict__ list) {
  if ((int)blockIdx.x >= n_items) return;
  const int j = list ? list[blockIdx.x] : (int)blockIdx.x;
  const Job job = jobs[j];
  JobGeo g;
  g.cl = job.clip[0];
  g.ct = job.clip[1];
  g.cr = job.clip[2];
  g.cb = job.clip[3];
  if (job.kind == kJobRect) {
    if (Lane() == 0) {
      g.l = g.t = g.r = g.b = 0.f;
      g.edges = 0;
      g.overflow = 0;
      geo[j] = g;
    }
    return;
  }
  const GeomSrc src = job.geom == 0 ? scene : outline;
  const uint8_t* verbs = src.verbs + job.verb_begin;
  const float* pts = src.pts + 2 * (int64_t)job.pt_begin;
  const float* w = src.weights + job.w_begin;
  int nv = job.verb_end - job.verb_begin;
  const bool write = kMode == 1;
  EdgeSink sink;
  sink.out = write ? edges + edge_base[j] : nullptr;
  sink.cap = write ? edge_cap[j] : 0;
  sink.minx = sink.miny = INFINITY;
  sink.maxx = sink.maxy = -INFINITY;
  sink.clipped = !ClipIsCull(job.clip);
  sink.scan_y0 = kMinS32;
  sink.scan_y1 = kMaxS32;
  int lane = Lane();
  // pass 0 (clipped jobs only): the quad path's bounds; pass 1: edges
  for (int pass = sink.clipped ? 0 : 1; pass < 2; pass++) {
  if (pass == 1 && sink.clipped) {
    // RastePath's scan bounds: the path bounds intersected with the clip
    // (Rect::Intersect), floored / ceiled; CanBeIgnored snaps their top
    // and bottom as SetLine snaps y
    float l = WaveMinF(sink.minx), t = WaveMinF(sink.miny);
    float r = WaveMaxF(sink.maxx), b = WaveMaxF(sink.maxy);
    float il = fmaxf(l, job.clip[0]), ir = fminf(r, job.clip[2]);
    float it = fmaxf(t, job.clip[1]), ib = fminf(b, job.clip[3]);
    if (il < ir && it < ib) {
      sink.scan_y0 = SnapY(ScalarToFixedAA(floorf(it)));
      sink.scan_y1 = SnapY(ScalarToFixedAA(ceilf(ib)));
    } else {
      sink.scan_y0 = kMaxS32;  // no scan rows: every edge is dropped
      sink.scan_y1 = kMinS32;
    }
    sink.minx = sink.miny = INFINITY;
    sink.maxx = sink.maxy = -INFINITY;
  }
  const bool bounds_only = pass == 0;
  int pt_carry = 0, w_carry = 0, edge_carry = 0;
  int move_carry = -1;  // point index of the last move before this chunk
  for (int base = 0; base < nv; base += kWave) {
    int k = base + lane;
    int v = k < nv ? verbs[k] : -1;
    int np = v >= 0 ? VerbPoints(v) : 0;
    int nw = v == HIPR_VERB_CONIC ? 1 : 0;
    int pi_incl = WaveInclusiveSum(np);
    int wi_incl = WaveInclusiveSum(nw);
    int pi = pt_carry + pi_incl - np;  // first point of verb k
    int wi = w_carry + wi_incl - nw;
    int move_here = v == HIPR_VERB_MOVE ? pi : -1;
    int move_pt = WaveInclusiveMax(move_here);
    if (move_pt < 0) move_pt = move_carry;
    // Bounds: every move point except a trailing move (Path::ComputePtBounds
    // and Path::Iter's kDone on a final move).
    if (v == HIPR_VERB_MOVE && k != nv - 1) {
      sink.Bound(job.m.Map(P2{pts[2 * pi], pts[2 * pi + 1]}));
    }
    bool seg = v >= 0 && IsSegment(v);
    P2 qstart{0.f, 0.f}, mv{0.f, 0.f};
    bool last = false;
    if (seg) {
      int pv = verbs[k - 1];
      if (pv == HIPR_VERB_MOVE) {
        qstart = P2{pts[2 * (pi - 1)], pts[2 * (pi - 1) + 1]};
      } else {
        qstart = SegmentQEnd(pv, pts, pi - VerbPoints(pv));
      }
      last = (k + 1 >= nv) || !IsSegment(verbs[k + 1]);
      mv = job.m.Map(P2{pts[2 * move_pt], pts[2 * move_pt + 1]});
    }
    // Pass 1 (both modes): count this lane's edges and extend the bounds.
    sink.pos = 0;
    const bool clip_now = sink.clipped;
    if (bounds_only) sink.clipped = false;  // the cheap count; nothing kept
    if (seg) {
      P2 qend = VisitSegment(v, pts, pi, w, wi, qstart, job.m, &sink, false, true);
      if (last && !bounds_only) sink.Line(job.m.Map(qend), mv, false);
    }
    sink.clipped = clip_now;
    int mine = bounds_only ? 0 : sink.pos;
    int incl = WaveInclusiveSum(mine);
    // Pass 2 (write mode): build the records at the compacted positions.
    if (write && seg && !bounds_only) {
      sink.pos = edge_carry + incl - mine;
      P2 qend = VisitSegment(v, pts, pi, w, wi, qstart, job.m, &sink, true, false);
      if (last) sink.Line(job.m.Map(qend), mv, true);
    }
    edge_carry += __builtin_amdgcn_readlane(incl, kWave - 1);
    pt_carry += __builtin_amdgcn_readlane(pi_incl, kWave - 1);
    w_carry += __builtin_amdgcn_readlane(wi_incl, kWave - 1);
    move_carry = __builtin_amdgcn_readlane(move_pt, kWave - 1);
  }
  if (pass == 1) {
  float l = WaveMinF(sink.minx), t = WaveMinF(sink.miny);
  float r = WaveMaxF(sink.maxx), b = WaveMaxF(sink.maxy);
  if (lane == 0) {
    g.l = l;
    g.t = t;
    g.r = r;
    g.b = b;
    g.edges = edge_carry;
    g.overflow = write && edge_carry > edge_cap[j] ? 1 : 0;
    geo[j] = g;
  }
  }
  }
}

}  // namespace hipr

#endif  // SKITY_HIP_FLATTEN_HPP_
