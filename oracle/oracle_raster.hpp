// oracle_raster.hpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of skity's software DrawPath path, used solely as the
// parity checker by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg.  Nothing in the product library links, loads or calls it.
//
// Every function cites the reference file:line whose behaviour it restates.
// It is compiled with -ffp-contract=off so that float expressions round
// exactly as the reference's x86-64 build (no FMA contraction) does.
#ifndef ORACLE_RASTER_HPP_
#define ORACLE_RASTER_HPP_

#include <cstdint>
#include <vector>

#include "../include/skity_hip.h"

namespace orc {

typedef int32_t Fixed;  // 16.16 (src/render/sw/sw_subpixel.hpp:19)
typedef int32_t FDot6;  // 26.6

struct Pt {
  float x, y;
};

// A path in the reference Path layout: verbs, points, conic weights
// (include/skity/graphic/path.hpp:587-592).
struct PathData {
  std::vector<uint8_t> verbs;
  std::vector<Pt> pts;
  std::vector<float> weights;
  int fill = HIPR_FILL_WINDING;
  // last_move_to_index_ bookkeeping of the Path builder
  // (src/graphic/path.cc:494-505,837-859,1327-1339).
  int last_move_index = -1;  // ~index when the contour was closed
  bool has_last_move = false;

  void MoveTo(Pt p);
  void LineTo(Pt p);
  void QuadTo(Pt p1, Pt p2);
  void Close();
  void InjectMoveToIfNeeded();
};

// Edge record as built by SWEdge::SetLine / SWQuadEdge::SetQuad
// (src/render/sw/sw_edge.hpp:18-84).  Also the 16-int exchange layout used
// by hipr_debug_build_edges.
struct Edge {
  Fixed x, y, dx, dy, upper_x, upper_y, lower_y;
  int8_t curve_count;
  uint8_t curve_shift;
  int8_t winding;
  bool is_quad;
  Fixed qx, qy, qdx, qdy, qddx, qddy, q_first_y, q_last_x, q_last_y;
  Fixed snapped_x, snapped_y;
  // linked-list slots for the active edge list walk
  int prev, next;
};

struct Span {
  int32_t x, y, len, cover;
  // 1 if the span comes from SpanBuilder's row accumulator (its flush), 0 if
  // it went straight to RealSpanBuilder (sw_raster.cc:19-136).
  int32_t accumulated;
};

// Convert a scene path into PathData.
PathData ScenePath(const hipr_scene* s, int p);

// Stroke::QuadPath (src/geometry/stroke.cc:910-962).
PathData QuadPath(const PathData& src);

// SWEdgeBuilder::BuildEdges (src/render/sw/sw_edge.cc:299-336) for the
// given integer scan bounds.  Edges are returned in emission order.
void BuildEdges(const PathData& qpath, float top, float bottom,
                std::vector<Edge>* edges);

// SWRaster::RastePath (src/render/sw/sw_raster.cc:731-786) with an affine
// canvas matrix m = {sx, kx, tx, ky, sy, ty} (nullptr: identity), applied
// after the quad conversion as Path::CopyWithMatrix does
// (src/graphic/path.cc:1274-1295, src/geometry/matrix.cc:462-465).
// Returns the spans in emission order; *bounds receives the floored/ceiled
// path bounds.  clip = {l, t, r, b}: the canvas's global clip bounds the scan
// bounds are intersected with (nullptr: kMaxCullRect, +-1e9).
void RastePath(const PathData& path, std::vector<Span>* spans,
               float bounds_ltrb[4], std::vector<Edge>* edges_out,
               const float* m = nullptr, const float* clip = nullptr);

// Path bounds (Path::ComputePtBounds, src/graphic/path.cc:1347-1360).
bool PathBounds(const PathData& p, float ltrb[4]);

}  // namespace orc

#endif  // ORACLE_RASTER_HPP_
