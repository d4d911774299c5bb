// The 32- and 64-slot walks: paths whose rows keep more than 1024 active
// edges (WalkEdges keeps its active list as a linked list of any length,
// src/render/sw/sw_raster.cc:546-677).  The active list stays in the
// K x 64-lane register layout of the narrower walks; at these widths most of
// it lives in scratch memory, which is the price of a rare path, not of the
// frame.  These jobs are walked unbanded (no schedule pass), so only the
// count and write modes exist.  Compiled as its own translation unit so the
// two large instantiations build beside skity_hip.hip.
#define HIPR_TEMPLATES_ONLY
#include "../../include/skity_hip.h"
#include "hipr_walk.hpp"

namespace hipr {

template <int K>
static hipError_t LaunchTier(int mode, unsigned grid, size_t lds, hipStream_t st,
                             const WalkParams& P) {
  if (mode == kWalkCount)
    hipLaunchKernelGGL((WalkKernel<K, kWalkCount>), dim3(grid), dim3(kWave), lds, st, P);
  else
    hipLaunchKernelGGL((WalkKernel<K, kWalkWrite>), dim3(grid), dim3(kWave), lds, st, P);
  return hipGetLastError();
}

hipError_t LaunchWideWalk(int K, int mode, unsigned grid, size_t lds, hipStream_t st,
                          const WalkParams& P) {
  if (mode != kWalkCount && mode != kWalkWrite) return hipErrorInvalidValue;
  if (K == 32) return LaunchTier<32>(mode, grid, lds, st, P);
  if (K == 64) return LaunchTier<64>(mode, grid, lds, st, P);
  return hipErrorInvalidValue;
}

}  // namespace hipr
