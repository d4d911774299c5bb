// hipr_common.hpp -- job/canvas descriptors and wave-level helpers shared by
// the gfx950 kernels.
#ifndef SKITY_HIP_COMMON_HPP_
#define SKITY_HIP_COMMON_HPP_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hipr_geom.hpp"

namespace hipr {

constexpr int kWave = 64;
constexpr int kTile = 16;  // coverage tiles are 16x16 pixels

// Geometry of a set of paths in skity's Path layout (structure of arrays).
struct GeomSrc {
  const uint8_t* verbs;
  const float* pts;  // x,y pairs
  const float* weights;
};

enum JobKind : int32_t { kJobFill = 0, kJobStroke = 1, kJobImage = 2 };

// One rasterisation: a fill or stroke outline of a path, drawn onto canvas
// `canvas` with the canvas translation (tx, ty) applied after the quad
// conversion; or (kJobImage) the blurred layer `image` composited over its
// filter bounds.
struct Job {
  int32_t path;         // index of the DrawPath call (paint)
  int32_t canvas;
  int32_t kind;
  int32_t fill;         // HIPR_FILL_*
  int32_t geom;         // 0: scene arrays, 1: stroke outline arrays
  int32_t verb_begin, verb_end;
  int32_t pt_begin, w_begin;
  float tx, ty;
  int32_t image;        // layer canvas index for kJobImage, else -1
  int32_t paint;        // index into the device paint table
  int32_t pad;
};

// Per-job outputs of the flatten stage.
struct JobGeo {
  float l, t, r, b;     // bounds of the quad path (after translation)
  int32_t edges;        // edges built
  int32_t overflow;     // 1 if more edges than capacity
};

// Per-job outputs of the walk stage.
struct JobCov {
  int32_t left, top, right, bottom;   // floored/ceiled bounds (pixels)
  int32_t tiles;        // coverage tiles written
  int32_t planes;       // coverage planes written (>= tiles)
  int32_t max_active;
  int32_t flags;        // bit0 plane overflow, bit1 active overflow,
                        // bit2 stacked direct spans, bit3 open interval
};

struct Canvas {
  uint8_t* pixels;      // premultiplied RGBA8, row-major
  int32_t width, height;
  int32_t tiles_x, tiles_y;
  int32_t tile_begin;   // first global tile index
  int32_t job_begin, job_end;  // jobs drawn on this canvas, in order
};

// ---- wave helpers ------------------------------------------------------------
__device__ __forceinline__ int Lane() { return threadIdx.x & (kWave - 1); }

__device__ __forceinline__ int WaveInclusiveSum(int v) {
  int lane = Lane();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    int t = __shfl_up(v, o, kWave);
    if (lane >= o) v += t;
  }
  return v;
}
__device__ __forceinline__ int WaveInclusiveMax(int v) {
  int lane = Lane();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    int t = __shfl_up(v, o, kWave);
    if (lane >= o) v = v > t ? v : t;
  }
  return v;
}
__device__ __forceinline__ int WaveMin(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    int t = __shfl_xor(v, o, kWave);
    v = v < t ? v : t;
  }
  return v;
}
__device__ __forceinline__ int WaveMax(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    int t = __shfl_xor(v, o, kWave);
    v = v > t ? v : t;
  }
  return v;
}
__device__ __forceinline__ float WaveMinF(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, kWave));
  return v;
}
__device__ __forceinline__ float WaveMaxF(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, kWave));
  return v;
}
__device__ __forceinline__ int WaveSum(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

}  // namespace hipr

#endif  // SKITY_HIP_COMMON_HPP_
