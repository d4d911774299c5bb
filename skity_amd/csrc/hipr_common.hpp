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
// Widest canvas (and mask-filter layer): the walk keeps one LDS row of the
// visible columns, 4 bytes per pixel.
constexpr int kMaxCanvasSide = 16384;

// Geometry of a set of paths in skity's Path layout (structure of arrays).
struct GeomSrc {
  const uint8_t* verbs;
  const float* pts;  // x,y pairs
  const float* weights;
};

// kJobRect: coverage 255 over an integer rectangle of the canvas (the
// paint's L, T, R, B), which is what RastePath gives an axis-aligned
// rectangle with integer sides under the identity: a DrawPaint, or the
// composite of a blurred layer drawn without a canvas matrix.  It is not
// walked.
enum JobKind : int32_t { kJobFill = 0, kJobStroke = 1, kJobRect = 2 };

// An affine canvas matrix applied to quad-path points as Path::CopyWithMatrix
// does (src/graphic/path.cc:1274-1295): glm's mat4 * vec4(x, y, 0, 1) adds the
// column products pairwise, (sx x + kx y) + (0 + tx), each operation rounded.
struct Affine {
  float sx, kx, tx, ky, sy, ty;
  __host__ __device__ P2 Map(P2 p) const {
    return P2{(sx * p.x + kx * p.y) + (0.f + tx), (ky * p.x + sy * p.y) + (0.f + ty)};
  }
};

// One rasterisation: a fill or stroke outline of a path, drawn onto canvas
// `canvas` through matrix `m` applied after the quad conversion; or
// (kJobRect) a rectangle of full coverage.
struct Job {
  int32_t path;         // index of the draw (paint)
  int32_t canvas;
  int32_t kind;
  int32_t fill;         // HIPR_FILL_*
  int32_t geom;         // 0: scene arrays, 1: stroke outline arrays
  int32_t verb_begin, verb_end;
  int32_t pt_begin, w_begin;
  Affine m;
  int32_t image;        // layer canvas index of an image paint, else -1
  int32_t paint;        // index into the device paint table
  int32_t flags;        // kJob* bits, set per draw
  // scan clip {l, t, r, b}: the canvas's global clip bounds (kMaxCullRect
  // when the draw has none)
  float clip[4] = {-1e9f, -1e9f, 1e9f, 1e9f};
};
// The job keeps zero-cover direct spans and three planes per tile (its blend
// mode changes pixels under a transparent source on this target).
constexpr int32_t kJobKeepZero = 1;

// Per-job outputs of the flatten stage.
struct JobGeo {
  float l, t, r, b;     // bounds of the quad path (after translation)
  int32_t edges;        // edges built
  int32_t overflow;     // 1 if more edges than capacity
  float cl, ct, cr, cb; // the job's scan clip (Job::clip)
};
constexpr float kCullExtent = 1e9f;  // kMaxCullRect = {-1e9, -1e9, 1e9, 1e9}
__host__ __device__ inline bool ClipIsCull(const float c[4]) {
  return c[0] <= -kCullExtent && c[1] <= -kCullExtent && c[2] >= kCullExtent &&
         c[3] >= kCullExtent;
}

// Per-job outputs of the walk stage.
struct JobCov {
  int32_t left, top, right, bottom;   // floored/ceiled bounds (pixels)
  int32_t tiles;        // coverage tiles written
  int32_t planes;       // coverage planes written (>= tiles)
  int32_t max_active;
  int32_t flags;        // bit0 plane overflow, bit1 active overflow,
                        // bit2 stacked direct spans, bit3 open interval
  int32_t steps;        // walk sub-steps taken
  int32_t time_10ns;    // wall time of the job's walk, in 10 ns units
  int32_t phase[5];     // of which: roles+trapezoids, quad updates, re-sort,
                        // row flushes, edge insertion (10 ns units)
};

// Canvas::format bits: the target is not premultiplied (kOpaque or
// kUnpremul), and its bytes are B, G, R, A.
constexpr int32_t kFormatUnpremul = 1;
constexpr int32_t kFormatBGRA = 2;

struct Canvas {
  uint8_t* pixels;      // RGBA8 (or BGRA8), row-major
  int32_t width, height;
  int32_t tiles_x, tiles_y;
  int32_t tile_begin;   // first global tile index
  int32_t job_begin, job_end;  // jobs drawn on this canvas, in order
  int32_t format;              // kFormat* bits (0: premultiplied RGBA)
  int32_t pad;
};

// ---- wave helpers ------------------------------------------------------------
__device__ __forceinline__ int Lane() { return threadIdx.x & (kWave - 1); }

// Orders the LDS accesses of a one-wavefront workgroup.  The lanes of a wave
// issue their LDS instructions in one stream, so a compiler fence is enough;
// unlike __syncthreads() it does not wait for outstanding global stores.
__device__ __forceinline__ void WaveSync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

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

// ---- DPP forms (GFX9 data-parallel primitives; every lane must be active) --
// Inclusive scans: row_shr 1, 2, 4, 8 inside each row of 16 lanes, then
// row_bcast 15 (into rows 1, 3) and row_bcast 31 (into rows 2, 3).  A lane
// whose source is outside the row keeps the identity passed as `old`.
constexpr int kDppRowShr = 0x110, kDppWaveShl1 = 0x130, kDppWaveShr1 = 0x138,
              kDppBcast15 = 0x142, kDppBcast31 = 0x143;

__device__ __forceinline__ int DppScanSum(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, kDppRowShr | 1, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, kDppRowShr | 2, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, kDppRowShr | 4, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, kDppRowShr | 8, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, kDppBcast15, 0xa, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, kDppBcast31, 0xc, 0xf, false);
  return v;
}
__device__ __forceinline__ int DppMax2(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ int DppScanMax(int v) {
  const int lo = (int)0x80000000;
  v = DppMax2(v, __builtin_amdgcn_update_dpp(lo, v, kDppRowShr | 1, 0xf, 0xf, false));
  v = DppMax2(v, __builtin_amdgcn_update_dpp(lo, v, kDppRowShr | 2, 0xf, 0xf, false));
  v = DppMax2(v, __builtin_amdgcn_update_dpp(lo, v, kDppRowShr | 4, 0xf, 0xf, false));
  v = DppMax2(v, __builtin_amdgcn_update_dpp(lo, v, kDppRowShr | 8, 0xf, 0xf, false));
  v = DppMax2(v, __builtin_amdgcn_update_dpp(lo, v, kDppBcast15, 0xa, 0xf, false));
  v = DppMax2(v, __builtin_amdgcn_update_dpp(lo, v, kDppBcast31, 0xc, 0xf, false));
  return v;
}
// Lane i receives lane i-1 (lane 0 receives `fill`).
__device__ __forceinline__ int DppShiftUp(int v, int fill) {
  return __builtin_amdgcn_update_dpp(fill, v, kDppWaveShr1, 0xf, 0xf, false);
}
// Lane i receives lane i+1 (lane 63 receives `fill`).
__device__ __forceinline__ int DppShiftDown(int v, int fill) {
  return __builtin_amdgcn_update_dpp(fill, v, kDppWaveShl1, 0xf, 0xf, false);
}

// The forms the kernels call: DPP unless HIPR_USE_DPP is 0.
#ifndef HIPR_USE_DPP
#define HIPR_USE_DPP 1
#endif
__device__ __forceinline__ int ScanSum(int v) {
#if HIPR_USE_DPP
  return DppScanSum(v);
#else
  return WaveInclusiveSum(v);
#endif
}
__device__ __forceinline__ int ScanMax(int v) {
#if HIPR_USE_DPP
  return DppScanMax(v);
#else
  return WaveInclusiveMax(v);
#endif
}
__device__ __forceinline__ int ShiftUp(int v, int fill) {
#if HIPR_USE_DPP
  return DppShiftUp(v, fill);
#else
  int t = __shfl_up(v, 1, kWave);
  return Lane() == 0 ? fill : t;
#endif
}
__device__ __forceinline__ int ShiftDown(int v, int fill) {
#if HIPR_USE_DPP
  return DppShiftDown(v, fill);
#else
  int t = __shfl_down(v, 1, kWave);
  return Lane() == kWave - 1 ? fill : t;
#endif
}

// Non-template kernels are defined once, by the library's main translation
// unit; a second unit (hipr_walk_wide.hip) defines HIPR_TEMPLATES_ONLY.
#ifndef HIPR_TEMPLATES_ONLY
// Checks the DPP forms against the shuffle forms: per wave of 64 inputs,
// out[0..5][64] = DppScanSum, WaveInclusiveSum, DppScanMax, WaveInclusiveMax,
// DppShiftUp(fill -7), DppShiftDown(fill -9).
__global__ void __launch_bounds__(64) WaveOpsCheckKernel(const int* in, int* out) {
  const int lane = Lane();
  const int v = in[blockIdx.x * kWave + lane];
  int* o = out + blockIdx.x * 6 * kWave;
  o[0 * kWave + lane] = DppScanSum(v);
  o[1 * kWave + lane] = WaveInclusiveSum(v);
  o[2 * kWave + lane] = DppScanMax(v);
  o[3 * kWave + lane] = WaveInclusiveMax(v);
  o[4 * kWave + lane] = DppShiftUp(v, -7);
  o[5 * kWave + lane] = DppShiftDown(v, -9);
}
#endif  // HIPR_TEMPLATES_ONLY

}  // namespace hipr

#endif  // SKITY_HIP_COMMON_HPP_
