// hipr_mlp.hpp -- the learned pre-sizer: a 16-64-64-2 multilayer perceptron
// that predicts, per raster job, the number of edges the flatten stage builds
// and the number of coverage tile planes the walk writes, so that the exact
// counting passes (and their host synchronisations) can be skipped.
//
// Arithmetic: FP16 weights and activations as MFMA operands, FP32
// accumulation (v_mfma_f32_16x16x16_f16), FP32 master weights, Adam.
// Everything is computed in the transposed form Z^T = W^T * X^T so that one
// layer's 16x16 output block (MFMA D layout: lane l holds rows 4*(l/16)+i,
// column l%16) is, unchanged, the next layer's B operand (lane l holds
// k = 4*(l/16)+i, n = l%16).  A wavefront processes 16 samples per tile.
//
// Parameter vector (float32, kMlpTotal entries):
//   W1[16][64] b1[64] W2[64][64] b2[64] W3[64][2] b3[2]   (trained)
//   feature mean[16], feature scale[16], target mean[2], target scale[2]
#ifndef SKITY_HIP_MLP_HPP_
#define SKITY_HIP_MLP_HPP_

#include "hipr_common.hpp"

namespace hipr {

constexpr int kMlpIn = 16, kMlpHid = 64, kMlpOut = 2;
constexpr int kW1 = 0;
constexpr int kB1 = kW1 + kMlpIn * kMlpHid;
constexpr int kW2 = kB1 + kMlpHid;
constexpr int kB2 = kW2 + kMlpHid * kMlpHid;
constexpr int kW3 = kB2 + kMlpHid;
constexpr int kB3 = kW3 + kMlpHid * kMlpOut;
constexpr int kMlpTrained = kB3 + kMlpOut;  // 5378
constexpr int kFMu = kMlpTrained;
constexpr int kFSig = kFMu + kMlpIn;
constexpr int kYMu = kFSig + kMlpIn;
constexpr int kYSig = kYMu + kMlpOut;
constexpr int kMlpTotal = kYSig + kMlpOut;  // 5414

// FP16 operand blocks (256 halves each), in the order they are consumed.
constexpr int kOpL1 = 0;                 // 4 blocks: W1^T, output block mb
constexpr int kOpL2 = kOpL1 + 4;         // 16 blocks: W2^T (mb, kb)
constexpr int kOpL3 = kOpL2 + 16;        // 4 blocks: W3^T, k block kb
constexpr int kOpB3 = kOpL3 + 4;         // 4 blocks: W3 (backward), block mb
constexpr int kOpB2 = kOpB3 + 4;         // 16 blocks: W2 (backward) (mb, kb)
constexpr int kOpBlocks = kOpB2 + 16;    // 44

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float4_t Mfma16(half4_t a, half4_t b, float4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}

// Builds the FP16 operand blocks from the FP32 parameters.  One thread per
// (block, lane, element): 44 * 64 * 4 threads.
__global__ void MlpPackKernel(const float* __restrict__ prm, _Float16* __restrict__ ops) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= kOpBlocks * 256) return;
  int blk = t >> 8, l = (t >> 2) & 63, i = t & 3;
  int r4 = 4 * (l >> 4) + i, c16 = l & 15;
  float v = 0.f;
  if (blk < kOpL2) {  // A[m = h][k = f] = W1[f][h]
    int mb = blk - kOpL1;
    v = prm[kW1 + r4 * kMlpHid + mb * 16 + c16];
  } else if (blk < kOpL3) {  // A[m = h][k = k] = W2[k][h]
    int mb = (blk - kOpL2) >> 2, kb = (blk - kOpL2) & 3;
    v = prm[kW2 + (kb * 16 + r4) * kMlpHid + mb * 16 + c16];
  } else if (blk < kOpB3) {  // A[m = o][k = k] = W3[k][o]
    int kb = blk - kOpL3;
    v = c16 < kMlpOut ? prm[kW3 + (kb * 16 + r4) * kMlpOut + c16] : 0.f;
  } else if (blk < kOpB2) {  // A[m = h][k = o] = W3[h][o]
    int mb = blk - kOpB3;
    v = r4 < kMlpOut ? prm[kW3 + (mb * 16 + c16) * kMlpOut + r4] : 0.f;
  } else {  // A[m = k][kk = h] = W2[k][h]
    int mb = (blk - kOpB2) >> 2, kb = (blk - kOpB2) & 3;
    v = prm[kW2 + (mb * 16 + c16) * kMlpHid + kb * 16 + r4];
  }
  ops[t] = (_Float16)v;
}

struct MlpTile {
  float4_t z1[4], z2[4], z3;  // pre-activations, D layout
  half4_t h1[4], h2[4];       // activations as B operands
};

__device__ __forceinline__ half4_t LoadOp(const _Float16* ops, int blk, int lane) {
  return *(const half4_t*)(ops + blk * 256 + lane * 4);
}
__device__ __forceinline__ half4_t ToHalf4(float4_t v) {
  half4_t h;
  h[0] = (_Float16)v[0];
  h[1] = (_Float16)v[1];
  h[2] = (_Float16)v[2];
  h[3] = (_Float16)v[3];
  return h;
}

// Normalised features of 16 samples as the B operand (lane l: sample l%16,
// features 4*(l/16)+i).  Samples at or beyond n are zero.
__device__ __forceinline__ half4_t LoadFeatures(const float* __restrict__ feats, int64_t s0,
                                                int64_t n, const float* __restrict__ prm,
                                                int lane) {
  int64_t s = s0 + (lane & 15);
  int f0 = 4 * (lane >> 4);
  half4_t x;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    float v = 0.f;
    if (s < n) v = (feats[s * kMlpIn + f0 + i] - prm[kFMu + f0 + i]) * prm[kFSig + f0 + i];
    x[i] = (_Float16)v;
  }
  return x;
}

__device__ void MlpForwardTile(const _Float16* __restrict__ ops, const float* __restrict__ prm,
                               half4_t x, int lane, MlpTile* T) {
  int r4 = 4 * (lane >> 4);
#pragma unroll
  for (int mb = 0; mb < 4; mb++) {
    float4_t c;
#pragma unroll
    for (int i = 0; i < 4; i++) c[i] = prm[kB1 + mb * 16 + r4 + i];
    c = Mfma16(LoadOp(ops, kOpL1 + mb, lane), x, c);
    T->z1[mb] = c;
    float4_t a;
#pragma unroll
    for (int i = 0; i < 4; i++) a[i] = c[i] > 0.f ? c[i] : 0.f;
    T->h1[mb] = ToHalf4(a);
  }
#pragma unroll
  for (int mb = 0; mb < 4; mb++) {
    float4_t c;
#pragma unroll
    for (int i = 0; i < 4; i++) c[i] = prm[kB2 + mb * 16 + r4 + i];
#pragma unroll
    for (int kb = 0; kb < 4; kb++) c = Mfma16(LoadOp(ops, kOpL2 + mb * 4 + kb, lane), T->h1[kb], c);
    T->z2[mb] = c;
    float4_t a;
#pragma unroll
    for (int i = 0; i < 4; i++) a[i] = c[i] > 0.f ? c[i] : 0.f;
    T->h2[mb] = ToHalf4(a);
  }
  float4_t c;
#pragma unroll
  for (int i = 0; i < 4; i++) c[i] = (r4 + i) < kMlpOut ? prm[kB3 + r4 + i] : 0.f;
#pragma unroll
  for (int kb = 0; kb < 4; kb++) c = Mfma16(LoadOp(ops, kOpL3 + kb, lane), T->h2[kb], c);
  T->z3 = c;
}

// Inference: out[s][o] = predicted target (de-normalised), s < n.
__global__ void __launch_bounds__(64)
    MlpInferKernel(const _Float16* __restrict__ ops, const float* __restrict__ prm,
                   const float* __restrict__ feats, int64_t n, float* __restrict__ out) {
  int lane = Lane();
  for (int64_t s0 = (int64_t)blockIdx.x * 16; s0 < n; s0 += (int64_t)gridDim.x * 16) {
    MlpTile T;
    MlpForwardTile(ops, prm, LoadFeatures(feats, s0, n, prm, lane), lane, &T);
    int64_t s = s0 + (lane & 15);
    if (lane < 16 && s < n) {
#pragma unroll
      for (int o = 0; o < kMlpOut; o++) out[s * kMlpOut + o] = T.z3[o] * prm[kYSig + o] + prm[kYMu + o];
    }
  }
}

// Forward + backward over `per_wave` samples per wavefront (a multiple of
// 16).  Each wave writes its partial gradient (of the mean squared error in
// normalised target space, already divided by n_total) to
// grads[wave][kMlpTrained] and its summed squared error to loss[wave].
__global__ void __launch_bounds__(64)
    MlpGradKernel(const _Float16* __restrict__ ops, const float* __restrict__ prm,
                  const float* __restrict__ feats, const float* __restrict__ targets,
                  int64_t n, int per_wave, float inv_n_total, float* __restrict__ grads,
                  float* __restrict__ loss) {
  __shared__ float sX[16][kMlpIn];
  __shared__ float sH1[16][kMlpHid];
  __shared__ float sH2[16][kMlpHid];
  __shared__ float sD1[16][kMlpHid];
  __shared__ float sD2[16][kMlpHid];
  __shared__ float sD3[16][kMlpOut];
  const int lane = Lane();
  const int r4 = 4 * (lane >> 4), c16 = lane & 15;
  float gW1[kMlpIn], gW2[kMlpHid], gW3[kMlpOut];
  float gB1 = 0.f, gB2 = 0.f, gB3 = 0.f;
#pragma unroll
  for (int k = 0; k < kMlpIn; k++) gW1[k] = 0.f;
#pragma unroll
  for (int k = 0; k < kMlpHid; k++) gW2[k] = 0.f;
#pragma unroll
  for (int o = 0; o < kMlpOut; o++) gW3[o] = 0.f;
  float sq = 0.f;
  int64_t begin = (int64_t)blockIdx.x * per_wave;
  int64_t end = begin + per_wave < n ? begin + per_wave : n;
  for (int64_t s0 = begin; s0 < end; s0 += 16) {
    half4_t x = LoadFeatures(feats, s0, end, prm, lane);
    MlpTile T;
    MlpForwardTile(ops, prm, x, lane, &T);
    // dL/dz3 (x n_total): 2 (z3 - y_normalised) for valid samples
    int64_t s = s0 + c16;
    float4_t d3 = {0.f, 0.f, 0.f, 0.f};
    if (lane < 16 && s < end) {
#pragma unroll
      for (int o = 0; o < kMlpOut; o++) {
        float yn = (targets[s * kMlpOut + o] - prm[kYMu + o]) / prm[kYSig + o];
        float e = T.z3[o] - yn;
        sq += e * e;
        d3[o] = 2.f * e;
      }
    }
    half4_t b3 = ToHalf4(d3);
    float4_t dz2[4], dz1[4];
    half4_t b2[4];
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
      float4_t c = {0.f, 0.f, 0.f, 0.f};
      c = Mfma16(LoadOp(ops, kOpB3 + mb, lane), b3, c);
#pragma unroll
      for (int i = 0; i < 4; i++) c[i] = T.z2[mb][i] > 0.f ? c[i] : 0.f;
      dz2[mb] = c;
      b2[mb] = ToHalf4(c);
    }
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
      float4_t c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < 4; kb++) c = Mfma16(LoadOp(ops, kOpB2 + mb * 4 + kb, lane), b2[kb], c);
#pragma unroll
      for (int i = 0; i < 4; i++) c[i] = T.z1[mb][i] > 0.f ? c[i] : 0.f;
      dz1[mb] = c;
    }
    // stage the tile (sample-major) for the weight-gradient products
#pragma unroll
    for (int i = 0; i < 4; i++) sX[c16][r4 + i] = (float)x[i];
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        int h = mb * 16 + r4 + i;
        sH1[c16][h] = (float)T.h1[mb][i];
        sH2[c16][h] = (float)T.h2[mb][i];
        sD1[c16][h] = dz1[mb][i];
        sD2[c16][h] = dz2[mb][i];
      }
    }
    if (lane < 16) {
#pragma unroll
      for (int o = 0; o < kMlpOut; o++) sD3[c16][o] = d3[o];
    }
    __syncthreads();
    // lane owns column h = lane of W1, W2 and row k = lane of W3
    for (int t = 0; t < 16; t++) {
      float d1 = sD1[t][lane], d2 = sD2[t][lane];
#pragma unroll
      for (int k = 0; k < kMlpIn; k++) gW1[k] += sX[t][k] * d1;
#pragma unroll
      for (int k = 0; k < kMlpHid; k++) gW2[k] += sH1[t][k] * d2;
      float h2 = sH2[t][lane];
#pragma unroll
      for (int o = 0; o < kMlpOut; o++) gW3[o] += h2 * sD3[t][o];
      gB1 += d1;
      gB2 += d2;
      if (lane < kMlpOut) gB3 += sD3[t][lane];
    }
    __syncthreads();
  }
  float* g = grads + (int64_t)blockIdx.x * kMlpTrained;
#pragma unroll
  for (int k = 0; k < kMlpIn; k++) g[kW1 + k * kMlpHid + lane] = gW1[k] * inv_n_total;
#pragma unroll
  for (int k = 0; k < kMlpHid; k++) g[kW2 + k * kMlpHid + lane] = gW2[k] * inv_n_total;
#pragma unroll
  for (int o = 0; o < kMlpOut; o++) g[kW3 + lane * kMlpOut + o] = gW3[o] * inv_n_total;
  g[kB1 + lane] = gB1 * inv_n_total;
  g[kB2 + lane] = gB2 * inv_n_total;
  if (lane < kMlpOut) g[kB3 + lane] = gB3 * inv_n_total;
  float tot = sq;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o, kWave);
  if (lane == 0) loss[blockIdx.x] = tot;
}

// Sums the per-wave gradients (and losses) and, if lr > 0, applies one Adam
// step to the trained parameters.  One thread per trained parameter.
// Thread 0 also writes the batch's mean loss (sum of the wave losses times
// inv_count) to *loss_out when loss_out is not null.
__global__ void MlpAdamKernel(float* __restrict__ prm, float* __restrict__ m,
                              float* __restrict__ v, const float* __restrict__ grads,
                              int n_waves, float* __restrict__ grad_sum, float lr, float b1,
                              float b2, float eps, float corr1, float corr2,
                              const float* __restrict__ wave_loss, float inv_count,
                              float* __restrict__ loss_out) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0 && loss_out) {
    float l = 0.f;
    for (int w = 0; w < n_waves; w++) l += wave_loss[w];
    *loss_out = l * inv_count;
  }
  if (p >= kMlpTrained) return;
  float g = 0.f;
  for (int w = 0; w < n_waves; w++) g += grads[(int64_t)w * kMlpTrained + p];
  if (grad_sum) grad_sum[p] = g;
  if (lr > 0.f) {
    float mm = b1 * m[p] + (1.f - b1) * g;
    float vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    prm[p] -= lr * (mm / corr1) / (sqrtf(vv / corr2) + eps);
  }
}

// ---- features ---------------------------------------------------------------
// Per raster job, from the source path (and the paint for strokes):
//  0-3 log1p of the line / quad / conic / cubic counts
//  4   log1p contours
//  5,6 log1p bbox width, height (control points, after the job translation)
//  7   log1p of the 16x16 tiles the bbox covers on the canvas
//  8   log1p control-polygon length
//  9   log1p summed |turning angle| of the control polygon (radians)
//  10  log1p stroke width (0 for fills)
//  11  1 for strokes
//  12  1 for even-odd fills
//  13  log1p of the largest cubic control-polygon deviation
//  14  canvas matrix scale (1: the scene's matrix is the identity)
//  15  1 (bias feature)
struct FeatureParams {
  const Job* jobs;
  int32_t n_jobs;
  GeomSrc scene;
  const hipr_paint* paints;
  const Canvas* canvases;
  float* out;  // [n_jobs][16]
};

__global__ void __launch_bounds__(64) FeaturesKernel(FeatureParams P) {
  int j = blockIdx.x;
  if (j >= P.n_jobs) return;
  const int lane = Lane();
  const Job job = P.jobs[j];
  float* out = P.out + (int64_t)j * kMlpIn;
  if (job.kind == kJobRect) {
    if (lane < kMlpIn) out[lane] = 0.f;
    return;
  }
  const uint8_t* verbs = P.scene.verbs + job.verb_begin;
  const float* pts = P.scene.pts + 2 * (int64_t)job.pt_begin;
  int nv = job.verb_end - job.verb_begin;
  int cnt[4] = {0, 0, 0, 0}, contours = 0;
  float minx = INFINITY, miny = INFINITY, maxx = -INFINITY, maxy = -INFINITY;
  float len = 0.f, turn = 0.f, dev = 0.f;
  int pt_carry = 0;
  for (int base = 0; base < nv; base += kWave) {
    int k = base + lane;
    int v = k < nv ? verbs[k] : -1;
    int np = v == HIPR_VERB_MOVE || v == HIPR_VERB_LINE ? 1
             : v == HIPR_VERB_QUAD || v == HIPR_VERB_CONIC ? 2
             : v == HIPR_VERB_CUBIC ? 3 : 0;
    int incl = WaveInclusiveSum(np);
    int pi = pt_carry + incl - np;
    if (v == HIPR_VERB_MOVE) contours++;
    if (v >= HIPR_VERB_LINE && v <= HIPR_VERB_CUBIC) {
      cnt[v - 1]++;
      // control polygon of the segment, from the previous point
      float px = pts[2 * (pi - 1)], py = pts[2 * (pi - 1) + 1];
      float pdx = 0.f, pdy = 0.f;
      for (int q = 0; q < np; q++) {
        float x = pts[2 * (pi + q)], y = pts[2 * (pi + q) + 1];
        float dx = x - px, dy = y - py;
        len += sqrtf(dx * dx + dy * dy);
        if (q > 0) {
          float cr = pdx * dy - pdy * dx, dt = pdx * dx + pdy * dy;
          turn += fabsf(atan2f(cr, dt));
        }
        pdx = dx;
        pdy = dy;
        px = x;
        py = y;
      }
      if (v == HIPR_VERB_CUBIC) {
        float x0 = pts[2 * (pi - 1)], y0 = pts[2 * (pi - 1) + 1];
        float ex = (3.f * pts[2 * (pi + 1)] - pts[2 * (pi + 2)]) - (3.f * pts[2 * pi] - x0);
        float ey = (3.f * pts[2 * (pi + 1) + 1] - pts[2 * (pi + 2) + 1]) -
                   (3.f * pts[2 * pi + 1] - y0);
        dev = fmaxf(dev, sqrtf(ex * ex + ey * ey));
      }
    }
    for (int q = 0; q < np; q++) {
      P2 d = job.m.Map(P2{pts[2 * (pi + q)], pts[2 * (pi + q) + 1]});
      float x = d.x, y = d.y;
      minx = fminf(minx, x);
      maxx = fmaxf(maxx, x);
      miny = fminf(miny, y);
      maxy = fmaxf(maxy, y);
    }
    pt_carry += __shfl(incl, kWave - 1, kWave);
  }
  for (int c = 0; c < 4; c++) cnt[c] = WaveSum(cnt[c]);
  contours = WaveSum(contours);
  minx = WaveMinF(minx);
  miny = WaveMinF(miny);
  maxx = WaveMaxF(maxx);
  maxy = WaveMaxF(maxy);
  for (int o = 32; o > 0; o >>= 1) {
    len += __shfl_xor(len, o, kWave);
    turn += __shfl_xor(turn, o, kWave);
    dev = fmaxf(dev, __shfl_xor(dev, o, kWave));
  }
  if (lane == 0) {
    const Canvas cv = P.canvases[job.canvas];
    bool ok = minx <= maxx && miny <= maxy;
    float w = ok ? maxx - minx : 0.f, h = ok ? maxy - miny : 0.f;
    float tiles = 0.f;
    if (ok) {
      float x0 = fmaxf(floorf(minx), 0.f), x1 = fminf(ceilf(maxx), (float)cv.width);
      float y0 = fmaxf(floorf(miny), 0.f), y1 = fminf(ceilf(maxy), (float)cv.height);
      if (x0 < x1 && y0 < y1) {
        int c0 = (int)x0 >> 4, c1 = ((int)x1 - 1) >> 4, r0 = (int)y0 >> 4, r1 = ((int)y1 - 1) >> 4;
        tiles = (float)((c1 - c0 + 1) * (r1 - r0 + 1));
      }
    }
    bool stroke = job.kind == kJobStroke;
    float sw = stroke ? P.paints[job.path].stroke_width : 0.f;
    out[0] = log1pf((float)cnt[0]);
    out[1] = log1pf((float)cnt[1]);
    out[2] = log1pf((float)cnt[2]);
    out[3] = log1pf((float)cnt[3]);
    out[4] = log1pf((float)contours);
    out[5] = log1pf(w);
    out[6] = log1pf(h);
    out[7] = log1pf(tiles);
    out[8] = log1pf(len);
    out[9] = log1pf(turn);
    out[10] = log1pf(sw);
    out[11] = stroke ? 1.f : 0.f;
    out[12] = (!stroke && job.fill == HIPR_FILL_EVEN_ODD) ? 1.f : 0.f;
    out[13] = log1pf(dev);
    out[14] = 1.f;
    out[15] = 1.f;
  }
}

}  // namespace hipr

#endif  // SKITY_HIP_MLP_HPP_
