// hipr_blur_mfma.hpp -- StackBlur passes as banded-matrix products on MFMA.
//
// One pass of SWStackBlur (src/render/sw/sw_stack_blur.cc:18-284) is, per
// channel, out[a] = (S[a] * mul[r]) >> shr[r] with
//   S[a] = sum_{d=-r..r} (r + 1 - |d|) * in[clamp(a + d)]
// along a line (a row for the first pass, a column for the second).  For 16
// consecutive outputs a0..a0+15 this is a product with a banded Toeplitz
// matrix: S = sum_j A_j * B_j, where A_j holds 16 input positions
// a0 + 16 j - 16 J + k (J = ceil(r / 16), j = 0..2J) and
// B_j[k][n] = w(16 j - 16 J + k - n), w(d) = max(0, r + 1 - |d|).
//
// Exactness: pixels (<= 255) and weights (<= 255) are exact in fp16; every
// product (< 2^16) and every partial sum (<= 255 (r+1)^2 < 2^24 for
// r <= 254) is exact in fp32, whatever the summation order, so the MFMA sums
// equal the integer sums and the output equals the scalar kernel's.
//
// Tile: 4 lines x 64 outputs per wavefront.  MFMA row m = line * 4 + channel,
// so the D fragment of lane l holds the four channels of one pixel
// (line l / 16, output a0 + l % 16), stored as one RGBA8 word.
#ifndef SKITY_HIP_BLUR_MFMA_HPP_
#define SKITY_HIP_BLUR_MFMA_HPP_

#include "hipr_blur.hpp"
#include "hipr_mlp.hpp"

namespace hipr {

constexpr int kBlurLines = 4;     // lines per tile
constexpr int kBlurOut = 64;      // outputs per line per tile
constexpr int kBlurMaxJ = 16;     // ceil(254 / 16)
constexpr int kBlurSpan = kBlurOut + 2 * 16 * kBlurMaxJ;  // staged positions per line

struct BlurTileJob {
  BlurJob b;
  int64_t tile_begin;  // first global tile of this job
  int32_t line_tiles, out_tiles;
};

// pass 0: rows (src -> dst); pass 1: columns.  One wavefront per tile.
template <int kPass>
__device__ __forceinline__ void StackBlurMfmaTiles(const BlurTileJob* __restrict__ jobs,
                                                   int n_jobs, int64_t total_tiles,
                                                   _Float16* stage) {
  const int lane = Lane();
  for (int64_t t = blockIdx.x; t < total_tiles; t += gridDim.x) {
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (jobs[mid].tile_begin <= t) {
        lo = mid;
      } else {
        hi = mid - 1;
      }
    }
    const BlurTileJob tj = jobs[lo];
    const BlurJob b = tj.b;
    const int64_t lt = t - tj.tile_begin;
    // Row pass: consecutive tiles walk along a row band.  Column pass:
    // consecutive tiles take neighbouring column groups at the same rows, so
    // the workgroups running at one time read (and write) whole memory lines
    // of each row between them and the L2 keeps them.
    const int line0 = (int)(kPass == 0 ? lt / tj.out_tiles : lt % tj.line_tiles) * kBlurLines;
    const int a_tile = (int)(kPass == 0 ? lt % tj.out_tiles : lt / tj.line_tiles) * kBlurOut;
    const int n_lines = kPass == 0 ? b.height : b.width;
    const int len = kPass == 0 ? b.width : b.height;
    const int r = b.radius;
    const int J = (r + 15) >> 4;
    const int span = kBlurOut + 32 * J;
    uint32_t mul, shr;
    StackBlurMulShr(r, &mul, &shr);
    // stage [line][channel][position], positions a_tile - 16 J ..
    for (int e = lane; e < kBlurLines * span; e += kWave) {
      // row pass: lanes take consecutive positions of one row; column pass:
      // four lanes take the tile's four columns of one row (16 contiguous
      // bytes), the wave 16 rows
      int ln = kPass == 0 ? e / span : e % kBlurLines;
      int pos = kPass == 0 ? e % span : e / kBlurLines;
      int line = line0 + ln < n_lines ? line0 + ln : n_lines - 1;
      int a = a_tile - 16 * J + pos;
      a = a < 0 ? 0 : (a >= len ? len - 1 : a);
      int64_t px = kPass == 0 ? (int64_t)line * b.width + a : (int64_t)a * b.width + line;
      uint32_t v = ((const uint32_t*)b.src)[px];
#pragma unroll
      for (int c = 0; c < 4; c++)
        stage[(ln * 4 + c) * span + pos] = (_Float16)(float)((v >> (8 * c)) & 0xFF);
    }
    WaveSync();
    const int am = lane & 15;          // A row m = line * 4 + channel
    const int ak = 4 * (lane >> 4);    // A / B k base
    const int bn = lane & 15;          // B column n
    const _Float16* arow = stage + am * span;
    // The band block B_j is the same for the four 16-output blocks of the
    // tile: build it once per j and feed it to the four accumulators.
    float4_t acc[kBlurOut / 16];
#pragma unroll
    for (int blk = 0; blk < kBlurOut / 16; blk++) acc[blk] = float4_t{0.f, 0.f, 0.f, 0.f};
    for (int j = 0; j <= 2 * J; j++) {
      half4_t w;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        int d = 16 * j - 16 * J + ak + i - bn;
        d = d < 0 ? -d : d;
        w[i] = (_Float16)(float)(d <= r ? r + 1 - d : 0);
      }
#pragma unroll
      for (int blk = 0; blk < kBlurOut / 16; blk++) {
        half4_t a = *(const half4_t*)(arow + 16 * blk + 16 * j + ak);
        acc[blk] = Mfma16(a, w, acc[blk]);
      }
    }
#pragma unroll
    for (int blk = 0; blk < kBlurOut / 16; blk++) {
      // lane: line = lane / 16, output a_tile + 16 blk + lane % 16, channels i
      const int line = line0 + (lane >> 4);
      const int a = a_tile + 16 * blk + (lane & 15);
      if (line < n_lines && a < len) {
        uint32_t o = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
          uint64_t S = (uint64_t)acc[blk][i];  // an exact integer below 2^24
          o |= (uint32_t)(((S * mul) >> shr) & 0xFF) << (8 * i);
        }
        int64_t px = kPass == 0 ? (int64_t)line * b.width + a : (int64_t)a * b.width + line;
        ((uint32_t*)b.dst)[px] = o;
      }
    }
    WaveSync();
  }
}

template <int kPass>
__global__ void __launch_bounds__(64)
    StackBlurMfmaKernel(const BlurTileJob* __restrict__ jobs, int n_jobs, int64_t total_tiles) {
  __shared__ _Float16 stage[kBlurLines * 4 * kBlurSpan];
  StackBlurMfmaTiles<kPass>(jobs, n_jobs, total_tiles, stage);
}

// One image, its job passed by value (no job table in device memory).
template <int kPass>
__global__ void __launch_bounds__(64) StackBlurMfmaOneKernel(BlurTileJob job, int64_t total_tiles) {
  __shared__ _Float16 stage[kBlurLines * 4 * kBlurSpan];
  StackBlurMfmaTiles<kPass>(&job, 1, total_tiles, stage);
}

}  // namespace hipr

#endif  // SKITY_HIP_BLUR_MFMA_HPP_
