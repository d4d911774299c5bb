// hipr_blur.hpp -- stage (d): the mask-filter blur of a premultiplied layer.
//
// SWStackBlur::Blur (src/render/sw/sw_stack_blur.cc:18-284) computes, per
// channel, out[x] = (S[x] * mul[r]) >> shr[r] with
//   S[x] = sum_{k=-r..r} (r + 1 - |k|) * in[clamp(x + k)],
// first along rows, then along columns of the row result.  (Its column pass
// seeds two running sums from the wrong channel of the top pixel; that only
// matters when the top row is not transparent, which a mask-filter layer's
// top row always is, because the layer extends round(radius) >= 2 pixels
// beyond the drawn path.)
//
// Two exact implementations:
//   * StackBlurLineKernel: one thread per line, running sums in integers;
//   * the MFMA form in hipr_blur_mfma.hpp: S for a block of 16 outputs is a
//     banded-matrix product of fp16 pixels (<= 255, exact) with fp16 weights
//     (<= 255, exact) accumulated in fp32 (every partial sum < 2^24, exact).
#ifndef SKITY_HIP_BLUR_HPP_
#define SKITY_HIP_BLUR_HPP_

#include "hipr_common.hpp"

namespace hipr {

// The published StackBlur normalisation (the reference's tables list exactly
// these values for r in [0, 254]): shr = smallest s with 2^s > 256 (r+1)^2,
// mul = ceil(2^s / (r+1)^2).
__host__ __device__ inline void StackBlurMulShr(int r, uint32_t* mul, uint32_t* shr) {
  uint64_t d = (uint64_t)(r + 1) * (uint64_t)(r + 1);
  int s = 0;
  while (!(((uint64_t)1 << s) > 256 * d)) s++;
  *shr = (uint32_t)s;
  *mul = (uint32_t)((((uint64_t)1 << s) + d - 1) / d);
}

struct BlurJob {
  const uint8_t* src;
  uint8_t* dst;
  int32_t width, height;
  int32_t radius;  // already rounded and clamped to 254, >= 2
};

// MaskFilterOnFilter's post-passes for the blur styles other than kNormal
// (src/effect/mask_filter.cc:60-103), per pixel of a layer: `raw` is the
// layer before the blur, `blurred` the blur's output (updated in place).
struct BlurStyleJob {
  const uint8_t* raw;
  uint8_t* blurred;
  int64_t pixels;
  int32_t style;  // HIPR_BLUR_SOLID / OUTER / INNER
  int32_t pad;
};
__global__ void __launch_bounds__(256)
    BlurStyleKernel(const BlurStyleJob* __restrict__ jobs, int n_jobs) {
  for (int k = blockIdx.y; k < n_jobs; k += gridDim.y) {
  const BlurStyleJob jb = jobs[k];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < jb.pixels;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t raw = ((const uint32_t*)jb.raw)[i];
    uint32_t blur = ((const uint32_t*)jb.blurred)[i];
    uint32_t raw_a = raw >> 24, blur_a = blur >> 24, out = blur;
    if (jb.style == HIPR_BLUR_SOLID) {
      if (raw_a > 0) out = raw;
    } else if (jb.style == HIPR_BLUR_OUTER) {
      if (raw_a > 0 && raw_a >= blur_a) out = 0;
    } else if (jb.style == HIPR_BLUR_INNER) {
      if (raw_a > 0) {
        // AlphaMulQ(blur, a_factor * raw_a * blur_a): the float product,
        // left to right, converted to the integer scale by truncation
        const float a_factor = 1.f / 255.f;
        uint32_t scale = (uint32_t)(a_factor * (float)raw_a * (float)blur_a);
        const uint32_t mask = 0xFF00FF;
        uint32_t rb = ((blur & mask) * scale) >> 8;
        uint32_t ag = ((blur >> 8) & mask) * scale;
        out = (rb & mask) | (ag & ~mask);
      } else {
        out = 0;
      }
    }
    ((uint32_t*)jb.blurred)[i] = out;
  }
  }
}

// pass 0: rows (src -> dst); pass 1: columns (src -> dst).  One thread per
// line; lines of all blur jobs are numbered consecutively (line_begin).
__global__ void __launch_bounds__(256)
    StackBlurLineKernel(const BlurJob* __restrict__ jobs, const int64_t* __restrict__ line_begin,
                        int n_jobs, int64_t total_lines, int pass) {
  int64_t gl = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gl >= total_lines) return;
  int lo = 0, hi = n_jobs - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (line_begin[mid] <= gl) {
      lo = mid;
    } else {
      hi = mid - 1;
    }
  }
  const BlurJob b = jobs[lo];
  int line = (int)(gl - line_begin[lo]);
  int n = pass == 0 ? b.width : b.height;
  int64_t stride = pass == 0 ? 4 : (int64_t)b.width * 4;
  int64_t off = pass == 0 ? (int64_t)line * b.width * 4 : (int64_t)line * 4;
  const uint8_t* src = b.src + off;
  uint8_t* dst = b.dst + off;
  int r = b.radius;
  uint32_t mul, shr;
  StackBlurMulShr(r, &mul, &shr);
  auto at = [&](int i) -> uint32_t {
    i = i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
    return *(const uint32_t*)(src + i * stride);
  };
  // S(x) = sum_k (r+1-|k|) p[x+k];  S(x+1) = S(x) + sum_{k=1..r+1} p[x+k]
  //                                          - sum_{k=-r..0} p[x+k]
  uint64_t S[4] = {0, 0, 0, 0}, in_s[4] = {0, 0, 0, 0}, out_s[4] = {0, 0, 0, 0};
  for (int k = -r; k <= r; k++) {
    uint32_t p = at(k);
    uint32_t w = (uint32_t)(r + 1 - (k < 0 ? -k : k));
    for (int c = 0; c < 4; c++) {
      uint32_t v = (p >> (8 * c)) & 0xFF;
      S[c] += (uint64_t)w * v;
      if (k >= 1) in_s[c] += v;
      if (k <= 0) out_s[c] += v;
    }
  }
  for (int x = 0; x < n; x++) {
    uint32_t o = 0;
    for (int c = 0; c < 4; c++) o |= (uint32_t)(((S[c] * mul) >> shr) & 0xFF) << (8 * c);
    // the output line may alias nothing: src and dst are distinct buffers
    *(uint32_t*)(dst + x * stride) = o;
    uint32_t pin = at(x + r + 1), pout = at(x - r), pmid = at(x + 1);
    for (int c = 0; c < 4; c++) {
      uint64_t vi = (pin >> (8 * c)) & 0xFF, vo = (pout >> (8 * c)) & 0xFF,
               vm = (pmid >> (8 * c)) & 0xFF;
      in_s[c] += vi;            // window (x+1 .. x+r+1]
      S[c] = S[c] + in_s[c] - out_s[c];
      out_s[c] -= vo;           // drop x-r, then add x+1
      out_s[c] += vm;
      in_s[c] -= vm;
    }
  }
}

}  // namespace hipr

#endif  // SKITY_HIP_BLUR_HPP_
