// hipr_blit.hpp -- analytic trapezoid coverage of one scanline step.
//
// blit_trapezoid_row / blit_aaa_trapezoid_row and their alpha helpers
// (src/render/sw/sw_raster.cc:138-544), evaluated per pixel in closed form.
// Coverage goes into one LDS delta row, prefix-summed when the pixel row is
// complete.  Each 32-bit entry carries two fields whose prefix sums are
// non-negative and below 2^16, so one integer prefix sum S decodes as
// A = S & 0xFFFF, D = S >> 16:
//   D: spans the reference sends straight to RealSpanBuilder (full-height
//      steps whose edges are not too close); each adds (1 << 12) + cover, so
//      D carries the number of such spans above bit 12;
//   A: spans the reference accumulates in SpanBuilder; saturating addition
//      of non-negative alphas equals min(255, sum), so plain sums suffice.
// Spans with cover 0 blend nothing and are skipped.
#ifndef SKITY_HIP_BLIT_HPP_
#define SKITY_HIP_BLIT_HPP_

#include "hipr_common.hpp"

namespace hipr {

typedef uint8_t Alpha;

constexpr int kDirectCount = 1 << 12;  // one direct span, in the D field

// The row covers the visible pixels [left, left + width) of the path's
// canvas.  A span's two delta entries are clamped into [0, width]: deltas left
// of the canvas fold into entry 0 and deltas right of it into entry `width`,
// which no visible pixel's prefix sum reads, so every visible pixel sums the
// same deltas as over the path's full width.
struct RowSink {
  int32_t* d;
  int32_t left;   // pixel x of delta index 0 (the first visible column)
  int32_t width;  // visible columns; entry `width` is the end-of-row entry
  // Direct spans of cover 0 blend nothing in src-over, and are dropped; a
  // job whose blend mode does change pixels under a transparent source keeps
  // them (they still count as a direct span of the pixel).
  bool keep_zero;
  __device__ __forceinline__ int Clamp(int i) const {
    return i < 0 ? 0 : (i > width ? width : i);
  }
  // kClamp = false: the caller guarantees left <= x and x + len <= left +
  // width (the per-pixel loop of BlitAAARow, which runs over visible
  // columns only), so the two entries need no clamping.
  template <bool kClamp = true>
  __device__ __forceinline__ void Direct(int x, int len, int v) {
    if ((v == 0 && !keep_zero) || len <= 0) return;
    int i = x - left, e = x - left + len;
    if (kClamp) {
      i = Clamp(i);
      e = Clamp(e);
      if (i == e) return;
    }
    int inc = (kDirectCount + v) << 16;
    atomicAdd(&d[i], inc);
    atomicAdd(&d[e], -inc);
  }
  template <bool kClamp = true>
  __device__ __forceinline__ void Acc(int x, int len, int v) {
    if (v == 0 || len <= 0) return;
    int i = x - left, e = x - left + len;
    if (kClamp) {
      i = Clamp(i);
      e = Clamp(e);
      if (i == e) return;
    }
    atomicAdd(&d[i], v);
    atomicAdd(&d[e], -v);
  }
  template <bool kClamp = true>
  __device__ __forceinline__ void Put(bool direct, int x, int len, int v) {
    if (direct) {
      Direct<kClamp>(x, len, v);
    } else {
      Acc<kClamp>(x, len, v);
    }
  }
};

__device__ __forceinline__ Alpha PartialAlphaA(Alpha a, Alpha full) {
  return (Alpha)((a * full) >> 8);
}
__device__ __forceinline__ Alpha TrapezoidToAlpha(Fixed l1, Fixed l2) {
  Fixed area = (l1 + l2) / 2;
  return (Alpha)(area >> 8);
}
__device__ __forceinline__ Alpha PartialTriangleToAlpha(Fixed a, Fixed b) {
  Fixed area = (a >> 11) * (a >> 11) * (b >> 11);
  return (Alpha)((area >> 8) & 0xFF);
}
__device__ __forceinline__ Fixed ApproxIntersection(Fixed l1, Fixed r1, Fixed l2,
                                                    Fixed r2) {
  if (l1 > r1) {
    Fixed t = l1;
    l1 = r1;
    r1 = t;
  }
  if (l2 > r2) {
    Fixed t = l2;
    l2 = r2;
    r2 = t;
  }
  return (Max32(l1, l2) + Min32(r1, r2)) / 2;
}

// Entry i of compute_alpha_below_line(l, r, dY, full) (sw_raster.cc:343-368),
// R = ceil(r) > 0.
__device__ __forceinline__ Alpha AlphaBelow(int i, int R, Fixed l, Fixed r, Fixed dY,
                                            Alpha full) {
  if (R == 1) return PartialAlphaA(TrapezoidToAlpha(l, r), full);
  Fixed last = r - ((R - 1) << 16);
  Fixed last_h = FixedMul(last, dY);
  if (i == R - 1) return (Alpha)(FixedMul(last, last_h) >> 9);
  if (i == 0) return (Alpha)(full - PartialTriangleToAlpha(kFixed1 - l, dY));
  // alpha16 starts at last_h + dY/2 for i = R-2 and grows by dY per step left.
  Fixed a16 = (Fixed)((uint32_t)(last_h + (dY >> 1)) + (uint32_t)(R - 2 - i) * (uint32_t)dY);
  return (Alpha)((a16 >> 8) & 0xFF);
}
// Entry i of compute_alpha_above_line (sw_raster.cc:312-339).
__device__ __forceinline__ Alpha AlphaAbove(int i, int R, Fixed l, Fixed r, Fixed dY,
                                            Alpha full) {
  if (R == 1) return PartialAlphaA((Alpha)(((R << 17) - l - r) >> 9), full);
  Fixed first = kFixed1 - l;
  Fixed first_h = FixedMul(first, dY);
  if (i == 0) return (Alpha)(FixedMul(first, first_h) >> 9);
  if (i == R - 1) {
    Fixed last = r - ((R - 1) << 16);
    return (Alpha)(full - PartialTriangleToAlpha(last, dY));
  }
  Fixed a16 = (Fixed)((uint32_t)(first_h + (dY >> 1)) + (uint32_t)(i - 1) * (uint32_t)dY);
  return (Alpha)(a16 >> 8);
}

// blit_aaa_trapezoid_row (sw_raster.cc:370-455), one pixel at a time.
__device__ void BlitAAARow(RowSink* s, Fixed ul, Fixed ur, Fixed ll, Fixed lr, Fixed lDY,
                           Fixed rDY, Alpha full, bool no_real) {
  int L = FixedFloorToInt(ul), R = FixedCeilToInt(lr);
  int len = R - L;
  bool direct = full == 0xFF && !no_real;
  if (len == 1) {
    Alpha a = TrapezoidToAlpha(ur - ul, lr - ll);
    if (direct) {
      s->Direct(L, 1, a);
    } else {
      s->Acc(L, 1, PartialAlphaA(a, full));
    }
    return;
  }
  int uL = FixedFloorToInt(ul), lL = FixedCeilToInt(ll);
  int uR = FixedFloorToInt(ur), lR = FixedCeilToInt(lr);
  bool left_two = uL + 2 == lL;
  bool right_two = uR + 2 == lR;
  Fixed lfirst = 0, lsecond = 0, rfirst = 0, rsecond = 0;
  Alpha la1 = 0, la2 = 0, ra1 = 0, ra2 = 0;
  if (left_two) {
    lfirst = IntToFixed(uL) + kFixed1 - ul;
    lsecond = ll - ul - lfirst;
    la1 = (Alpha)(full - PartialTriangleToAlpha(lfirst, lDY));
    la2 = PartialTriangleToAlpha(lsecond, lDY);
  }
  if (right_two) {
    rfirst = IntToFixed(uR) + kFixed1 - ur;
    rsecond = lr - ur - rfirst;
    ra1 = PartialTriangleToAlpha(rfirst, rDY);
    ra2 = (Alpha)(full - PartialTriangleToAlpha(rsecond, rDY));
  }
  Fixed bl_l = ul - IntToFixed(uL), bl_r = ll - IntToFixed(uL);
  int bl_R = FixedCeilToInt(bl_r);
  Fixed ab_l = ur - IntToFixed(uR), ab_r = lr - IntToFixed(uR);
  int ab_R = FixedCeilToInt(ab_r);
  // Run of equal full alphas between the two AA regions goes out as one
  // delta span; every other pixel individually.  Pixels outside the visible
  // columns add nothing any visible pixel reads (RowSink), so they are
  // skipped.
  int run_begin = -1, run_len = 0;
  const int i_end = Min32(len, s->left + s->width - L);
  for (int i = Max32(0, s->left - L); i < i_end; i++) {
    int x = L + i;
    Alpha a = full;
    bool touched = false;
    if (left_two) {
      if (i == 0) {
        a = a > la1 ? (Alpha)(a - la1) : 0;
        touched = true;
      } else if (i == 1) {
        a = a > la2 ? (Alpha)(a - la2) : 0;
        touched = true;
      }
    } else if (x >= uL && x < lL) {
      Alpha t = AlphaBelow(x - uL, bl_R, bl_l, bl_r, lDY, full);
      a = a > t ? (Alpha)(a - t) : 0;
      touched = true;
    }
    if (right_two) {
      if (i == len - 2) {
        a = a > ra1 ? (Alpha)(a - ra1) : 0;
        touched = true;
      } else if (i == len - 1) {
        a = a > ra2 ? (Alpha)(a - ra2) : 0;
        touched = true;
      }
    } else if (x >= uR && x < lR) {
      Alpha t = AlphaAbove(x - uR, ab_R, ab_l, ab_r, rDY, full);
      a = a > t ? (Alpha)(a - t) : 0;
      touched = true;
    }
    if (!touched) {
      if (run_begin < 0) run_begin = x;
      run_len++;
    } else {
      if (run_len > 0) {
        s->Put<false>(direct, run_begin, run_len, full);
        run_begin = -1;
        run_len = 0;
      }
      s->Put<false>(direct, x, 1, a);
    }
  }
  if (run_len > 0) s->Put<false>(direct, run_begin, run_len, full);
}

// blit_trapezoid_row (sw_raster.cc:457-544).
__device__ void BlitTrapezoidRow(RowSink* s, Fixed ul, Fixed ur, Fixed ll, Fixed lr,
                                 Fixed lDY, Fixed rDY, Alpha full, bool no_real) {
  if (ul > ur) return;
  if (ll > lr) ll = lr = ApproxIntersection(ul, ll, ur, lr);
  if (ul == ur && ll == lr) return;
  if (ul > ll) {
    Fixed t = ul;
    ul = ll;
    ll = t;
  }
  if (ur > lr) {
    Fixed t = ur;
    ur = lr;
    lr = t;
  }
  bool direct = full == 0xFF && !no_real;
  Fixed join_left = FixedCeilToFixed(ll);
  Fixed join_rite = FixedFloorToFixed(ur);
  if (join_left <= join_rite) {
    if (ul < join_left) {
      int len = FixedCeilToInt(join_left - ul);
      if (len == 1) {
        Alpha a = TrapezoidToAlpha(join_left - ul, join_left - ll);
        if (direct) {
          s->Direct(ul >> 16, 1, a);
        } else {
          s->Acc(ul >> 16, 1, PartialAlphaA(a, full));
        }
      } else if (len == 2) {
        Fixed first = join_left - kFixed1 - ul;
        Fixed second = ll - ul - first;
        Alpha a1 = PartialTriangleToAlpha(first, lDY);
        Alpha a2 = (Alpha)(full - PartialTriangleToAlpha(second, lDY));
        s->Put(direct, ul >> 16, 1, a1);
        s->Put(direct, (ul >> 16) + 1, 1, a2);
      } else {
        BlitAAARow(s, ul, join_left, ll, join_left, lDY, kMaxS32, full, no_real);
      }
    }
    if (join_left < join_rite) {
      s->Put(direct, FixedFloorToInt(join_left), FixedFloorToInt(join_rite - join_left),
             full);
    }
    if (lr > join_rite) {
      int len = FixedCeilToInt(lr - join_rite);
      if (len == 1) {
        Alpha a = TrapezoidToAlpha(ur - join_rite, lr - join_rite);
        if (direct) {
          s->Direct(join_rite >> 16, 1, a);
        } else {
          s->Acc(join_rite >> 16, 1, PartialAlphaA(a, full));
        }
      } else if (len == 2) {
        Fixed first = join_rite + kFixed1 - ur;
        Fixed second = lr - ur - first;
        Alpha a1 = (Alpha)(full - PartialTriangleToAlpha(first, rDY));
        Alpha a2 = PartialTriangleToAlpha(second, rDY);
        s->Put(direct, join_rite >> 16, 1, a1);
        s->Put(direct, (join_rite >> 16) + 1, 1, a2);
      } else {
        BlitAAARow(s, join_rite, ur, join_rite, lr, kMaxS32, rDY, full, no_real);
      }
    }
  } else {
    BlitAAARow(s, ul, ur, ll, lr, lDY, rDY, full, no_real);
  }
}

}  // namespace hipr

#endif  // SKITY_HIP_BLIT_HPP_
