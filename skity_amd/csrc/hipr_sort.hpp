// hipr_sort.hpp -- per-job edge ordering for the walk.
//
// SortEdges (src/render/sw/sw_raster.cc:679-704) orders edges by
// (upper_y, x, dx).  The walk only needs them grouped by upper_y in emission
// order (it orders each group of newly active edges by (x, dx, position)
// itself), so the key is (upper_y, emission index): a bitonic sort of 64-bit
// keys in LDS, followed by a gather of the records into a work copy that the
// walk may mutate (quad edges advance their forward differences in place).
#ifndef SKITY_HIP_SORT_HPP_
#define SKITY_HIP_SORT_HPP_

#include "hipr_common.hpp"

namespace hipr {

constexpr int kSortThreads = 256;
constexpr int kSortLdsKeys = 8192;  // 64 KiB of keys

__device__ __forceinline__ uint64_t EdgeKey(const EdgeRec& e, int idx) {
  // upper_y is a 16.16 multiple of 1/4; bias to make it unsigned.
  return ((uint64_t)((uint32_t)e.upper_y ^ 0x80000000u) << 32) | (uint32_t)idx;
}

// keys: per-job scratch of next_pow2(n) entries (used when n exceeds LDS).
// active_bound (may be null) receives an upper bound of the job's active
// edge list: the largest number of edges whose row range (upper_y to lower_y,
// or to the last step's y for a quad edge) meets one bin of rows (a bin is one pixel row when the job spans at most
// 16384 rows).  The walk keeps exactly the edges with upper_y <= y < lower_y
// active, so this bounds every sub-step's list.
__global__ void __launch_bounds__(kSortThreads)
    SortGatherKernel(const int32_t* __restrict__ job_list, int n_list,
                     const JobGeo* __restrict__ geo, const int64_t* __restrict__ edge_base,
                     const EdgeRec* __restrict__ edges, EdgeRec* __restrict__ wedges,
                     uint64_t* __restrict__ gkeys, int32_t* __restrict__ active_bound) {
  __shared__ uint64_t skeys[kSortLdsKeys];
  int li = blockIdx.x;
  if (li >= n_list) return;
  int j = job_list[li];
  int n = geo[j].edges;
  if (n <= 0) {
    if (active_bound && threadIdx.x == 0) active_bound[j] = 0;
    return;
  }
  int64_t base = edge_base[j];
  const EdgeRec* src = edges + base;
  EdgeRec* dst = wedges + base;
  int np2 = 1;
  while (np2 < n) np2 <<= 1;
  bool in_lds = np2 <= kSortLdsKeys;
  uint64_t* keys = in_lds ? skeys : gkeys + base * 2;  // scratch sized 2n >= np2
  for (int i = threadIdx.x; i < np2; i += blockDim.x) {
    keys[i] = i < n ? EdgeKey(src[i], i) : ~0ull;
  }
  __syncthreads();
  for (int k = 2; k <= np2; k <<= 1) {
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
      for (int i = threadIdx.x; i < np2; i += blockDim.x) {
        int ixj = i ^ jj;
        if (ixj > i) {
          uint64_t a = keys[i], b = keys[ixj];
          bool up = (i & k) == 0;
          if ((a > b) == up) {
            keys[i] = b;
            keys[ixj] = a;
          }
        }
      }
      if (!in_lds) __threadfence_block();
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    int idx = (int)(uint32_t)(keys[i] & 0xFFFFFFFFu);
    dst[i] = src[idx];
  }
  if (!active_bound) return;
  // ---- row-overlap bound (LDS reused as a histogram of row bins) ----
  __syncthreads();
  constexpr int kBins = kSortLdsKeys * 2;
  int32_t* hist = (int32_t*)skeys;
  const JobGeo g = geo[j];
  const int top = (int)floorf(g.t), bottom = (int)ceilf(g.b);
  const int H = bottom > top ? bottom - top : 1;
  const int bin = (H + kBins - 3) / (kBins - 2);  // rows per bin, >= 1
  const int nb = H / bin + 2;                     // <= kBins
  for (int i = threadIdx.x; i < nb; i += blockDim.x) hist[i] = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const EdgeRec& e = src[i];
    // a quad edge stays in the list until its last step (q_last_y)
    Fixed last = EdgeIsQuad(e.packed) && e.q_last_y > e.lower_y ? e.q_last_y : e.lower_y;
    int r0 = (e.upper_y >> 16) - top, r1 = ((last - 1) >> 16) - top;
    r0 = r0 < 0 ? 0 : (r0 >= H ? H - 1 : r0);
    r1 = r1 < r0 ? r0 : (r1 >= H ? H - 1 : r1);
    atomicAdd(&hist[r0 / bin], 1);
    atomicAdd(&hist[r1 / bin + 1], -1);
  }
  __syncthreads();
  // chunked scan: each thread owns a contiguous chunk of bins
  __shared__ int32_t part[kSortThreads];
  const int chunk = (nb + kSortThreads - 1) / kSortThreads;
  const int c0 = threadIdx.x * chunk, c1 = c0 + chunk < nb ? c0 + chunk : nb;
  int sum = 0;
  for (int i = c0; i < c1; i++) sum += hist[i];
  part[threadIdx.x] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int t = 0; t < kSortThreads; t++) {
      int v = part[t];
      part[t] = run;
      run += v;
    }
  }
  __syncthreads();
  int run = part[threadIdx.x], mx = 0;
  for (int i = c0; i < c1; i++) {
    run += hist[i];
    mx = run > mx ? run : mx;
  }
  __syncthreads();
  part[threadIdx.x] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    int m = 0;
    for (int t = 0; t < kSortThreads; t++) m = part[t] > m ? part[t] : m;
    active_bound[j] = m;
  }
}

}  // namespace hipr

#endif  // SKITY_HIP_SORT_HPP_
