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
__global__ void __launch_bounds__(kSortThreads)
    SortGatherKernel(const int32_t* __restrict__ job_list, int n_list,
                     const JobGeo* __restrict__ geo, const int64_t* __restrict__ edge_base,
                     const EdgeRec* __restrict__ edges, EdgeRec* __restrict__ wedges,
                     uint64_t* __restrict__ gkeys) {
  __shared__ uint64_t skeys[kSortLdsKeys];
  int li = blockIdx.x;
  if (li >= n_list) return;
  int j = job_list[li];
  int n = geo[j].edges;
  if (n <= 0) return;
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
}

}  // namespace hipr

#endif  // SKITY_HIP_SORT_HPP_
