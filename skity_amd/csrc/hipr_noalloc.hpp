// hipr_noalloc.hpp -- StackBlur of device-resident images without any device
// allocation, for callers that cannot allocate (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).
//
// The intermediate image of the two separable passes lives in `scratch`, a
// file-scope device array of kNoAllocScratchBytes that every GPU the library
// runs on holds until the process ends.  In host code `scratch` names a host
// stand-in, not device memory: its device address is looked up once per
// device with hipGetSymbolAddress (the device current) and only that address
// reaches kernels and copies.  Every call on a GPU shares that GPU's array:
// calls that overlap on one GPU race.  Nothing in it needs clearing: the row
// pass writes every pixel of the intermediate before the column pass reads it.
#ifndef SKITY_HIP_NOALLOC_HPP_
#define SKITY_HIP_NOALLOC_HPP_

#include <mutex>

#include "hipr_blur_mfma.hpp"

namespace hipr {

// Largest image: 4,194,304 pixels (2048 x 2048 RGBA8) of intermediate.
constexpr int64_t kNoAllocMaxPixels = 4194304;
constexpr size_t kNoAllocScratchBytes = (size_t)kNoAllocMaxPixels * 4;

static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

constexpr int kNoAllocMaxDevices = 64;
static uint8_t* g_scratch_ptr[kNoAllocMaxDevices] = {};
static std::mutex g_scratch_mutex;

// Device address of `scratch` on `device` (which must be current); looked
// up on the first call for that device.
inline hipError_t NoAllocScratch(int device, uint8_t** out) {
  if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = (uint8_t*)p;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

}  // namespace hipr

#endif  // SKITY_HIP_NOALLOC_HPP_
