// Host-side helpers shared by the renderer and the visibility model: error
// reporting for the C ABI and growable device buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "gbr.h"

namespace gbr {

// records msg for gbr_last_error() and returns s
gbr_status fail(gbr_status s, const std::string& msg);

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return ::gbr::fail(e_ == hipErrorOutOfMemory ? GBR_ERR_OOM : GBR_ERR_DEVICE,     \
                         std::string(#expr) + ": " + hipGetErrorString(e_));           \
  } while (0)

#define TRY(expr)                    \
  do {                               \
    gbr_status s_ = (expr);          \
    if (s_ != GBR_OK) return s_;     \
  } while (0)

template <class T>
struct DevBuf {
  T* ptr = nullptr;
  size_t cap = 0;
  gbr_status reserve(size_t n, bool keep = false, hipStream_t s = nullptr) {
    if (n <= cap && ptr) return GBR_OK;
    size_t ncap = std::max<size_t>(n + n / 2, 64);
    T* np = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&np), ncap * sizeof(T));
    if (e != hipSuccess)
      return ::gbr::fail(GBR_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    if (keep && ptr && cap) {
      (void)hipMemcpyAsync(np, ptr, cap * sizeof(T), hipMemcpyDeviceToDevice, s);
      (void)hipStreamSynchronize(s);
    }
    if (ptr) (void)hipFree(ptr);
    ptr = np;
    cap = ncap;
    return GBR_OK;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
  }
};

template <class T>
gbr_status upload(DevBuf<T>& b, const std::vector<T>& v, hipStream_t s) {
  TRY(b.reserve(std::max<size_t>(v.size(), 1)));
  if (!v.empty()) HIP_TRY(hipMemcpyAsync(b.ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  return GBR_OK;
}

}  // namespace gbr
