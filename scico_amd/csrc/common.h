// Shared plumbing for the scico_amd C ABI: status codes, the thread-local
// error message, HIP error checks and small device helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/scico_amd.h"

namespace scxr {

void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);

// Return SCXR_ERR_HIP with a message if a HIP call failed.
#define SCXR_HIP(call)                                                        \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      return ::scxr::fail(SCXR_ERR_HIP, "%s failed: %s (%s:%d)", #call,      \
                          hipGetErrorString(e_), __FILE__, __LINE__);         \
  } while (0)

// Check the launch that was just issued.
#define SCXR_LAUNCHED(name)                                                   \
  do {                                                                        \
    hipError_t e_ = hipGetLastError();                                        \
    if (e_ != hipSuccess)                                                     \
      return ::scxr::fail(SCXR_ERR_HIP, "launch of %s failed: %s", name,     \
                          hipGetErrorString(e_));                             \
  } while (0)

#define SCXR_REQUIRE(cond, code, ...)                                         \
  do {                                                                        \
    if (!(cond)) return ::scxr::fail(code, __VA_ARGS__);                     \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned ceil_div(long long a, long long b) {
  return static_cast<unsigned>((a + b - 1) / b);
}

// Device-side predicate: a null pointer means "always run"; otherwise the
// kernel is a no-op while *pred == 0.  Used to replay fixed-length
// iteration sequences (CG with a device-side stopping test) without host
// synchronisation.
__device__ __forceinline__ bool pred_off(const int32_t* pred) {
  return pred != nullptr && *pred == 0;
}

// Correctly rounded float32 quotient d / w given rw = RN64(1 / w).
// RN32(RN64(d * rw)) == RN32(d / w) for float32 d, w: the relative error
// of the float64 product is below 2^-52, while a float32 quotient d / w
// lies at least 2^-49 (relative) away from every float32 rounding
// midpoint, so both round the same way.
__device__ __forceinline__ float div_rn(float d, double rw) {
  return static_cast<float>(static_cast<double>(d) * rw);
}

}  // namespace scxr
