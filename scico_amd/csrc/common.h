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

// min(d / w, 1) for d >= 0 with the quotient rounded as by div_rn.  The
// median of (q, 0, 1) equals min(q, 1) for q >= 0, and the compiler turns
// it into the clamp bit of the float64-to-float32 conversion, so the
// clamp costs no instruction of its own.
__device__ __forceinline__ float div_rn_sat(float d, double rw) {
  return __builtin_amdgcn_fmed3f(div_rn(d, rw), 0.0f, 1.0f);
}

// floor(x) converted to int32 in one instruction (for |x| < 2^31).
__device__ __forceinline__ int cvt_flr_i32(float x) {
  int r;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

// Sum of one double per thread over a 256-thread workgroup (fixed order:
// wave shuffles, then the 4 wave sums left to right).  Thread 0 gets the
// result.  `sh` holds 4 doubles of LDS.
__device__ __forceinline__ double block_sum_f64(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = ((sh[0] + sh[1]) + sh[2]) + sh[3];
  return r;
}

// Every thread of a 256-thread workgroup gets the same fixed-order sum of
// parts[0..n).  Used to finish a reduction redundantly in each consumer
// workgroup instead of in a separate launch.
__device__ __forceinline__ double all_sum_f64(const double* __restrict__ parts, int n,
                                              double* sh) {
  double v = 0.0;
  for (int e = threadIdx.x; e < n; e += 256) v += parts[e];
  v = block_sum_f64(v, sh);
  __syncthreads();
  if (threadIdx.x == 0) sh[0] = v;
  __syncthreads();
  const double r = sh[0];
  __syncthreads();
  return r;
}

// Raw buffer access (gfx9-family buffer word 3): a wave-uniform base and
// row offset live in SGPRs, the lane offset in one VGPR, so a load costs no
// per-lane address arithmetic; loads at or beyond `bytes` return zero.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), static_cast<short>(0), bytes,
                                           0x00020000);
}
__device__ __forceinline__ float4 bload4(__amdgpu_buffer_rsrc_t r, unsigned voff, int soff) {
  return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}
__device__ __forceinline__ float bload1(__amdgpu_buffer_rsrc_t r, unsigned voff, int soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

}  // namespace scxr
