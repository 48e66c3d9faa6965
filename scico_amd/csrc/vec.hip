// Elementwise vector kernels and deterministic float64 reductions used by
// the generic LinearOperator algebra, functionals and solvers.
//
// Reductions: pass 1 writes one float64 partial per workgroup (fixed grid
// of at most RED_BLOCKS workgroups, fixed traversal order), pass 2 sums the
// partials in index order in a single workgroup.  No atomics, so results
// are bitwise reproducible.
#include "common.h"
#include "vec_impl.h"

#include <atomic>
#include <mutex>

using namespace scxr;

namespace scxr {
namespace vec {

__global__ __launch_bounds__(256) void axpby_kernel(int64_t n, float a, const float* __restrict__ x,
                                                    float b, const float* __restrict__ y,
                                                    float* __restrict__ z, const int32_t* pred) {
  if (pred_off(pred)) return;
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < n; e += 256ll * gridDim.x)
    z[e] = __fadd_rn(__fmul_rn(a, x[e]), __fmul_rn(b, y[e]));
}

__global__ __launch_bounds__(256) void add_kernel(int64_t n, const float* __restrict__ x,
                                                  float sign, const float* __restrict__ y,
                                                  float* __restrict__ z, const int32_t* pred) {
  if (pred_off(pred)) return;
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < n; e += 256ll * gridDim.x)
    z[e] = sign > 0.0f ? __fadd_rn(x[e], y[e]) : __fsub_rn(x[e], y[e]);
}

__global__ __launch_bounds__(256) void scale_kernel(int64_t n, float a, const float* __restrict__ x,
                                                    float* __restrict__ z, const int32_t* pred) {
  if (pred_off(pred)) return;
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < n; e += 256ll * gridDim.x)
    z[e] = __fmul_rn(a, x[e]);
}

__global__ __launch_bounds__(256) void fill_kernel(int64_t n, float v, float* __restrict__ z,
                                                   const int32_t* pred) {
  if (pred_off(pred)) return;
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < n; e += 256ll * gridDim.x) z[e] = v;
}

__global__ __launch_bounds__(256) void copy_kernel(int64_t n, const float* __restrict__ x,
                                                   float* __restrict__ z, const int32_t* pred) {
  if (pred_off(pred)) return;
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < n; e += 256ll * gridDim.x) z[e] = x[e];
}

__global__ __launch_bounds__(256) void mul_kernel(int64_t n, const float* __restrict__ x,
                                                  const float* __restrict__ y,
                                                  float* __restrict__ z, const int32_t* pred) {
  if (pred_off(pred)) return;
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < n; e += 256ll * gridDim.x)
    z[e] = __fmul_rn(x[e], y[e]);
}

// z = (x >= thr) ? v : 0
__global__ __launch_bounds__(256) void ge_kernel(int64_t n, const float* __restrict__ x, float thr,
                                                 float v, float* __restrict__ z,
                                                 const int32_t* pred) {
  if (pred_off(pred)) return;
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < n; e += 256ll * gridDim.x)
    z[e] = (x[e] >= thr) ? v : 0.0f;
}

// z = x * (s / *d) with a device float64 divisor (power iteration).
__global__ __launch_bounds__(256) void scale_div_kernel(int64_t n, const float* __restrict__ x,
                                                        const double* __restrict__ d,
                                                        float* __restrict__ z,
                                                        const int32_t* pred) {
  if (pred_off(pred)) return;
  const double dv = *d;
  const float inv = dv != 0.0 ? static_cast<float>(1.0 / dv) : 0.0f;
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < n; e += 256ll * gridDim.x)
    z[e] = dv != 0.0 ? static_cast<float>(static_cast<double>(x[e]) / dv) : x[e] * inv;
}

// Element term of each reduction kind.
__device__ __forceinline__ double red_term(int kind, const float* x, const float* y, int64_t e,
                                           int64_t n, int K) {
  switch (kind) {
    case RED_DOT:
      return static_cast<double>(x[e]) * static_cast<double>(y[e]);
    case RED_SUMSQ: {
      const double v = x[e];
      return v * v;
    }
    case RED_SQDIST: {
      const double v = static_cast<double>(x[e]) - static_cast<double>(y[e]);
      return v * v;
    }
    case RED_SUM:
      return x[e];
    case RED_L1:
      return fabs(static_cast<double>(x[e]));
    case RED_L21: {  // x is (K, n): sum over columns of the l2 norm over axis 0
      double s = 0.0;
      for (int k = 0; k < K; ++k) {
        const double v = x[static_cast<int64_t>(k) * n + e];
        s += v * v;
      }
      return sqrt(s);
    }
    default:
      return 0.0;
  }
}

__global__ __launch_bounds__(256) void reduce_pass1(int kind, int64_t n, int K,
                                                    const float* __restrict__ x,
                                                    const float* __restrict__ y,
                                                    double* __restrict__ partial,
                                                    const int32_t* pred) {
  if (pred_off(pred)) return;
  __shared__ double sh[4];
  double s = 0.0;
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < n; e += 256ll * gridDim.x)
    s += red_term(kind, x, y, e, n, K);
  s = block_sum_f64(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void reduce_pass2(int nparts, const double* __restrict__ partial,
                                                    double scale, int post,
                                                    double* __restrict__ out,
                                                    const int32_t* pred) {
  if (pred_off(pred)) return;
  __shared__ double sh[4];
  double s = 0.0;
  for (int e = threadIdx.x; e < nparts; e += 256) s += partial[e];
  s = block_sum_f64(s, sh);
  if (threadIdx.x == 0) {
    s *= scale;
    if (post == POST_SQRT) s = sqrt(s);
    *out = s;
  }
}

unsigned grid_for(int64_t n) {
  const int64_t b = (n + 1023) / 1024;
  return static_cast<unsigned>(b < 1 ? 1 : (b > 2048 ? 2048 : b));
}

int reduce(int kind, int64_t n, int K, const float* x, const float* y, double scale, int post,
           double* out, double* ws, const int32_t* pred, hipStream_t s) {
  const int64_t b = (n + 1023) / 1024;
  const unsigned g = static_cast<unsigned>(b < 1 ? 1 : (b > RED_BLOCKS ? RED_BLOCKS : b));
  hipLaunchKernelGGL(reduce_pass1, dim3(g), dim3(256), 0, s, kind, n, K, x, y, ws, pred);
  SCXR_LAUNCHED("reduce pass 1");
  hipLaunchKernelGGL(reduce_pass2, dim3(1), dim3(256), 0, s, static_cast<int>(g), ws, scale, post,
                     out, pred);
  SCXR_LAUNCHED("reduce pass 2");
  return SCXR_OK;
}

int sum_partials(int nparts, const double* parts, double scale, int post, double* out,
                 const int32_t* pred, hipStream_t s) {
  hipLaunchKernelGGL(reduce_pass2, dim3(1), dim3(256), 0, s, nparts, parts, scale, post, out,
                     pred);
  SCXR_LAUNCHED("sum partials");
  return SCXR_OK;
}

// Workspace of scxr_vec_reduce_noalloc: the first pass's float64 partials
// (one per workgroup, at most RED_BLOCKS), 8 KiB held by every device the
// library runs on until the process ends.  Host code never passes this
// name to a kernel: the name is a host-side stand-in, the device address
// comes from hipGetSymbolAddress, once per device.
constexpr int NOALLOC_PARTIALS = RED_BLOCKS;
static __device__ __align__(256) double scratch[NOALLOC_PARTIALS];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static std::atomic<double*> g_scratch_ptr[NOALLOC_MAX_DEVICES];

// Device address of the current device's copy of `scratch`.
static int scratch_of_current_device(double** out) {
  int dev = 0;
  SCXR_HIP(hipGetDevice(&dev));
  SCXR_REQUIRE(dev >= 0 && dev < NOALLOC_MAX_DEVICES, SCXR_ERR_UNSUPPORTED,
               "device ordinal %d is beyond the %d devices the static workspace supports", dev,
               NOALLOC_MAX_DEVICES);
  double* p = g_scratch_ptr[dev].load(std::memory_order_acquire);
  if (!p) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    p = g_scratch_ptr[dev].load(std::memory_order_relaxed);
    if (!p) {
      void* addr = nullptr;
      SCXR_HIP(hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch)));
      SCXR_REQUIRE(addr != nullptr, SCXR_ERR_HIP, "no device address for the static workspace");
      p = static_cast<double*>(addr);
      g_scratch_ptr[dev].store(p, std::memory_order_release);
    }
  }
  *out = p;
  return SCXR_OK;
}

}  // namespace vec
}  // namespace scxr

extern "C" {

int scxr_vec_axpby(int64_t n, float a, const float* x, float b, const float* y, float* z,
                   const int32_t* pred, void* stream) {
  SCXR_REQUIRE(n >= 0 && (n == 0 || (x && y && z)), SCXR_ERR_ARG, "bad axpby arguments");
  if (n == 0) return SCXR_OK;
  hipLaunchKernelGGL(vec::axpby_kernel, dim3(vec::grid_for(n)), dim3(256), 0, as_stream(stream), n,
                     a, x, b, y, z, pred);
  SCXR_LAUNCHED("axpby");
  return SCXR_OK;
}

int scxr_vec_add(int64_t n, const float* x, float sign, const float* y, float* z,
                 const int32_t* pred, void* stream) {
  SCXR_REQUIRE(n >= 0 && (n == 0 || (x && y && z)), SCXR_ERR_ARG, "bad add arguments");
  if (n == 0) return SCXR_OK;
  hipLaunchKernelGGL(vec::add_kernel, dim3(vec::grid_for(n)), dim3(256), 0, as_stream(stream), n,
                     x, sign, y, z, pred);
  SCXR_LAUNCHED("add");
  return SCXR_OK;
}

int scxr_vec_scale(int64_t n, float a, const float* x, float* z, const int32_t* pred,
                   void* stream) {
  SCXR_REQUIRE(n >= 0 && (n == 0 || (x && z)), SCXR_ERR_ARG, "bad scale arguments");
  if (n == 0) return SCXR_OK;
  hipLaunchKernelGGL(vec::scale_kernel, dim3(vec::grid_for(n)), dim3(256), 0, as_stream(stream), n,
                     a, x, z, pred);
  SCXR_LAUNCHED("scale");
  return SCXR_OK;
}

int scxr_vec_fill(int64_t n, float v, float* z, const int32_t* pred, void* stream) {
  SCXR_REQUIRE(n >= 0 && (n == 0 || z), SCXR_ERR_ARG, "bad fill arguments");
  if (n == 0) return SCXR_OK;
  hipLaunchKernelGGL(vec::fill_kernel, dim3(vec::grid_for(n)), dim3(256), 0, as_stream(stream), n,
                     v, z, pred);
  SCXR_LAUNCHED("fill");
  return SCXR_OK;
}

int scxr_vec_copy(int64_t n, const float* x, float* z, const int32_t* pred, void* stream) {
  SCXR_REQUIRE(n >= 0 && (n == 0 || (x && z)), SCXR_ERR_ARG, "bad copy arguments");
  if (n == 0) return SCXR_OK;
  hipLaunchKernelGGL(vec::copy_kernel, dim3(vec::grid_for(n)), dim3(256), 0, as_stream(stream), n,
                     x, z, pred);
  SCXR_LAUNCHED("copy");
  return SCXR_OK;
}

int scxr_vec_mul(int64_t n, const float* x, const float* y, float* z, const int32_t* pred,
                 void* stream) {
  SCXR_REQUIRE(n >= 0 && (n == 0 || (x && y && z)), SCXR_ERR_ARG, "bad mul arguments");
  if (n == 0) return SCXR_OK;
  hipLaunchKernelGGL(vec::mul_kernel, dim3(vec::grid_for(n)), dim3(256), 0, as_stream(stream), n,
                     x, y, z, pred);
  SCXR_LAUNCHED("mul");
  return SCXR_OK;
}

int scxr_vec_ge(int64_t n, const float* x, float thr, float v, float* z, const int32_t* pred,
                void* stream) {
  SCXR_REQUIRE(n >= 0 && (n == 0 || (x && z)), SCXR_ERR_ARG, "bad ge arguments");
  if (n == 0) return SCXR_OK;
  hipLaunchKernelGGL(vec::ge_kernel, dim3(vec::grid_for(n)), dim3(256), 0, as_stream(stream), n,
                     x, thr, v, z, pred);
  SCXR_LAUNCHED("ge");
  return SCXR_OK;
}

int scxr_vec_div_dev(int64_t n, const float* x, const double* d, float* z, const int32_t* pred,
                     void* stream) {
  SCXR_REQUIRE(n >= 0 && (n == 0 || (x && d && z)), SCXR_ERR_ARG, "bad div_dev arguments");
  if (n == 0) return SCXR_OK;
  hipLaunchKernelGGL(vec::scale_div_kernel, dim3(vec::grid_for(n)), dim3(256), 0,
                     as_stream(stream), n, x, d, z, pred);
  SCXR_LAUNCHED("div_dev");
  return SCXR_OK;
}

int scxr_vec_reduce(int32_t kind, int64_t n, int32_t K, const float* x, const float* y,
                    double scale, int32_t post, double* out, double* ws, const int32_t* pred,
                    void* stream) {
  SCXR_REQUIRE(kind >= vec::RED_DOT && kind <= vec::RED_L21, SCXR_ERR_ARG, "bad reduction kind");
  SCXR_REQUIRE(n >= 0 && x && out && ws, SCXR_ERR_ARG, "bad reduction arguments");
  SCXR_REQUIRE((kind != vec::RED_DOT && kind != vec::RED_SQDIST) || y, SCXR_ERR_ARG,
               "reduction needs a second operand");
  SCXR_REQUIRE(kind != vec::RED_L21 || K >= 1, SCXR_ERR_ARG, "L21 reduction needs K >= 1");
  return vec::reduce(kind, n, K, x, y, scale, post, out, ws, pred, as_stream(stream));
}

int scxr_vec_reduce_noalloc(int32_t kind, int64_t n, int32_t K, const float* x, const float* y,
                            double scale, int32_t post, double* out, const int32_t* pred,
                            void* stream) {
  SCXR_REQUIRE(kind >= vec::RED_DOT && kind <= vec::RED_L21, SCXR_ERR_ARG, "bad reduction kind");
  SCXR_REQUIRE(n >= 0 && x && out, SCXR_ERR_ARG, "bad reduction arguments");
  SCXR_REQUIRE((kind != vec::RED_DOT && kind != vec::RED_SQDIST) || y, SCXR_ERR_ARG,
               "reduction needs a second operand");
  SCXR_REQUIRE(kind != vec::RED_L21 || K >= 1, SCXR_ERR_ARG, "L21 reduction needs K >= 1");
  // The first pass's grid, hence the partial count, is that of vec::reduce:
  // one workgroup per 1024 elements, at least 1 and at most RED_BLOCKS.
  static_assert(vec::NOALLOC_PARTIALS >= vec::RED_BLOCKS, "workspace below the grid cap");
  const int64_t blocks = (n + 1023) / 1024;
  const int64_t parts = blocks < 1 ? 1 : (blocks > vec::RED_BLOCKS ? vec::RED_BLOCKS : blocks);
  SCXR_REQUIRE(parts <= vec::NOALLOC_PARTIALS, SCXR_ERR_ARG,
               "reduction needs %lld partials, the static workspace holds %d",
               static_cast<long long>(parts), vec::NOALLOC_PARTIALS);
  double* ws = nullptr;
  const int rc = vec::scratch_of_current_device(&ws);
  if (rc) return rc;
  // Every partial the second pass reads is written by this call's first
  // pass, so nothing left by an earlier call needs clearing.
  return vec::reduce(kind, n, K, x, y, scale, post, out, ws, pred, as_stream(stream));
}

}  // extern "C"
