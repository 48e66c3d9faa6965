// Internal interface of the vector kernels (shared with solver kernels).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace scxr {
namespace vec {

enum { RED_DOT = 0, RED_SUMSQ = 1, RED_SQDIST = 2, RED_SUM = 3, RED_L1 = 4, RED_L21 = 5 };
enum { POST_NONE = 0, POST_SQRT = 1 };

// Upper bound on first-pass workgroups (size of the float64 workspace).
constexpr int RED_BLOCKS = 1024;

unsigned grid_for(int64_t n);
// out = post(scale * sum(parts[0..nparts))), one workgroup, fixed order.
int sum_partials(int nparts, const double* parts, double scale, int post, double* out,
                 const int32_t* pred, hipStream_t s);
int reduce(int kind, int64_t n, int K, const float* x, const float* y, double scale, int post,
           double* out, double* ws, const int32_t* pred, hipStream_t s);

}  // namespace vec
}  // namespace scxr
