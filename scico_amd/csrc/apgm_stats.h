// Statistics row of the fused AcceleratedPGM-TV iteration (itstat fields
// Objective f + g and Residual, scico/optimize/_pgm.py:151-156, 242-248),
// reduced by one workgroup of T threads from per-workgroup partials.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace scxr {

// Sums the double partials obj[nobj], tv[ntv] and res[nres] and stores
//   hist[2 row] = obj_scale * sum(obj) + lam * sum(tv),
//   hist[2 row + 1] = sqrt(sum(res)),
// at row = counter[0] when row < hist_cap; counter[0] then counts the row.
// Every thread of the workgroup must call it.  Thread t adds the elements
// t + T i of each array in index order, in batches of 8 T whose loads of
// all three arrays are issued together; the waves then reduce by shuffles
// and thread 0 adds the waves in order.
template <int T>
__device__ __forceinline__ void apgm_stats_block(const double* __restrict__ obj, int nobj,
                                                 double obj_scale, const double* __restrict__ tv,
                                                 int ntv, double lam,
                                                 const double* __restrict__ res, int nres,
                                                 double* __restrict__ hist,
                                                 int32_t* __restrict__ counter, int hist_cap) {
  static_assert(T % 64 == 0, "whole waves");
  __shared__ double sh[3][T / 64];
  const double* src[3] = {obj, tv, res};
  const int n[3] = {nobj, ntv, nres};
  const int nmax = max(nobj, max(ntv, nres));
  double v[3] = {0.0, 0.0, 0.0};
  for (int e0 = 0; e0 < nmax; e0 += 8 * T) {
    double b[3][8];
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int e = e0 + static_cast<int>(threadIdx.x) + T * i;
        b[q][i] = e < n[q] ? src[q][e] : 0.0;
      }
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int i = 0; i < 8; ++i) v[q] += b[q][i];
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_down(v[q], off, 64);
    if (lane == 0) sh[q][wave] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double so = 0.0, st = 0.0, sr = 0.0;
    for (int w = 0; w < T / 64; ++w) {
      so += sh[0][w];
      st += sh[1][w];
      sr += sh[2][w];
    }
    const int row = counter[0];
    if (row < hist_cap) {
      hist[2 * row] = obj_scale * so + lam * st;
      hist[2 * row + 1] = sqrt(sr);
    }
    counter[0] = row + 1;
  }
}

}  // namespace scxr
