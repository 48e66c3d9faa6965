// Device stencils of the finite-difference operator
// (scico/linop/_diff.py:236-272: append=0 or circular) on row-major
// images / volumes, shared by the FD kernels, the fused ADMM kernels and
// the projector epilogues.
#pragma once

#include <hip/hip_runtime.h>

namespace scxr {
namespace fd {

// Forward difference along one axis at index k of an axis of length n,
// given the values at k and k+1 (or at 0 when circular and k = n-1).
//   append=0: x[k+1] - x[k] for k < n-1, else 0
//   circular: x[(k+1) % n] - x[k]
__device__ __forceinline__ float diff_fwd(float xk, float xnext, int k, int n, bool circ) {
  if (!circ && k == n - 1) return 0.0f;
  return __fsub_rn(xnext, xk);
}

// Transpose of the difference along one axis evaluated at index k from
// the difference values g[k-1] (gm) and g[k] (g0):
//   append=0: (D^T g)[0] = -g[0], (D^T g)[k] = g[k-1] - g[k], (D^T g)[n-1] = g[n-2]
//   circular: (D^T g)[k] = g[(k-1) % n] - g[k]
__device__ __forceinline__ float diff_adj(float gm, float g0, int k, int n, bool circ) {
  if (circ) return __fsub_rn(gm, g0);
  const float a = (k >= 1) ? gm : 0.0f;
  const float b = (k < n - 1) ? g0 : 0.0f;
  return __fsub_rn(a, b);
}

// (C^T C x)[i, j] for the 2D stack C = [D_0; D_1] (sum of the per-axis
// transposes in axis order, scico/linop/_stack.py:71-72).
__device__ __forceinline__ float gram2d(const float* __restrict__ x, int M, int N, int i, int j,
                                        bool circ) {
  const size_t c = static_cast<size_t>(i) * N + j;
  const float xc = x[c];
  // axis 0
  const int ip = (i + 1 < M) ? i + 1 : 0;
  const int im = (i >= 1) ? i - 1 : M - 1;
  const float g0 = diff_fwd(xc, x[static_cast<size_t>(ip) * N + j], i, M, circ);
  const float xm0 = x[static_cast<size_t>(im) * N + j];
  const float gm0 = diff_fwd(xm0, xc, im, M, circ);
  const float t0 = diff_adj(gm0, g0, i, M, circ);
  // axis 1
  const int jp = (j + 1 < N) ? j + 1 : 0;
  const int jm = (j >= 1) ? j - 1 : N - 1;
  const float g1 = diff_fwd(xc, x[static_cast<size_t>(i) * N + jp], j, N, circ);
  const float xm1 = x[static_cast<size_t>(i) * N + jm];
  const float gm1 = diff_fwd(xm1, xc, jm, N, circ);
  const float t1 = diff_adj(gm1, g1, j, N, circ);
  return __fadd_rn(t0, t1);
}

// (C^T w)[i, j] for w = a - b with a, b stacked (2, M, N) (b may be null).
__device__ __forceinline__ float adj2d_diff(const float* __restrict__ a,
                                            const float* __restrict__ b, int M, int N, int i,
                                            int j, bool circ) {
  const size_t MN = static_cast<size_t>(M) * N;
  auto w = [&](int comp, int r, int c) {
    const size_t e = comp * MN + static_cast<size_t>(r) * N + c;
    return b ? __fsub_rn(a[e], b[e]) : a[e];
  };
  const int im = (i >= 1) ? i - 1 : M - 1;
  const int jm = (j >= 1) ? j - 1 : N - 1;
  const float t0 = diff_adj(w(0, im, j), w(0, i, j), i, M, circ);
  const float t1 = diff_adj(w(1, i, jm), w(1, i, j), j, N, circ);
  return __fadd_rn(t0, t1);
}

}  // namespace fd
}  // namespace scxr
