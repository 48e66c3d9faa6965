// Solver kernels: finite differences, L1 / L21 proximal maps, conjugate
// gradient with device-resident scalars, and the fused ADMM z/u update
// with its iteration statistics.
//
// CG (scico/solver.py:326-405) keeps every scalar on the device in a
// float64 state block, so a fixed-length sequence of iterations can be
// enqueued (or captured in a graph) without host synchronisation:
//   state[ST_BN]      ||b||
//   state[ST_TOL2]    max(tol * ||b||, atol)^2
//   state[ST_NITER]   iterations performed
//   state[ST_RELRES]  sqrt(num) / ||b||
//   state[ST_NUM + k] r^T r after k iterations
//   state[ST_NUM + (maxiter + 1) + k] p^T A p of iteration k
//   flags[k]          1 while iteration k is to run ("ii < maxiter and
//                     num > tol^2"), else 0; every kernel of iteration k is
//                     predicated on flags[k].
#include <cmath>

#include "apgm_stats.h"
#include "common.h"
#include "fd_impl.h"
#include "solver_impl.h"
#include "vec_impl.h"

using namespace scxr;

namespace scxr {
namespace solver {

// ------------------------------------------------------ finite differences --

// out (ndim, d0, d1, d2) = stack of per-axis differences of x (d0, d1, d2);
// component c differences axis (3 - ndim) + c.
__global__ __launch_bounds__(256) void fd_fwd_kernel(int ndim, int d0, int d1, int d2,
                                                     const float* __restrict__ x,
                                                     float* __restrict__ out, int circ,
                                                     const int32_t* pred) {
  if (pred_off(pred)) return;
  const int64_t n = static_cast<int64_t>(d0) * d1 * d2;
  const int dims[3] = {d0, d1, d2};
  const int64_t strides[3] = {static_cast<int64_t>(d1) * d2, d2, 1};
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < n; e += 256ll * gridDim.x) {
    const int k0 = static_cast<int>(e / strides[0]);
    const int k1 = static_cast<int>((e / d2) % d1);
    const int k2 = static_cast<int>(e % d2);
    const int idx[3] = {k0, k1, k2};
    const float xc = x[e];
    for (int c = 0; c < ndim; ++c) {
      const int ax = 3 - ndim + c;
      const int k = idx[ax], len = dims[ax];
      const int kn = (k + 1 < len) ? k + 1 : 0;
      const float xn = x[e + static_cast<int64_t>(kn - k) * strides[ax]];
      out[c * n + e] = fd::diff_fwd(xc, xn, k, len, circ != 0);
    }
  }
}

// out = sum_c D_c^T g[c]   (scico/linop/_stack.py:71-72)
__global__ __launch_bounds__(256) void fd_adj_kernel(int ndim, int d0, int d1, int d2,
                                                     const float* __restrict__ g,
                                                     float* __restrict__ out, int circ,
                                                     float scale, const int32_t* pred) {
  if (pred_off(pred)) return;
  const int64_t n = static_cast<int64_t>(d0) * d1 * d2;
  const int dims[3] = {d0, d1, d2};
  const int64_t strides[3] = {static_cast<int64_t>(d1) * d2, d2, 1};
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < n; e += 256ll * gridDim.x) {
    const int idx[3] = {static_cast<int>(e / strides[0]), static_cast<int>((e / d2) % d1),
                        static_cast<int>(e % d2)};
    float acc = 0.0f;
    for (int c = 0; c < ndim; ++c) {
      const int ax = 3 - ndim + c;
      const int k = idx[ax], len = dims[ax];
      const int km = (k >= 1) ? k - 1 : len - 1;
      const float* gc = g + c * n;
      const float t = fd::diff_adj(gc[e + static_cast<int64_t>(km - k) * strides[ax]], gc[e], k,
                                   len, circ != 0);
      acc = (c == 0) ? t : __fadd_rn(acc, t);
    }
    out[e] = scale == 1.0f ? acc : __fmul_rn(scale, acc);
  }
}

// out = rho * C^T C x  (gram of the stacked differences)
__global__ __launch_bounds__(256) void fd_gram_kernel(int ndim, int d0, int d1, int d2,
                                                      const float* __restrict__ x,
                                                      float* __restrict__ out, int circ,
                                                      float rho, const int32_t* pred) {
  if (pred_off(pred)) return;
  const int64_t n = static_cast<int64_t>(d0) * d1 * d2;
  const int dims[3] = {d0, d1, d2};
  const int64_t strides[3] = {static_cast<int64_t>(d1) * d2, d2, 1};
  const bool cc = circ != 0;
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < n; e += 256ll * gridDim.x) {
    const int idx[3] = {static_cast<int>(e / strides[0]), static_cast<int>((e / d2) % d1),
                        static_cast<int>(e % d2)};
    const float xc = x[e];
    float acc = 0.0f;
    for (int c = 0; c < ndim; ++c) {
      const int ax = 3 - ndim + c;
      const int k = idx[ax], len = dims[ax];
      const int kp = (k + 1 < len) ? k + 1 : 0;
      const int km = (k >= 1) ? k - 1 : len - 1;
      const int64_t st = strides[ax];
      const float g0 = fd::diff_fwd(xc, x[e + (kp - k) * st], k, len, cc);
      const float gm = fd::diff_fwd(x[e + (km - k) * st], xc, km, len, cc);
      const float t = fd::diff_adj(gm, g0, k, len, cc);
      acc = (c == 0) ? t : __fadd_rn(acc, t);
    }
    out[e] = __fmul_rn(rho, acc);
  }
}

// -------------------------------------------------------------- proxes --

// L21 over axis 0 of v (K, n)   (scico/functional/_norm.py:254-263)
__device__ __forceinline__ void l21_shrink(float* v, int K, float lam) {
  float ss = 0.0f;
  for (int k = 0; k < K; ++k) ss = (k == 0) ? __fmul_rn(v[0], v[0]) : __fadd_rn(ss, __fmul_rn(v[k], v[k]));
  const float len = sqrtf(ss);
  float nl = __fsub_rn(len, lam);
  nl = __fmul_rn(0.5f, __fadd_rn(nl, fabsf(nl)));
  for (int k = 0; k < K; ++k) {
    const float dir = (len != 0.0f) ? v[k] / len : 0.0f;  // no_nan_divide
    v[k] = __fmul_rn(nl, dir);
  }
}

// L1   (scico/functional/_norm.py:110-116)
__device__ __forceinline__ float l1_shrink(float v, float lam) {
  float t = __fsub_rn(fabsf(v), lam);
  t = __fmul_rn(0.5f, __fadd_rn(t, fabsf(t)));
  const float sg = (v > 0.0f) ? 1.0f : ((v < 0.0f) ? -1.0f : 0.0f);
  return __fmul_rn(sg, t);
}

__global__ __launch_bounds__(256) void l21_prox_kernel(int K, int64_t n, const float* __restrict__ v,
                                                       float lam, float* __restrict__ out,
                                                       const int32_t* pred) {
  if (pred_off(pred)) return;
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < n; e += 256ll * gridDim.x) {
    float w[MAX_K];
    for (int k = 0; k < K; ++k) w[k] = v[k * n + e];
    l21_shrink(w, K, lam);
    for (int k = 0; k < K; ++k) out[k * n + e] = w[k];
  }
}

__global__ __launch_bounds__(256) void l1_prox_kernel(int64_t n, const float* __restrict__ v,
                                                      float lam, float* __restrict__ out,
                                                      const int32_t* pred) {
  if (pred_off(pred)) return;
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < n; e += 256ll * gridDim.x)
    out[e] = l1_shrink(v[e], lam);
}

// ------------------------------------------------------------------ CG --

// r = b - A x0; p = r; partials of b^T b and r^T r.
// Kernel bodies take their (virtual) block index vb and grid size vg, so
// that the batched kernels run exactly the same per-solve arithmetic.
__device__ __forceinline__ void cg_start_body(int64_t n, const float* __restrict__ b,
                                              const float* __restrict__ Ax,
                                              float* __restrict__ r, float* __restrict__ p,
                                              double* __restrict__ ws, int vb, int vg) {
  __shared__ double sh[4];
  double bb = 0.0, rr = 0.0;
  for (int64_t e = vb * 256ll + threadIdx.x; e < n; e += 256ll * vg) {
    const float be = b[e];
    const float re = __fsub_rn(be, Ax[e]);
    r[e] = re;
    p[e] = re;
    bb += static_cast<double>(be) * be;
    rr += static_cast<double>(re) * re;
  }
  bb = block_sum_f64(bb, sh);
  if (threadIdx.x == 0) ws[vb] = bb;
  rr = block_sum_f64(rr, sh);
  if (threadIdx.x == 0) ws[vg + vb] = rr;
}

__global__ __launch_bounds__(256) void cg_start_kernel(int64_t n, const float* __restrict__ b,
                                                       const float* __restrict__ Ax,
                                                       float* __restrict__ r,
                                                       float* __restrict__ p,
                                                       double* __restrict__ ws) {
  cg_start_body(n, b, Ax, r, p, ws, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ void cg_start_fin_body(int nparts, const double* __restrict__ ws,
                                                  double* __restrict__ state,
                                                  int32_t* __restrict__ flags, int maxiter,
                                                  double tol, double atol) {
  __shared__ double sh[4];
  const double bb = all_sum_f64(ws, nparts, sh);
  const double rr = all_sum_f64(ws + nparts, nparts, sh);
  for (int k = 1 + threadIdx.x; k <= maxiter; k += 256) flags[k] = 0;
  if (threadIdx.x == 0) {
    const double bn = sqrt(bb);
    const double t = fmax(tol * bn, atol);
    state[ST_BN] = bn;
    state[ST_TOL2] = t * t;
    state[ST_NITER] = 0.0;
    state[ST_RELRES] = sqrt(rr) / bn;
    state[ST_NUM] = rr;
    flags[0] = (maxiter > 0 && rr > t * t) ? 1 : 0;
  }
}

__global__ __launch_bounds__(256) void cg_start_fin(int nparts, const double* __restrict__ ws,
                                                    double* __restrict__ state,
                                                    int32_t* __restrict__ flags, int maxiter,
                                                    double tol, double atol) {
  cg_start_fin_body(nparts, ws, state, flags, maxiter, tol, atol);
}

// alpha = num_k / p^T A p; x += alpha p; r -= alpha A p; partials of r^T r.
__device__ __forceinline__ void cg_xr_body(int64_t n, int k, int maxiter, float* __restrict__ x,
                                           float* __restrict__ r, const float* __restrict__ p,
                                           const float* __restrict__ Ap,
                                           double* __restrict__ state,
                                           const int32_t* __restrict__ flags,
                                           const double* __restrict__ pap_parts, int pap_n,
                                           double* __restrict__ ws_rr, int vb, int vg) {
  if (flags[k] == 0) return;
  __shared__ double sh[4];
  const double pap = all_sum_f64(pap_parts, pap_n, sh);
  const float alpha = static_cast<float>(static_cast<float>(state[ST_NUM + k]) /
                                         static_cast<float>(pap));
  double rr = 0.0;
  for (int64_t e = vb * 256ll + threadIdx.x; e < n; e += 256ll * vg) {
    const float pe = p[e];
    x[e] = __fadd_rn(x[e], __fmul_rn(alpha, pe));
    const float re = __fsub_rn(r[e], __fmul_rn(alpha, Ap[e]));
    r[e] = re;
    rr += static_cast<double>(re) * re;
  }
  rr = block_sum_f64(rr, sh);
  if (threadIdx.x == 0) {
    ws_rr[vb] = rr;
    if (vb == 0) state[ST_NUM + (maxiter + 1) + k] = pap;
  }
}

__global__ __launch_bounds__(256) void cg_xr_kernel(int64_t n, int k, int maxiter,
                                                    float* __restrict__ x, float* __restrict__ r,
                                                    const float* __restrict__ p,
                                                    const float* __restrict__ Ap,
                                                    double* __restrict__ state,
                                                    const int32_t* __restrict__ flags,
                                                    const double* __restrict__ pap_parts,
                                                    int pap_n, double* __restrict__ ws_rr) {
  cg_xr_body(n, k, maxiter, x, r, p, Ap, state, flags, pap_parts, pap_n, ws_rr, blockIdx.x,
             gridDim.x);
}

// beta = num_{k+1} / num_k; p = r + beta p; stopping test for k + 1.
__device__ __forceinline__ void cg_p_body(int64_t n, int k, int maxiter,
                                          const float* __restrict__ r, float* __restrict__ p,
                                          double* __restrict__ state, int32_t* __restrict__ flags,
                                          const double* __restrict__ ws_rr, int rr_n, int vb,
                                          int vg) {
  if (flags[k] == 0) return;
  __shared__ double sh[4];
  const double rr = all_sum_f64(ws_rr, rr_n, sh);
  const float num_new = static_cast<float>(rr);
  const float beta = num_new / static_cast<float>(state[ST_NUM + k]);
  for (int64_t e = vb * 256ll + threadIdx.x; e < n; e += 256ll * vg)
    p[e] = __fadd_rn(r[e], __fmul_rn(beta, p[e]));
  if (vb == 0 && threadIdx.x == 0) {
    state[ST_NUM + k + 1] = rr;
    state[ST_NITER] = k + 1;
    state[ST_RELRES] = sqrt(rr) / state[ST_BN];
    flags[k + 1] = (k + 1 < maxiter && rr > state[ST_TOL2]) ? 1 : 0;
  }
}

__global__ __launch_bounds__(256) void cg_p_kernel(int64_t n, int k, int maxiter,
                                                   const float* __restrict__ r,
                                                   float* __restrict__ p,
                                                   double* __restrict__ state,
                                                   int32_t* __restrict__ flags,
                                                   const double* __restrict__ ws_rr, int rr_n) {
  cg_p_body(n, k, maxiter, r, p, state, flags, ws_rr, rr_n, blockIdx.x, gridDim.x);
}

// cg_p_body over a 32 x 32 tile of an M x N image (grid: ceil(N/32) x
// ceil(M/32), 256 threads), also writing the transposed direction pT
// (N x M) that the 2D projector's column-fast views stream: p is updated
// with the same operations, so it is bitwise cg_p_kernel's.
__global__ __launch_bounds__(256) void cg_p_t_kernel(int M, int N, int k, int maxiter,
                                                     const float* __restrict__ r,
                                                     float* __restrict__ p,
                                                     float* __restrict__ pT,
                                                     double* __restrict__ state,
                                                     int32_t* __restrict__ flags,
                                                     const double* __restrict__ ws_rr, int rr_n) {
  if (flags[k] == 0) return;
  __shared__ double sh[4];
  __shared__ float tile[32][33];
  const double rr = all_sum_f64(ws_rr, rr_n, sh);
  const float num_new = static_cast<float>(rr);
  const float beta = num_new / static_cast<float>(state[ST_NUM + k]);
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  const int i0 = blockIdx.y * 32, j0 = blockIdx.x * 32;
  for (int q = ty; q < 32; q += 8) {
    const int i = i0 + q, j = j0 + tx;
    if (i < M && j < N) {
      const int64_t e = static_cast<int64_t>(i) * N + j;
      const float v = __fadd_rn(r[e], __fmul_rn(beta, p[e]));
      p[e] = v;
      tile[q][tx] = v;
    }
  }
  __syncthreads();
  for (int q = ty; q < 32; q += 8) {
    const int j = j0 + q, i = i0 + tx;
    if (i < M && j < N) pT[static_cast<int64_t>(j) * M + i] = tile[tx][q];
  }
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
    state[ST_NUM + k + 1] = rr;
    state[ST_NITER] = k + 1;
    state[ST_RELRES] = sqrt(rr) / state[ST_BN];
    flags[k + 1] = (k + 1 < maxiter && rr > state[ST_TOL2]) ? 1 : 0;
  }
}

// ---------------------------------------------------------------- ADMM --

// One fused pass of the ADMM z/u update for C = FiniteDifference (2D,
// append=0 or circular) and g = lam * (L21 over the stack axis | L1):
//   Cx = C x; z_old = z; z = prox_g(Cx + u, lam / rho); u = (u + Cx) - z
// (scico/optimize/_admm.py:364-378), plus per-workgroup partials of
// ||C x - z||^2 (primal residual) and g(z) / lam.
__device__ __forceinline__ void admm_zu2d_body(int M, int N, const float* __restrict__ x,
                                               float* __restrict__ z, float* __restrict__ u,
                                               float* __restrict__ zold, float thr, int circ,
                                               int iso, double* __restrict__ ws_prml,
                                               double* __restrict__ ws_g, int vb, int vg) {
  __shared__ double sh[4];
  const int64_t n = static_cast<int64_t>(M) * N;
  const bool cc = circ != 0;
  double prml = 0.0, gsum = 0.0;
  for (int64_t e = vb * 256ll + threadIdx.x; e < n; e += 256ll * vg) {
    const int i = static_cast<int>(e / N), j = static_cast<int>(e - static_cast<int64_t>(i) * N);
    const float xc = x[e];
    const int ip = (i + 1 < M) ? i + 1 : 0, jp = (j + 1 < N) ? j + 1 : 0;
    float cx[2];
    cx[0] = fd::diff_fwd(xc, x[static_cast<int64_t>(ip) * N + j], i, M, cc);
    cx[1] = fd::diff_fwd(xc, x[static_cast<int64_t>(i) * N + jp], j, N, cc);
    float v[2], uo[2];
    for (int c = 0; c < 2; ++c) {
      const float zc = z[c * n + e];
      zold[c * n + e] = zc;
      uo[c] = u[c * n + e];
      v[c] = __fadd_rn(cx[c], uo[c]);
    }
    if (iso) {
      l21_shrink(v, 2, thr);
      gsum += sqrt(static_cast<double>(v[0]) * v[0] + static_cast<double>(v[1]) * v[1]);
    } else {
      v[0] = l1_shrink(v[0], thr);
      v[1] = l1_shrink(v[1], thr);
      gsum += fabs(static_cast<double>(v[0])) + fabs(static_cast<double>(v[1]));
    }
    for (int c = 0; c < 2; ++c) {
      z[c * n + e] = v[c];
      u[c * n + e] = __fsub_rn(__fadd_rn(uo[c], cx[c]), v[c]);
      const double d = static_cast<double>(__fsub_rn(cx[c], v[c]));
      prml += d * d;
    }
  }
  prml = block_sum_f64(prml, sh);
  if (threadIdx.x == 0) ws_prml[vb] = prml;
  gsum = block_sum_f64(gsum, sh);
  if (threadIdx.x == 0) ws_g[vb] = gsum;
}

__global__ __launch_bounds__(256) void admm_zu2d_kernel(int M, int N,
                                                        const float* __restrict__ x,
                                                        float* __restrict__ z,
                                                        float* __restrict__ u,
                                                        float* __restrict__ zold, float thr,
                                                        int circ, int iso,
                                                        double* __restrict__ ws_prml,
                                                        double* __restrict__ ws_g) {
  admm_zu2d_body(M, N, x, z, u, zold, thr, circ, iso, ws_prml, ws_g, blockIdx.x, gridDim.x);
}

// Partials of ||rho * C^T (z - z_old)||^2 (dual residual, _admm.py:279-295).
__device__ __forceinline__ void admm_dual2d_body(int M, int N, const float* __restrict__ z,
                                                 const float* __restrict__ zold, float rho,
                                                 int circ, double* __restrict__ ws, int vb,
                                                 int vg) {
  __shared__ double sh[4];
  const int64_t n = static_cast<int64_t>(M) * N;
  double s = 0.0;
  for (int64_t e = vb * 256ll + threadIdx.x; e < n; e += 256ll * vg) {
    const int i = static_cast<int>(e / N), j = static_cast<int>(e - static_cast<int64_t>(i) * N);
    const float t = __fmul_rn(rho, fd::adj2d_diff(z, zold, M, N, i, j, circ != 0));
    s += static_cast<double>(t) * t;
  }
  s = block_sum_f64(s, sh);
  if (threadIdx.x == 0) ws[vb] = s;
}

// ------------------------------------------------- nd ADMM-TV (any A) --
// Fused passes of the ADMM-TV iteration for an operator A without
// projector epilogues (3D transform, ...), over all axes of an
// ndim-D array (ndim <= 3, extents padded to 3 with leading 1s).

struct NdIdx {
  int dims[3];
  int64_t st[3];
  int64_t n;
  __device__ NdIdx(int d0, int d1, int d2) {
    dims[0] = d0;
    dims[1] = d1;
    dims[2] = d2;
    st[0] = static_cast<int64_t>(d1) * d2;
    st[1] = d2;
    st[2] = 1;
    n = static_cast<int64_t>(d0) * d1 * d2;
  }
  __device__ void split(int64_t e, int idx[3]) const {
    if (n <= 0xffffffffll) {  // 32-bit arithmetic below 2^32 elements
      const unsigned u = static_cast<unsigned>(e), d2 = static_cast<unsigned>(dims[2]);
      const unsigned q = u / d2;
      idx[2] = static_cast<int>(u - q * d2);
      idx[1] = static_cast<int>(q % static_cast<unsigned>(dims[1]));
      idx[0] = static_cast<int>(q / static_cast<unsigned>(dims[1]));
      return;
    }
    idx[0] = static_cast<int>(e / st[0]);
    idx[1] = static_cast<int>((e / dims[2]) % dims[1]);
    idx[2] = static_cast<int>(e % dims[2]);
  }
};

// Coordinates along a grid-stride walk e, e + S, e + 2 S, ...: split once,
// then advanced by S with one carry per axis (S split as sx, sx[1] < d1,
// sx[2] < d2), so the loop divides nothing per element.
struct NdWalk {
  int idx[3] = {0, 0, 0}, sx[3] = {0, 0, 0};
  __device__ NdWalk(const NdIdx& t, int64_t e, int64_t S) {
    if (e < t.n) t.split(e, idx);
    if (S < t.n) t.split(S, sx);
  }
  __device__ __forceinline__ void next(const NdIdx& t) {
    idx[2] += sx[2];
    if (idx[2] >= t.dims[2]) {
      idx[2] -= t.dims[2];
      ++idx[1];
    }
    idx[1] += sx[1];
    if (idx[1] >= t.dims[1]) {
      idx[1] -= t.dims[1];
      ++idx[0];
    }
    idx[0] += sx[0];
  }
};

// (C^T w)[e] for w = a - b stacked (ndim, n) (b may be null), axis order sum.
__device__ __forceinline__ float adj_nd(const float* __restrict__ a, const float* __restrict__ b,
                                        const NdIdx& t, int ndim, int64_t e, const int idx[3],
                                        bool cc) {
  float acc = 0.0f;
  for (int c = 0; c < ndim; ++c) {
    const int ax = 3 - ndim + c;
    const int k = idx[ax], len = t.dims[ax];
    const int km = (k >= 1) ? k - 1 : len - 1;
    const int64_t em = e + static_cast<int64_t>(km - k) * t.st[ax];
    const int64_t o = c * t.n;
    const float wm = b ? __fsub_rn(a[o + em], b[o + em]) : a[o + em];
    const float w0 = b ? __fsub_rn(a[o + e], b[o + e]) : a[o + e];
    const float v = fd::diff_adj(wm, w0, k, len, cc);
    acc = (c == 0) ? v : __fadd_rn(acc, v);
  }
  return acc;
}

// out = aty + rho C^T (z - u) for an M x N image: the operations of the 2D
// adjoint's right-hand-side epilogue (_admmaux.py:245-256), from a kept
// A^T y (workgroup vb of vg).
__device__ __forceinline__ void admm_rhs2d_body(int M, int N, const float* __restrict__ aty,
                                                const float* __restrict__ z,
                                                const float* __restrict__ u, float rho, int circ,
                                                float* __restrict__ out, int vb, int vg) {
  const int64_t n = static_cast<int64_t>(M) * N;
  for (int64_t e = vb * 256ll + threadIdx.x; e < n; e += 256ll * vg) {
    const int i = static_cast<int>(e / N), j = static_cast<int>(e - static_cast<int64_t>(i) * N);
    out[e] = __fadd_rn(aty[e], __fmul_rn(rho, fd::adj2d_diff(z, u, M, N, i, j, circ != 0)));
  }
}

__global__ __launch_bounds__(256) void admm_rhs2d_kernel(int M, int N,
                                                         const float* __restrict__ aty,
                                                         const float* __restrict__ z,
                                                         const float* __restrict__ u, float rho,
                                                         int circ, float* __restrict__ out) {
  admm_rhs2d_body(M, N, aty, z, u, rho, circ, out, blockIdx.x, gridDim.x);
}

// out = base + rho C^T (z - u)   (_admmaux.py:245-256); out may alias base.
__global__ __launch_bounds__(256) void admm_rhs_nd_kernel(int ndim, int d0, int d1, int d2,
                                                          const float* base,
                                                          const float* __restrict__ z,
                                                          const float* __restrict__ u, float rho,
                                                          int circ, float* out) {
  const NdIdx t(d0, d1, d2);
  const int64_t S = 256ll * gridDim.x;
  NdWalk wk(t, blockIdx.x * 256ll + threadIdx.x, S);
#pragma unroll 4
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < t.n; e += S, wk.next(t)) {
    const int* idx = wk.idx;
    const float c = __fmul_rn(rho, adj_nd(z, u, t, ndim, e, idx, circ != 0));
    out[e] = __fadd_rn(base[e], c);
  }
}

// out = rho C^T C x + ata (out may alias ata), with per-block partials of
// sum x * out when dot_ws is given (_admmaux.py:258-268, solver.py:392).
__global__ __launch_bounds__(256) void admm_lhs_nd_kernel(int ndim, int d0, int d1, int d2,
                                                          const float* __restrict__ x,
                                                          const float* ata, float rho, int circ,
                                                          float* out, double* __restrict__ dot_ws,
                                                          const int32_t* pred) {
  if (pred_off(pred)) return;
  __shared__ double sh[4];
  const NdIdx t(d0, d1, d2);
  const bool cc = circ != 0;
  // Each thread takes elements e, e + S, e + 2 S, ... (S the grid's thread
  // count) in that order, NW at a time: the loads of a batch are issued
  // before any of its sums (the pass is bound by memory latency, not by
  // arithmetic), and the coordinates advance by S (NdWalk).  Results and
  // the order of the dot partial are those of one element per step.
  constexpr int NW = 4;
  const int64_t S = 256ll * gridDim.x;
  int64_t e = blockIdx.x * 256ll + threadIdx.x;
  NdWalk wk(t, e, S);
  double dot = 0.0;
  for (; e < t.n; e += NW * S) {
    float xc[NW], at[NW], xp[NW][3], xm[NW][3];
    int id[NW][3];
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const int64_t ew = e + w * S;
#pragma unroll
      for (int a = 0; a < 3; ++a) id[w][a] = wk.idx[a];
      wk.next(t);
      if (ew < t.n) {
        xc[w] = x[ew];
        at[w] = ata[ew];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          if (c >= ndim) break;
          const int ax = 3 - ndim + c;
          const int k = id[w][ax], len = t.dims[ax];
          const int kp = (k + 1 < len) ? k + 1 : 0;
          const int km = (k >= 1) ? k - 1 : len - 1;
          const int64_t st = t.st[ax];
          xp[w][c] = x[ew + (kp - k) * st];
          xm[w][c] = x[ew + (km - k) * st];
        }
      }
    }
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const int64_t ew = e + w * S;
      if (ew >= t.n) break;
      float acc = 0.0f;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (c >= ndim) break;
        const int ax = 3 - ndim + c;
        const int k = id[w][ax], len = t.dims[ax];
        const int km = (k >= 1) ? k - 1 : len - 1;
        const float g0 = fd::diff_fwd(xc[w], xp[w][c], k, len, cc);
        const float gm = fd::diff_fwd(xm[w][c], xc[w], km, len, cc);
        const float v = fd::diff_adj(gm, g0, k, len, cc);
        acc = (c == 0) ? v : __fadd_rn(acc, v);
      }
      const float o = __fadd_rn(__fmul_rn(rho, acc), at[w]);
      out[ew] = o;
      dot += static_cast<double>(xc[w]) * o;
    }
  }
  if (dot_ws) {
    dot = block_sum_f64(dot, sh);
    if (threadIdx.x == 0) dot_ws[blockIdx.x] = dot;
  }
}

// z/u update of _admm.py:364-378 with C = FiniteDifference over all axes:
//   Cx = C x, v = Cx + u, z = prox(v, thr), u = u + Cx - z, zold = z_prev,
// plus per-block partials of ||Cx - z||^2 and of g(z) (L21 over the
// stack axis, or L1).
__global__ __launch_bounds__(256) void admm_zu_nd_kernel(int ndim, int d0, int d1, int d2,
                                                         const float* __restrict__ x,
                                                         float* __restrict__ z,
                                                         float* __restrict__ u,
                                                         float* __restrict__ zold, float thr,
                                                         int circ, int iso,
                                                         double* __restrict__ ws_prml,
                                                         double* __restrict__ ws_g) {
  __shared__ double sh[4];
  const NdIdx t(d0, d1, d2);
  const bool cc = circ != 0;
  double prml = 0.0, gsum = 0.0;
  const int64_t S = 256ll * gridDim.x;
  NdWalk wk(t, blockIdx.x * 256ll + threadIdx.x, S);
#pragma unroll 4
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < t.n; e += S, wk.next(t)) {
    const int* idx = wk.idx;
    const float xc = x[e];
    float cx[3], v[3], uo[3];
    for (int c = 0; c < ndim; ++c) {
      const int ax = 3 - ndim + c;
      const int k = idx[ax], len = t.dims[ax];
      const int kn = (k + 1 < len) ? k + 1 : 0;
      cx[c] = fd::diff_fwd(xc, x[e + static_cast<int64_t>(kn - k) * t.st[ax]], k, len, cc);
      const int64_t o = c * t.n + e;
      zold[o] = z[o];
      uo[c] = u[o];
      v[c] = __fadd_rn(cx[c], uo[c]);
    }
    if (iso) {
      l21_shrink(v, ndim, thr);
      double ss = 0.0;
      for (int c = 0; c < ndim; ++c) ss += static_cast<double>(v[c]) * v[c];
      gsum += sqrt(ss);
    } else {
      for (int c = 0; c < ndim; ++c) {
        v[c] = l1_shrink(v[c], thr);
        gsum += fabs(static_cast<double>(v[c]));
      }
    }
    for (int c = 0; c < ndim; ++c) {
      const int64_t o = c * t.n + e;
      z[o] = v[c];
      u[o] = __fsub_rn(__fadd_rn(uo[c], cx[c]), v[c]);
      const double d = static_cast<double>(__fsub_rn(cx[c], v[c]));
      prml += d * d;
    }
  }
  prml = block_sum_f64(prml, sh);
  if (threadIdx.x == 0) ws_prml[blockIdx.x] = prml;
  gsum = block_sum_f64(gsum, sh);
  if (threadIdx.x == 0) ws_g[blockIdx.x] = gsum;
}

// Partials of ||rho C^T (z - zold)||^2 (dual residual, _admm.py:279-295).
__global__ __launch_bounds__(256) void admm_dual_nd_kernel(int ndim, int d0, int d1, int d2,
                                                           const float* __restrict__ z,
                                                           const float* __restrict__ zold,
                                                           float rho, int circ,
                                                           double* __restrict__ ws) {
  __shared__ double sh[4];
  const NdIdx t(d0, d1, d2);
  double s = 0.0;
  const int64_t S = 256ll * gridDim.x;
  NdWalk wk(t, blockIdx.x * 256ll + threadIdx.x, S);
#pragma unroll 4
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < t.n; e += S, wk.next(t)) {
    const int* idx = wk.idx;
    const float v = __fmul_rn(rho, adj_nd(z, zold, t, ndim, e, idx, circ != 0));
    s += static_cast<double>(v) * v;
  }
  s = block_sum_f64(s, sh);
  if (threadIdx.x == 0) ws[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void admm_dual2d_kernel(int M, int N,
                                                          const float* __restrict__ z,
                                                          const float* __restrict__ zold,
                                                          float rho, int circ,
                                                          double* __restrict__ ws) {
  admm_dual2d_body(M, N, z, zold, rho, circ, ws, blockIdx.x, gridDim.x);
}

// Iteration statistics row (Objective, Prml Rsdl, Dual Rsdl, CG It, CG Res)
// appended at hist[5 * counter]; counter is a device int.
__device__ __forceinline__ void admm_itstat_body(
    const double* __restrict__ res, int nres, const double* __restrict__ g, int ng,
    const double* __restrict__ prml, int nprml, const double* __restrict__ dual, int ndual,
    const double* __restrict__ cg_state, double lam, double rho, double* __restrict__ hist,
    int32_t* __restrict__ counter, int hist_cap) {
  __shared__ double sh[4];
  const double r = all_sum_f64(res, nres, sh);
  const double gs = all_sum_f64(g, ng, sh);
  const double p = all_sum_f64(prml, nprml, sh);
  const double d = all_sum_f64(dual, ndual, sh);
  if (threadIdx.x == 0) {
    const int it = *counter;
    if (it < hist_cap) {
      double* row = hist + 5 * static_cast<int64_t>(it);
      row[0] = 0.5 * r + lam * gs;
      row[1] = sqrt(rho * p);
      row[2] = sqrt(d);
      row[3] = cg_state ? cg_state[ST_NITER] : 0.0;
      row[4] = cg_state ? cg_state[ST_RELRES] : 0.0;
    }
    *counter = it + 1;
  }
}

__global__ __launch_bounds__(256) void admm_itstat_kernel(
    const double* __restrict__ res, int nres, const double* __restrict__ g, int ng,
    const double* __restrict__ prml, int nprml, const double* __restrict__ dual, int ndual,
    const double* __restrict__ cg_state, double lam, double rho, double* __restrict__ hist,
    int32_t* __restrict__ counter, int hist_cap) {
  admm_itstat_body(res, nres, g, ng, prml, nprml, dual, ndual, cg_state, lam, rho, hist,
                   counter, hist_cap);
}

// One AcceleratedPGM statistics row (apgm_stats_block), reduced by one
// workgroup of 1024 threads.
constexpr int ITS_T = 1024;
__global__ __launch_bounds__(ITS_T) void apgm_itstat_kernel(
    const double* __restrict__ obj, int nobj, double obj_scale, const double* __restrict__ tv,
    int ntv, double lam, const double* __restrict__ res, int nres, double* __restrict__ hist,
    int32_t* __restrict__ counter, int hist_cap) {
  apgm_stats_block<ITS_T>(obj, nobj, obj_scale, tv, ntv, lam, res, nres, hist, counter,
                          hist_cap);
}

// ---------------------------------------------------------------- TV --

// Approximate TV prox of scico/functional/_tvnorm.py:183-236, all axes of
// an nd-array (nd <= 3, dims padded to 3 with leading 1s), evaluated per
// output pixel p:
//   vp = v edge-padded by one on every differenced axis (not if circular),
//   L_a(q) = s (vp(q) + vp(q + e_a)),  H_a(q) = s (vp(q + e_a) - vp(q))
//   (indices circular in the padded extent), s = 1/sqrt(2);
//   H' = shrink of H over the axes (L21, isotropic) or per axis (L1) with
//   threshold tau, on the unpadded region only;
//   out(p) = (1/K) sum_a [ s (L_a(p) + L_a(p - e_a)) + s (H'_a(p - e_a) - H'_a(p)) ].
struct TvDims {
  int d[3];    // array extents
  int P[3];    // padded (or circular) extents
  int first;   // first differenced axis (3 - nd)
  int circ;
};

__device__ __forceinline__ float tv_vp(const float* __restrict__ v, const TvDims& t, int q0, int q1,
                                       int q2) {
  // q in padded coordinates [0, P); edge replication maps P-1 -> d-1
  const int i0 = min(q0, t.d[0] - 1), i1 = min(q1, t.d[1] - 1), i2 = min(q2, t.d[2] - 1);
  return v[(static_cast<int64_t>(i0) * t.d[1] + i1) * t.d[2] + i2];
}

// Haar high-pass coefficients of all axes at q (shrunk if inside the
// region, threshold tau * w[q] when a weight map is given), and the
// low-pass of every axis at q.
__device__ __forceinline__ void tv_coeffs(const float* __restrict__ v, const TvDims& t,
                                          const int* q, float tau, int iso,
                                          const float* __restrict__ w, float* H, float* Lo) {
  const float s = 0.70710677f;  // float32(1 / sqrt(2))
  const float vq = tv_vp(v, t, q[0], q[1], q[2]);
  bool inside = true;
  for (int a = t.first; a < 3; ++a) {
    int qn[3] = {q[0], q[1], q[2]};
    qn[a] = (q[a] + 1 == t.P[a]) ? 0 : q[a] + 1;
    const float vn = tv_vp(v, t, qn[0], qn[1], qn[2]);
    H[a] = __fmul_rn(s, __fsub_rn(vn, vq));
    Lo[a] = __fmul_rn(s, __fadd_rn(vq, vn));
    if (!t.circ && q[a] >= t.d[a]) inside = false;
  }
  if (!inside) return;
  const float tq =
      w ? __fmul_rn(tau, w[(static_cast<int64_t>(q[0]) * t.d[1] + q[1]) * t.d[2] + q[2]]) : tau;
  if (iso) {
    float hv[3];
    int K = 0;
    for (int a = t.first; a < 3; ++a) hv[K++] = H[a];
    l21_shrink(hv, K, tq);
    K = 0;
    for (int a = t.first; a < 3; ++a) H[a] = hv[K++];
  } else {
    for (int a = t.first; a < 3; ++a) H[a] = l1_shrink(H[a], tq);
  }
}

// Pixel coordinates of linear index e (32-bit arithmetic when the volume
// has fewer than 2^32 pixels).
__device__ __forceinline__ void tv_split(int64_t e, const TvDims& t, bool small, int (&p)[3]) {
  if (small) {
    const unsigned u = static_cast<unsigned>(e), d2 = static_cast<unsigned>(t.d[2]);
    const unsigned d1 = static_cast<unsigned>(t.d[1]);
    const unsigned q = u / d2;
    p[2] = static_cast<int>(u - q * d2);
    p[1] = static_cast<int>(q % d1);
    p[0] = static_cast<int>(q / d1);
  } else {
    p[0] = static_cast<int>(e / (static_cast<int64_t>(t.d[1]) * t.d[2]));
    p[1] = static_cast<int>((e / t.d[2]) % t.d[1]);
    p[2] = static_cast<int>(e % t.d[2]);
  }
}

__global__ __launch_bounds__(256) void tv_prox_kernel(const float* __restrict__ v,
                                                      float* __restrict__ out, TvDims t,
                                                      float tau, float invK, int iso,
                                                      const float* __restrict__ w,
                                                      const int32_t* pred) {
  if (pred_off(pred)) return;
  const int64_t n = static_cast<int64_t>(t.d[0]) * t.d[1] * t.d[2];
  const float s = 0.70710677f;
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < n; e += 256ll * gridDim.x) {
    int p[3];
    tv_split(e, t, n <= 0xffffffffll, p);
    float Hp[3], Lp[3];
    tv_coeffs(v, t, p, tau, iso, w, Hp, Lp);
    float acc = 0.0f;
    for (int a = t.first; a < 3; ++a) {
      int q[3] = {p[0], p[1], p[2]};
      q[a] = (p[a] == 0) ? t.P[a] - 1 : p[a] - 1;
      float Hq[3], Lq[3];
      tv_coeffs(v, t, q, tau, iso, w, Hq, Lq);
      const float lo = __fmul_rn(s, __fadd_rn(Lp[a], Lq[a]));
      const float hi = __fmul_rn(s, __fsub_rn(Hq[a], Hp[a]));
      const float term = __fadd_rn(lo, hi);
      acc = (a == t.first) ? term : __fadd_rn(acc, term);
    }
    out[e] = __fmul_rn(invK, acc);
  }
}

// Two-dimensional (M, N) case on 16 x 16 output tiles.  The padded image
// around the tile (padded rows and columns i0 - 1 .. i0 + 16, taken modulo
// P and edge-replicated) is staged in LDS, the (shrunk) coefficients of the
// 17 x 17 padded locations the tile needs are computed once each instead
// of once per neighbouring output, and every output combines the
// coefficients at p, p - e0 and p - e1 with the same operations, in the
// same order, as tv_prox_kernel (so the results are bitwise equal).
//
// APGM (EXT = 1): the AcceleratedPGM step that follows the prox
// (scico/optimize/_pgm.py:221-240) and the TV value of its result, fused.
// The tile computes one extra row and column of outputs (the first of the
// next tiles, bitwise as they compute them) so that every pixel's forward
// differences are at hand.  With out holding x_prev and vm the previous
// extrapolated point v, each stored output x gives the partial sums of
// (x - v)^2 (the fixed-point residual) and of the TV norm of x (as
// tv_norm_kernel: sqrt of the squared differences or their absolute
// values, float64), one each per workgroup, and v <- x + c (x - x_prev),
// with the operations of the unfused path.
constexpr int TVT = 16;
template <int EXT>
__global__ __launch_bounds__(256) void tv_prox2d_kernel(const float* __restrict__ v,
                                                        float* __restrict__ out, TvDims t,
                                                        float tau, float invK, int iso,
                                                        const float* __restrict__ w,
                                                        const int32_t* pred,
                                                        float* __restrict__ vm, float c,
                                                        double* __restrict__ res_ws,
                                                        double* __restrict__ tv_ws) {
  if (pred_off(pred)) return;
  constexpr int NO = TVT + EXT;  // outputs per tile side
  __shared__ float vs[NO + 2][NO + 3];
  __shared__ float cH[2][NO + 1][NO + 2], cL[2][NO + 1][NO + 2];
  const float s = 0.70710677f;
  const int M = t.d[1], N = t.d[2], P0 = t.P[1], P1 = t.P[2];
  // tiles in row-major order, consecutive ones on one XCD (workgroups go to
  // the 8 XCDs in turn): neighbouring tiles share the cache lines of their
  // rows and halos (a bijection of the first 8 * floor(G / 8) workgroups)
  int bx = blockIdx.x, by = blockIdx.y;
  {
    const int G = gridDim.x * gridDim.y, G8 = (G / 8) * 8;
    int L = by * gridDim.x + bx;
    if (L < G8) L = (L % 8) * (G8 / 8) + L / 8;
    by = L / gridDim.x;
    bx = L - by * gridDim.x;
  }
  const int i0 = by * TVT, j0 = bx * TVT;
  const int tid = threadIdx.x;
  // q >= -1 here; the modulo is left for images smaller than a tile
  auto wrap = [](int q, int P) {
    if (q < 0) return q + P;
    if (q >= P) {
      q -= P;
      if (q >= P) q %= P;
    }
    return q;
  };
  for (int e = tid; e < (NO + 2) * (NO + 2); e += 256) {
    const int r = e / (NO + 2), cc = e - r * (NO + 2);
    const int qi = wrap(i0 - 1 + r, P0), qj = wrap(j0 - 1 + cc, P1);
    vs[r][cc] = v[static_cast<int64_t>(min(qi, M - 1)) * N + min(qj, N - 1)];
  }
  __syncthreads();
  for (int e = tid; e < (NO + 1) * (NO + 1); e += 256) {
    const int r = e / (NO + 1), cc = e - r * (NO + 1);
    const float vq = vs[r][cc], vn0 = vs[r + 1][cc], vn1 = vs[r][cc + 1];
    float H[2] = {__fmul_rn(s, __fsub_rn(vn0, vq)), __fmul_rn(s, __fsub_rn(vn1, vq))};
    const float L0 = __fmul_rn(s, __fadd_rn(vq, vn0)), L1 = __fmul_rn(s, __fadd_rn(vq, vn1));
    const int qi = wrap(i0 - 1 + r, P0), qj = wrap(j0 - 1 + cc, P1);
    if (t.circ || (qi < M && qj < N)) {
      const float tq = w ? __fmul_rn(tau, w[static_cast<int64_t>(qi) * N + qj]) : tau;
      if (iso) {
        l21_shrink(H, 2, tq);
      } else {
        H[0] = l1_shrink(H[0], tq);
        H[1] = l1_shrink(H[1], tq);
      }
    }
    cH[0][r][cc] = H[0];
    cH[1][r][cc] = H[1];
    cL[0][r][cc] = L0;
    cL[1][r][cc] = L1;
  }
  __syncthreads();
  // output (ty, tx) of the tile: p at (ty + 1, tx + 1), p - e0 at
  // (ty, tx + 1), p - e1 at (ty + 1, tx)
  auto output = [&](int ty, int tx) {
    const float lo0 = __fmul_rn(s, __fadd_rn(cL[0][ty + 1][tx + 1], cL[0][ty][tx + 1]));
    const float hi0 = __fmul_rn(s, __fsub_rn(cH[0][ty][tx + 1], cH[0][ty + 1][tx + 1]));
    const float lo1 = __fmul_rn(s, __fadd_rn(cL[1][ty + 1][tx + 1], cL[1][ty + 1][tx]));
    const float hi1 = __fmul_rn(s, __fsub_rn(cH[1][ty + 1][tx], cH[1][ty + 1][tx + 1]));
    const float acc = __fadd_rn(__fadd_rn(lo0, hi0), __fadd_rn(lo1, hi1));
    return __fmul_rn(invK, acc);
  };
  const int ty = tid / TVT, tx = tid - ty * TVT;
  const int pi = i0 + ty, pj = j0 + tx;
  if constexpr (EXT == 0) {
    if (pi < M && pj < N) out[static_cast<int64_t>(pi) * N + pj] = output(ty, tx);
  } else {
    __shared__ float xs[NO][NO + 1];
    for (int e = tid; e < NO * NO; e += 256) {
      const int r = e / NO, cc = e - r * NO;
      xs[r][cc] = output(r, cc);
    }
    __syncthreads();
    double res = 0.0, tv = 0.0;
    if (pi < M && pj < N) {
      const float xn = xs[ty][tx];
      const int64_t e = static_cast<int64_t>(pi) * N + pj;
      const float xo = out[e], vo = vm[e];
      const double d = static_cast<double>(xn) - static_cast<double>(vo);
      res = d * d;
      vm[e] = __fadd_rn(xn, __fmul_rn(c, __fsub_rn(xn, xo)));
      out[e] = xn;
      const bool circ = t.circ != 0;
      const float g0 = fd::diff_fwd(xn, xs[ty + 1][tx], pi, M, circ);
      const float g1 = fd::diff_fwd(xn, xs[ty][tx + 1], pj, N, circ);
      if (iso)
        tv = sqrt(static_cast<double>(g0) * g0 + static_cast<double>(g1) * g1);
      else
        tv = fabs(static_cast<double>(g0)) + fabs(static_cast<double>(g1));
    }
    __shared__ double sh[4];
    res = block_sum_f64(res, sh);
    __syncthreads();
    tv = block_sum_f64(tv, sh);
    if (tid == 0) {
      res_ws[by * gridDim.x + bx] = res;
      tv_ws[by * gridDim.x + bx] = tv;
    }
  }
}

// ------------------------------------------------- TV prox VJP (autodiff) --
// Cotangents of out = tv_prox(v; tau * w) with respect to v and w, for the
// differentiable unrolled reconstruction.  Pass 1, per padded coefficient
// location q: the cotangents of the low-pass Lo_a(q) and (through the
// shrink Jacobian) of the high-pass H_a(q), plus dw(q) for q inside;
// pass 2, per pixel p: dv(p) gathered over the padded locations that
// replicate p.  A cotangent gout(x) is zero outside [0, d).

__device__ __forceinline__ float tv_gout(const float* __restrict__ g, const TvDims& t, int q0,
                                         int q1, int q2) {
  if (q0 >= t.d[0] || q1 >= t.d[1] || q2 >= t.d[2]) return 0.0f;
  return g[(static_cast<int64_t>(q0) * t.d[1] + q1) * t.d[2] + q2];
}

__global__ __launch_bounds__(256) void tv_vjp_coef_kernel(
    const float* __restrict__ v, const float* __restrict__ gout, const float* __restrict__ w,
    TvDims t, float tau, float invK, int iso, float* __restrict__ HB, float* __restrict__ LB,
    float* __restrict__ gw) {
  const int64_t np = static_cast<int64_t>(t.P[0]) * t.P[1] * t.P[2];
  const float s = 0.70710677f;
  const float c = s * invK;
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < np; e += 256ll * gridDim.x) {
    const int q[3] = {static_cast<int>(e / (static_cast<int64_t>(t.P[1]) * t.P[2])),
                      static_cast<int>((e / t.P[2]) % t.P[1]), static_cast<int>(e % t.P[2])};
    const float vq = tv_vp(v, t, q[0], q[1], q[2]);
    const float g0 = tv_gout(gout, t, q[0], q[1], q[2]);
    float h[3] = {0.f, 0.f, 0.f}, hb[3] = {0.f, 0.f, 0.f};
    bool inside = true;
    for (int a = t.first; a < 3; ++a) {
      int qn[3] = {q[0], q[1], q[2]};
      qn[a] = (q[a] + 1 == t.P[a]) ? 0 : q[a] + 1;
      const float vn = tv_vp(v, t, qn[0], qn[1], qn[2]);
      h[a] = s * (vn - vq);
      const float gn = tv_gout(gout, t, qn[0], qn[1], qn[2]);
      LB[(a - t.first) * np + e] = c * (g0 + gn);
      hb[a] = c * (gn - g0);  // cotangent of the shrunk coefficient
      if (!t.circ && q[a] >= t.d[a]) inside = false;
    }
    float dtau = 0.0f;
    if (inside) {
      const float tq = w ? tau * w[(static_cast<int64_t>(q[0]) * t.d[1] + q[1]) * t.d[2] + q[2]]
                         : tau;
      if (iso) {
        float r2 = 0.0f;
        for (int a = t.first; a < 3; ++a) r2 += h[a] * h[a];
        const float r = sqrtf(r2);
        if (r > tq) {
          float hdot = 0.0f;
          for (int a = t.first; a < 3; ++a) hdot += h[a] * hb[a];
          const float k1 = 1.0f - tq / r, k2 = tq / (r * r * r);
          for (int a = t.first; a < 3; ++a) hb[a] = k1 * hb[a] + k2 * h[a] * hdot;
          dtau = -hdot / r;
        } else {
          for (int a = t.first; a < 3; ++a) hb[a] = 0.0f;
        }
      } else {
        for (int a = t.first; a < 3; ++a) {
          if (fabsf(h[a]) > tq) {
            dtau -= (h[a] > 0.0f ? hb[a] : -hb[a]);
          } else {
            hb[a] = 0.0f;
          }
        }
      }
      if (gw) gw[(static_cast<int64_t>(q[0]) * t.d[1] + q[1]) * t.d[2] + q[2]] = tau * dtau;
    }
    for (int a = t.first; a < 3; ++a) HB[(a - t.first) * np + e] = hb[a];
  }
}

__global__ __launch_bounds__(256) void tv_vjp_v_kernel(const float* __restrict__ HB,
                                                       const float* __restrict__ LB, TvDims t,
                                                       float* __restrict__ gv) {
  const int64_t n = static_cast<int64_t>(t.d[0]) * t.d[1] * t.d[2];
  const int64_t np = static_cast<int64_t>(t.P[0]) * t.P[1] * t.P[2];
  const float s = 0.70710677f;
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < n; e += 256ll * gridDim.x) {
    int p[3];
    tv_split(e, t, n <= 0xffffffffll, p);
    // padded locations replicating p: q_a in {p_a} or {p_a, d_a} on the last index
    int nalt[3];
    for (int a = 0; a < 3; ++a) nalt[a] = (!t.circ && a >= t.first && p[a] == t.d[a] - 1) ? 2 : 1;
    float acc = 0.0f;
    for (int i0 = 0; i0 < nalt[0]; ++i0)
      for (int i1 = 0; i1 < nalt[1]; ++i1)
        for (int i2 = 0; i2 < nalt[2]; ++i2) {
          const int q[3] = {p[0] + i0, p[1] + i1, p[2] + i2};
          const int64_t eq = (static_cast<int64_t>(q[0]) * t.P[1] + q[1]) * t.P[2] + q[2];
          for (int a = t.first; a < 3; ++a) {
            const int64_t o = (a - t.first) * np;
            int qm[3] = {q[0], q[1], q[2]};
            qm[a] = (q[a] == 0) ? t.P[a] - 1 : q[a] - 1;
            const int64_t em = (static_cast<int64_t>(qm[0]) * t.P[1] + qm[1]) * t.P[2] + qm[2];
            // vp(q) enters Lo_a(q), H_a(q) as the left value and
            // Lo_a(q - e_a), H_a(q - e_a) as the right value
            acc += s * (LB[o + eq] - HB[o + eq]) + s * (LB[o + em] + HB[o + em]);
          }
        }
    gv[e] = acc;
  }
}

// Partials of TV(x) = norm(FD(x)) with FD append=0 or circular
// (scico/functional/_tvnorm.py:89-115): L21 over the axes (iso) or L1.
__global__ __launch_bounds__(256) void tv_norm_kernel(const float* __restrict__ x, TvDims t,
                                                      int iso, double* __restrict__ ws,
                                                      const int32_t* pred) {
  if (pred_off(pred)) return;
  __shared__ double sh[4];
  const int64_t n = static_cast<int64_t>(t.d[0]) * t.d[1] * t.d[2];
  const int64_t st[3] = {static_cast<int64_t>(t.d[1]) * t.d[2], t.d[2], 1};
  double acc = 0.0;
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < n; e += 256ll * gridDim.x) {
    int p[3];
    tv_split(e, t, n <= 0xffffffffll, p);
    const float xc = x[e];
    double ss = 0.0;
    for (int a = t.first; a < 3; ++a) {
      const int kn = (p[a] + 1 < t.d[a]) ? p[a] + 1 : 0;
      const float g = fd::diff_fwd(xc, x[e + (kn - p[a]) * st[a]], p[a], t.d[a], t.circ != 0);
      if (iso)
        ss += static_cast<double>(g) * g;
      else
        acc += fabs(static_cast<double>(g));
    }
    if (iso) acc += sqrt(ss);
  }
  acc = block_sum_f64(acc, sh);
  if (threadIdx.x == 0) ws[blockIdx.x] = acc;
}

// ------------------------------------------------------------- batched --

struct SegBlock {
  int s, vb, vg;
  int64_t o, n;
  int M, N;
};

__device__ __forceinline__ SegBlock seg_block(const SegTabs& t) {
  SegBlock r;
  r.s = t.blk_seg[blockIdx.x];
  r.vb = blockIdx.x - t.blk_first[r.s];
  r.vg = t.G[r.s];
  r.o = t.off[r.s];
  const int2 mn = t.MN[r.s];
  r.M = mn.x;
  r.N = mn.y;
  r.n = static_cast<int64_t>(mn.x) * mn.y;
  return r;
}

__global__ __launch_bounds__(256) void cg_start_b(SegTabs t, const float* __restrict__ b,
                                                  const float* __restrict__ Ax,
                                                  float* __restrict__ r, float* __restrict__ p,
                                                  double* __restrict__ ws) {
  const SegBlock q = seg_block(t);
  cg_start_body(q.n, b + q.o, Ax + q.o, r + q.o, p + q.o, ws + 2 * t.woff[q.s], q.vb, q.vg);
}

__global__ __launch_bounds__(256) void cg_start_fin_b(SegTabs t, const double* __restrict__ ws,
                                                      double* __restrict__ state,
                                                      int32_t* __restrict__ flags, int maxiter,
                                                      double tol, double atol) {
  const int s = blockIdx.x;
  const int SS = ST_NUM + 2 * (maxiter + 1);
  cg_start_fin_body(t.G[s], ws + 2 * t.woff[s], state + static_cast<int64_t>(s) * SS,
                    flags + static_cast<int64_t>(s) * (maxiter + 1), maxiter, tol, atol);
}

__global__ __launch_bounds__(256) void cg_xr_b(SegTabs t, int k, int maxiter,
                                               float* __restrict__ x, float* __restrict__ r,
                                               const float* __restrict__ p,
                                               const float* __restrict__ Ap,
                                               double* __restrict__ state,
                                               const int32_t* __restrict__ flags,
                                               const double* __restrict__ pap,
                                               double* __restrict__ ws) {
  const SegBlock q = seg_block(t);
  const int SS = ST_NUM + 2 * (maxiter + 1);
  cg_xr_body(q.n, k, maxiter, x + q.o, r + q.o, p + q.o, Ap + q.o,
             state + static_cast<int64_t>(q.s) * SS,
             flags + static_cast<int64_t>(q.s) * (maxiter + 1), pap + t.pap_off[q.s],
             t.pap_n[q.s], ws + t.woff[q.s], q.vb, q.vg);
}

__global__ __launch_bounds__(256) void cg_p_b(SegTabs t, int k, int maxiter,
                                              const float* __restrict__ r, float* __restrict__ p,
                                              double* __restrict__ state,
                                              int32_t* __restrict__ flags,
                                              const double* __restrict__ ws) {
  const SegBlock q = seg_block(t);
  const int SS = ST_NUM + 2 * (maxiter + 1);
  cg_p_body(q.n, k, maxiter, r + q.o, p + q.o, state + static_cast<int64_t>(q.s) * SS,
            flags + static_cast<int64_t>(q.s) * (maxiter + 1), ws + t.woff[q.s], q.vg, q.vb,
            q.vg);
}

__global__ __launch_bounds__(256) void admm_zu_b(SegTabs t, const float* __restrict__ x,
                                                 float* __restrict__ z, float* __restrict__ u,
                                                 float* __restrict__ zold, float thr, int circ,
                                                 int iso, double* __restrict__ ws_prml,
                                                 double* __restrict__ ws_g) {
  const SegBlock q = seg_block(t);
  admm_zu2d_body(q.M, q.N, x + q.o, z + 2 * q.o, u + 2 * q.o, zold + 2 * q.o, thr, circ, iso,
                 ws_prml + t.woff[q.s], ws_g + t.woff[q.s], q.vb, q.vg);
}

__global__ __launch_bounds__(256) void admm_dual_b(SegTabs t, const float* __restrict__ z,
                                                   const float* __restrict__ zold, float rho,
                                                   int circ, double* __restrict__ ws) {
  const SegBlock q = seg_block(t);
  admm_dual2d_body(q.M, q.N, z + 2 * q.o, zold + 2 * q.o, rho, circ, ws + t.woff[q.s], q.vb,
                   q.vg);
}

__global__ __launch_bounds__(256) void admm_rhs_b(SegTabs t, const float* __restrict__ aty,
                                                  const float* __restrict__ z,
                                                  const float* __restrict__ u, float rho,
                                                  int circ, float* __restrict__ out) {
  const SegBlock q = seg_block(t);
  admm_rhs2d_body(q.M, q.N, aty + q.o, z + 2 * q.o, u + 2 * q.o, rho, circ, out + q.o, q.vb,
                  q.vg);
}

__global__ __launch_bounds__(256) void admm_itstat_b(
    SegTabs t, const double* __restrict__ res, const double* __restrict__ g,
    const double* __restrict__ prml, const double* __restrict__ dual,
    const double* __restrict__ cg_state, int maxiter, double lam, double rho,
    double* __restrict__ hist, int32_t* __restrict__ counters, int hist_cap) {
  const int s = blockIdx.x;
  const int SS = ST_NUM + 2 * (maxiter + 1);
  const int w = t.woff[s], G = t.G[s];
  admm_itstat_body(res + t.res_off[s], t.res_n[s], g + w, G, prml + w, G, dual + w, G,
                   cg_state ? cg_state + static_cast<int64_t>(s) * SS : nullptr, lam, rho,
                   hist + static_cast<int64_t>(s) * 5 * hist_cap, counters + s, hist_cap);
}

int batch_cg_start(const SegTabs& t, const float* b, const float* Ax, float* r, float* p,
                   double* state, int32_t* flags, int maxiter, double tol, double atol,
                   double* ws, hipStream_t s) {
  hipLaunchKernelGGL(cg_start_b, dim3(t.nblk), dim3(256), 0, s, t, b, Ax, r, p, ws);
  SCXR_LAUNCHED("batch cg start");
  hipLaunchKernelGGL(cg_start_fin_b, dim3(t.ns), dim3(256), 0, s, t, ws, state, flags, maxiter,
                     tol, atol);
  SCXR_LAUNCHED("batch cg start fin");
  return SCXR_OK;
}

int batch_cg_iter(const SegTabs& t, int k, int maxiter, float* x, float* r, float* p,
                  const float* Ap, double* state, int32_t* flags, const double* pap_parts,
                  double* ws, hipStream_t s) {
  hipLaunchKernelGGL(cg_xr_b, dim3(t.nblk), dim3(256), 0, s, t, k, maxiter, x, r, p, Ap, state,
                     flags, pap_parts, ws);
  SCXR_LAUNCHED("batch cg xr");
  hipLaunchKernelGGL(cg_p_b, dim3(t.nblk), dim3(256), 0, s, t, k, maxiter, r, p, state, flags,
                     ws);
  SCXR_LAUNCHED("batch cg p");
  return SCXR_OK;
}

int batch_admm_zu(const SegTabs& t, const float* x, float* z, float* u, float* zold, float thr,
                  int circ, int iso, double* ws_prml, double* ws_g, hipStream_t s) {
  hipLaunchKernelGGL(admm_zu_b, dim3(t.nblk), dim3(256), 0, s, t, x, z, u, zold, thr, circ, iso,
                     ws_prml, ws_g);
  SCXR_LAUNCHED("batch admm zu");
  return SCXR_OK;
}

int batch_admm_rhs(const SegTabs& t, const float* aty, const float* z, const float* u, float rho,
                   int circ, float* out, hipStream_t s) {
  hipLaunchKernelGGL(admm_rhs_b, dim3(t.nblk), dim3(256), 0, s, t, aty, z, u, rho, circ, out);
  SCXR_LAUNCHED("batch admm rhs");
  return SCXR_OK;
}

int batch_admm_dual(const SegTabs& t, const float* z, const float* zold, float rho, int circ,
                    double* ws, hipStream_t s) {
  hipLaunchKernelGGL(admm_dual_b, dim3(t.nblk), dim3(256), 0, s, t, z, zold, rho, circ, ws);
  SCXR_LAUNCHED("batch admm dual");
  return SCXR_OK;
}

int batch_admm_itstat(const SegTabs& t, const double* res, const double* g, const double* prml,
                      const double* dual, const double* cg_state, int maxiter, double lam,
                      double rho, double* hist, int32_t* counters, int hist_cap,
                      hipStream_t s) {
  hipLaunchKernelGGL(admm_itstat_b, dim3(t.ns), dim3(256), 0, s, t, res, g, prml, dual, cg_state,
                     maxiter, lam, rho, hist, counters, hist_cap);
  SCXR_LAUNCHED("batch admm itstat");
  return SCXR_OK;
}

unsigned elem_grid(int64_t n) {
  const int64_t b = (n + 1023) / 1024;
  return static_cast<unsigned>(b < 1 ? 1 : (b > ELEM_BLOCKS ? ELEM_BLOCKS : b));
}

}  // namespace solver
}  // namespace scxr

// ------------------------------------------------------------- C ABI --

extern "C" {

int scxr_fd_forward(int32_t ndim, const int32_t* shape, const float* x, float* out,
                    int32_t circular, const int32_t* pred, void* stream) {
  SCXR_REQUIRE(ndim >= 1 && ndim <= 3 && shape && x && out, SCXR_ERR_ARG, "bad fd arguments");
  int d[3] = {1, 1, 1};
  for (int k = 0; k < ndim; ++k) d[3 - ndim + k] = shape[k];
  const int64_t n = static_cast<int64_t>(d[0]) * d[1] * d[2];
  SCXR_REQUIRE(n > 0, SCXR_ERR_SHAPE, "empty array");
  hipLaunchKernelGGL(solver::fd_fwd_kernel, dim3(solver::elem_grid(n)), dim3(256), 0,
                     as_stream(stream), ndim, d[0], d[1], d[2], x, out, circular, pred);
  SCXR_LAUNCHED("fd forward");
  return SCXR_OK;
}

int scxr_fd_adjoint(int32_t ndim, const int32_t* shape, const float* g, float* out,
                    int32_t circular, float scale, const int32_t* pred, void* stream) {
  SCXR_REQUIRE(ndim >= 1 && ndim <= 3 && shape && g && out, SCXR_ERR_ARG, "bad fd arguments");
  int d[3] = {1, 1, 1};
  for (int k = 0; k < ndim; ++k) d[3 - ndim + k] = shape[k];
  const int64_t n = static_cast<int64_t>(d[0]) * d[1] * d[2];
  SCXR_REQUIRE(n > 0, SCXR_ERR_SHAPE, "empty array");
  hipLaunchKernelGGL(solver::fd_adj_kernel, dim3(solver::elem_grid(n)), dim3(256), 0,
                     as_stream(stream), ndim, d[0], d[1], d[2], g, out, circular, scale, pred);
  SCXR_LAUNCHED("fd adjoint");
  return SCXR_OK;
}

int scxr_fd_gram(int32_t ndim, const int32_t* shape, float rho, const float* x, float* out,
                 int32_t circular, const int32_t* pred, void* stream) {
  SCXR_REQUIRE(ndim >= 1 && ndim <= 3 && shape && x && out, SCXR_ERR_ARG, "bad fd arguments");
  int d[3] = {1, 1, 1};
  for (int k = 0; k < ndim; ++k) d[3 - ndim + k] = shape[k];
  const int64_t n = static_cast<int64_t>(d[0]) * d[1] * d[2];
  SCXR_REQUIRE(n > 0, SCXR_ERR_SHAPE, "empty array");
  hipLaunchKernelGGL(solver::fd_gram_kernel, dim3(solver::elem_grid(n)), dim3(256), 0,
                     as_stream(stream), ndim, d[0], d[1], d[2], x, out, circular, rho, pred);
  SCXR_LAUNCHED("fd gram");
  return SCXR_OK;
}

int scxr_l21_prox(int32_t K, int64_t n, const float* v, float lam, float* out,
                  const int32_t* pred, void* stream) {
  SCXR_REQUIRE(K >= 1 && K <= solver::MAX_K && n >= 0 && v && out, SCXR_ERR_ARG,
               "bad l21_prox arguments (K must be in [1, %d])", solver::MAX_K);
  if (n == 0) return SCXR_OK;
  hipLaunchKernelGGL(solver::l21_prox_kernel, dim3(solver::elem_grid(n)), dim3(256), 0,
                     as_stream(stream), K, n, v, lam, out, pred);
  SCXR_LAUNCHED("l21 prox");
  return SCXR_OK;
}

int scxr_l1_prox(int64_t n, const float* v, float lam, float* out, const int32_t* pred,
                 void* stream) {
  SCXR_REQUIRE(n >= 0 && v && out, SCXR_ERR_ARG, "bad l1_prox arguments");
  if (n == 0) return SCXR_OK;
  hipLaunchKernelGGL(solver::l1_prox_kernel, dim3(solver::elem_grid(n)), dim3(256), 0,
                     as_stream(stream), n, v, lam, out, pred);
  SCXR_LAUNCHED("l1 prox");
  return SCXR_OK;
}

int scxr_cg_start(int64_t n, const float* b, const float* Ax, float* r, float* p, double* state,
                  int32_t* flags, int32_t maxiter, double tol, double atol, double* ws,
                  void* stream) {
  SCXR_REQUIRE(n > 0 && b && Ax && r && p && state && flags && ws && maxiter >= 0, SCXR_ERR_ARG,
               "bad cg_start arguments");
  const unsigned g = solver::elem_grid(n);
  hipLaunchKernelGGL(solver::cg_start_kernel, dim3(g), dim3(256), 0, as_stream(stream), n, b, Ax,
                     r, p, ws);
  SCXR_LAUNCHED("cg start");
  hipLaunchKernelGGL(solver::cg_start_fin, dim3(1), dim3(256), 0, as_stream(stream),
                     static_cast<int>(g), ws, state, flags, maxiter, tol, atol);
  SCXR_LAUNCHED("cg start fin");
  return SCXR_OK;
}

int scxr_cg_iter(int64_t n, int32_t k, int32_t maxiter, float* x, float* r, float* p,
                 const float* Ap, double* state, int32_t* flags, const double* pap_parts,
                 int32_t pap_n, double* ws, void* stream) {
  SCXR_REQUIRE(n > 0 && x && r && p && Ap && state && flags && pap_parts && ws, SCXR_ERR_ARG,
               "bad cg_iter arguments");
  SCXR_REQUIRE(k >= 0 && k < maxiter && pap_n >= 1, SCXR_ERR_ARG, "bad cg_iter index");
  const unsigned g = solver::elem_grid(n);
  hipLaunchKernelGGL(solver::cg_xr_kernel, dim3(g), dim3(256), 0, as_stream(stream), n, k,
                     maxiter, x, r, p, Ap, state, flags, pap_parts, pap_n, ws);
  SCXR_LAUNCHED("cg xr");
  hipLaunchKernelGGL(solver::cg_p_kernel, dim3(g), dim3(256), 0, as_stream(stream), n, k, maxiter,
                     r, p, state, flags, ws, static_cast<int>(g));
  SCXR_LAUNCHED("cg p");
  return SCXR_OK;
}

int scxr_cg_iter_t(int32_t M, int32_t N, int32_t k, int32_t maxiter, float* x, float* r,
                   float* p, float* pT, const float* Ap, double* state, int32_t* flags,
                   const double* pap_parts, int32_t pap_n, double* ws, void* stream) {
  SCXR_REQUIRE(M > 0 && N > 0 && x && r && p && pT && Ap && state && flags && pap_parts && ws,
               SCXR_ERR_ARG, "bad cg_iter_t arguments");
  SCXR_REQUIRE(k >= 0 && k < maxiter && pap_n >= 1, SCXR_ERR_ARG, "bad cg_iter_t index");
  const int64_t n = static_cast<int64_t>(M) * N;
  const unsigned g = solver::elem_grid(n);
  hipLaunchKernelGGL(solver::cg_xr_kernel, dim3(g), dim3(256), 0, as_stream(stream), n, k,
                     maxiter, x, r, p, Ap, state, flags, pap_parts, pap_n, ws);
  SCXR_LAUNCHED("cg xr");
  hipLaunchKernelGGL(solver::cg_p_t_kernel, dim3(ceil_div(N, 32), ceil_div(M, 32)), dim3(256), 0,
                     as_stream(stream), M, N, k, maxiter, r, p, pT, state, flags, ws,
                     static_cast<int>(g));
  SCXR_LAUNCHED("cg p (with transposed copy)");
  return SCXR_OK;
}

int scxr_admm_rhs2d(int32_t M, int32_t N, const float* aty, const float* z, const float* u,
                    float rho, int32_t circular, float* out, void* stream) {
  SCXR_REQUIRE(M > 0 && N > 0 && aty && z && u && out, SCXR_ERR_ARG, "bad admm_rhs2d arguments");
  const int64_t n = static_cast<int64_t>(M) * N;
  hipLaunchKernelGGL(solver::admm_rhs2d_kernel, dim3(solver::elem_grid(n)), dim3(256), 0,
                     as_stream(stream), M, N, aty, z, u, rho, circular, out);
  SCXR_LAUNCHED("admm rhs2d");
  return SCXR_OK;
}

int scxr_admm_zu2d(int32_t M, int32_t N, const float* x, float* z, float* u, float* zold,
                   float thr, int32_t circular, int32_t isotropic, double* ws_prml,
                   double* ws_g, void* stream) {
  SCXR_REQUIRE(M > 0 && N > 0 && x && z && u && zold && ws_prml && ws_g, SCXR_ERR_ARG,
               "bad admm_zu2d arguments");
  const int64_t n = static_cast<int64_t>(M) * N;
  hipLaunchKernelGGL(solver::admm_zu2d_kernel, dim3(solver::elem_grid(n)), dim3(256), 0,
                     as_stream(stream), M, N, x, z, u, zold, thr, circular, isotropic, ws_prml,
                     ws_g);
  SCXR_LAUNCHED("admm zu2d");
  return SCXR_OK;
}

int scxr_admm_dual2d(int32_t M, int32_t N, const float* z, const float* zold, float rho,
                     int32_t circular, double* ws, void* stream) {
  SCXR_REQUIRE(M > 0 && N > 0 && z && zold && ws, SCXR_ERR_ARG, "bad admm_dual2d arguments");
  const int64_t n = static_cast<int64_t>(M) * N;
  hipLaunchKernelGGL(solver::admm_dual2d_kernel, dim3(solver::elem_grid(n)), dim3(256), 0,
                     as_stream(stream), M, N, z, zold, rho, circular, ws);
  SCXR_LAUNCHED("admm dual2d");
  return SCXR_OK;
}

static int nd_dims(int32_t ndim, const int32_t* shape, int d[3]) {
  SCXR_REQUIRE(ndim >= 1 && ndim <= 3 && shape, SCXR_ERR_ARG, "bad nd shape (ndim in 1..3)");
  d[0] = d[1] = d[2] = 1;
  for (int k = 0; k < ndim; ++k) {
    SCXR_REQUIRE(shape[k] > 0, SCXR_ERR_SHAPE, "empty array");
    d[3 - ndim + k] = shape[k];
  }
  return SCXR_OK;
}

int scxr_admm_rhs_nd(int32_t ndim, const int32_t* shape, const float* base, const float* z,
                     const float* u, float rho, int32_t circular, float* out, void* stream) {
  int d[3];
  int rc = nd_dims(ndim, shape, d);
  if (rc) return rc;
  SCXR_REQUIRE(base && z && u && out, SCXR_ERR_ARG, "bad admm_rhs_nd arguments");
  const int64_t n = static_cast<int64_t>(d[0]) * d[1] * d[2];
  hipLaunchKernelGGL(solver::admm_rhs_nd_kernel, dim3(solver::elem_grid(n)), dim3(256), 0,
                     as_stream(stream), ndim, d[0], d[1], d[2], base, z, u, rho, circular, out);
  SCXR_LAUNCHED("admm rhs nd");
  return SCXR_OK;
}

int scxr_admm_lhs_nd(int32_t ndim, const int32_t* shape, const float* x, const float* ata,
                     float rho, int32_t circular, float* out, double* dot_ws,
                     const int32_t* pred, void* stream) {
  int d[3];
  int rc = nd_dims(ndim, shape, d);
  if (rc) return rc;
  SCXR_REQUIRE(x && ata && out, SCXR_ERR_ARG, "bad admm_lhs_nd arguments");
  const int64_t n = static_cast<int64_t>(d[0]) * d[1] * d[2];
  hipLaunchKernelGGL(solver::admm_lhs_nd_kernel, dim3(solver::elem_grid(n)), dim3(256), 0,
                     as_stream(stream), ndim, d[0], d[1], d[2], x, ata, rho, circular, out,
                     dot_ws, pred);
  SCXR_LAUNCHED("admm lhs nd");
  return SCXR_OK;
}

int scxr_admm_zu_nd(int32_t ndim, const int32_t* shape, const float* x, float* z, float* u,
                    float* zold, float thr, int32_t circular, int32_t isotropic,
                    double* ws_prml, double* ws_g, void* stream) {
  int d[3];
  int rc = nd_dims(ndim, shape, d);
  if (rc) return rc;
  SCXR_REQUIRE(x && z && u && zold && ws_prml && ws_g, SCXR_ERR_ARG,
               "bad admm_zu_nd arguments");
  const int64_t n = static_cast<int64_t>(d[0]) * d[1] * d[2];
  hipLaunchKernelGGL(solver::admm_zu_nd_kernel, dim3(solver::elem_grid(n)), dim3(256), 0,
                     as_stream(stream), ndim, d[0], d[1], d[2], x, z, u, zold, thr, circular,
                     isotropic, ws_prml, ws_g);
  SCXR_LAUNCHED("admm zu nd");
  return SCXR_OK;
}

int scxr_admm_dual_nd(int32_t ndim, const int32_t* shape, const float* z, const float* zold,
                      float rho, int32_t circular, double* ws, void* stream) {
  int d[3];
  int rc = nd_dims(ndim, shape, d);
  if (rc) return rc;
  SCXR_REQUIRE(z && zold && ws, SCXR_ERR_ARG, "bad admm_dual_nd arguments");
  const int64_t n = static_cast<int64_t>(d[0]) * d[1] * d[2];
  hipLaunchKernelGGL(solver::admm_dual_nd_kernel, dim3(solver::elem_grid(n)), dim3(256), 0,
                     as_stream(stream), ndim, d[0], d[1], d[2], z, zold, rho, circular, ws);
  SCXR_LAUNCHED("admm dual nd");
  return SCXR_OK;
}

int scxr_admm_itstat(const double* res, int32_t nres, const double* g, int32_t ng,
                     const double* prml, int32_t nprml, const double* dual, int32_t ndual,
                     const double* cg_state, double lam, double rho, double* hist,
                     int32_t* counter, int32_t hist_cap, void* stream) {
  SCXR_REQUIRE(res && g && prml && dual && hist && counter, SCXR_ERR_ARG,
               "bad admm_itstat arguments");
  hipLaunchKernelGGL(solver::admm_itstat_kernel, dim3(1), dim3(256), 0, as_stream(stream), res,
                     nres, g, ng, prml, nprml, dual, ndual, cg_state, lam, rho, hist, counter,
                     hist_cap);
  SCXR_LAUNCHED("admm itstat");
  return SCXR_OK;
}

int32_t scxr_elem_blocks(int64_t n) { return static_cast<int32_t>(solver::elem_grid(n)); }

int scxr_apgm_itstat(const double* obj, int32_t nobj, double obj_scale, const double* tv,
                     int32_t ntv, double lam, const double* res, int32_t nres, double* hist,
                     int32_t* counter, int32_t hist_cap, void* stream) {
  SCXR_REQUIRE(obj && tv && res && hist && counter && nobj > 0 && ntv > 0 && nres > 0 &&
                   hist_cap > 0,
               SCXR_ERR_ARG, "bad apgm_itstat arguments");
  hipLaunchKernelGGL(solver::apgm_itstat_kernel, dim3(1), dim3(solver::ITS_T), 0,
                     as_stream(stream), obj,
                     nobj, obj_scale, tv, ntv, lam, res, nres, hist, counter, hist_cap);
  SCXR_LAUNCHED("apgm itstat");
  return SCXR_OK;
}

static int tv_dims(int32_t ndim, const int32_t* shape, int32_t circular, solver::TvDims& t) {
  SCXR_REQUIRE(ndim >= 1 && ndim <= 3 && shape, SCXR_ERR_ARG, "bad TV shape");
  for (int a = 0; a < 3; ++a) t.d[a] = 1;
  for (int k = 0; k < ndim; ++k) {
    SCXR_REQUIRE(shape[k] > 0, SCXR_ERR_SHAPE, "empty array");
    t.d[3 - ndim + k] = shape[k];
  }
  t.first = 3 - ndim;
  t.circ = circular != 0;
  for (int a = 0; a < 3; ++a) t.P[a] = (a >= t.first && !t.circ) ? t.d[a] + 1 : t.d[a];
  return SCXR_OK;
}

int scxr_tv_prox(int32_t ndim, const int32_t* shape, const float* v, float* out, float tau,
                 int32_t isotropic, int32_t circular, const int32_t* pred, void* stream) {
  SCXR_REQUIRE(v && out && v != out, SCXR_ERR_ARG, "bad tv_prox arguments (out must not alias v)");
  solver::TvDims t;
  const int rc = tv_dims(ndim, shape, circular, t);
  if (rc) return rc;
  const int64_t n = static_cast<int64_t>(t.d[0]) * t.d[1] * t.d[2];
  if (ndim == 2) {
    const dim3 g(ceil_div(t.d[2], solver::TVT), ceil_div(t.d[1], solver::TVT));
    hipLaunchKernelGGL(solver::tv_prox2d_kernel<0>, g, dim3(256), 0, as_stream(stream), v,
                       out, t, tau, 0.25f, isotropic, nullptr, pred, nullptr, 0.0f, nullptr,
                       nullptr);
  } else {
    hipLaunchKernelGGL(solver::tv_prox_kernel, dim3(solver::elem_grid(n)), dim3(256), 0,
                       as_stream(stream), v, out, t, tau, 1.0f / static_cast<float>(2 * ndim),
                       isotropic, nullptr, pred);
  }
  SCXR_LAUNCHED("tv prox");
  return SCXR_OK;
}

int32_t scxr_tv_prox_apgm_parts(int32_t M, int32_t N) {
  if (M <= 0 || N <= 0) return -1;
  return static_cast<int32_t>(ceil_div(M, solver::TVT) * ceil_div(N, solver::TVT));
}

int scxr_tv_prox_apgm(int32_t M, int32_t N, const float* u, float* x, float* v, float tau,
                      int32_t isotropic, int32_t circular, float c, double* res_ws,
                      double* tv_ws, const int32_t* pred, void* stream) {
  SCXR_REQUIRE(u && x && v && res_ws && tv_ws && u != x && u != v && x != v, SCXR_ERR_ARG,
               "bad tv_prox_apgm arguments (u, x and v must be distinct)");
  const int32_t shape[2] = {M, N};
  solver::TvDims t;
  const int rc = tv_dims(2, shape, circular, t);
  if (rc) return rc;
  const dim3 g(ceil_div(N, solver::TVT), ceil_div(M, solver::TVT));
  hipLaunchKernelGGL(solver::tv_prox2d_kernel<1>, g, dim3(256), 0, as_stream(stream), u, x, t,
                     tau, 0.25f, isotropic, nullptr, pred, v, c, res_ws, tv_ws);
  SCXR_LAUNCHED("tv prox apgm");
  return SCXR_OK;
}

int scxr_tv_prox_w(int32_t ndim, const int32_t* shape, const float* v, const float* w,
                   float* out, float tau, int32_t isotropic, int32_t circular,
                   const int32_t* pred, void* stream) {
  SCXR_REQUIRE(v && out && v != out, SCXR_ERR_ARG,
               "bad tv_prox_w arguments (out must not alias v)");
  solver::TvDims t;
  const int rc = tv_dims(ndim, shape, circular, t);
  if (rc) return rc;
  const int64_t n = static_cast<int64_t>(t.d[0]) * t.d[1] * t.d[2];
  if (ndim == 2) {
    const dim3 g(ceil_div(t.d[2], solver::TVT), ceil_div(t.d[1], solver::TVT));
    hipLaunchKernelGGL(solver::tv_prox2d_kernel<0>, g, dim3(256), 0, as_stream(stream), v,
                       out, t, tau, 0.25f, isotropic, w, pred, nullptr, 0.0f, nullptr, nullptr);
  } else {
    hipLaunchKernelGGL(solver::tv_prox_kernel, dim3(solver::elem_grid(n)), dim3(256), 0,
                       as_stream(stream), v, out, t, tau, 1.0f / static_cast<float>(2 * ndim),
                       isotropic, w, pred);
  }
  SCXR_LAUNCHED("tv prox w");
  return SCXR_OK;
}

int64_t scxr_tv_prox_vjp_ws(int32_t ndim, const int32_t* shape, int32_t circular) {
  solver::TvDims t;
  if (tv_dims(ndim, shape, circular, t) != SCXR_OK) return -1;
  return 2LL * ndim * t.P[0] * t.P[1] * t.P[2];
}

int scxr_tv_prox_vjp(int32_t ndim, const int32_t* shape, const float* v, const float* w,
                     const float* gout, float tau, int32_t isotropic, int32_t circular,
                     float* gv, float* gw, float* ws, void* stream) {
  SCXR_REQUIRE(v && gout && gv && ws, SCXR_ERR_ARG, "bad tv_prox_vjp arguments");
  SCXR_REQUIRE(!gw || w, SCXR_ERR_ARG, "tv_prox_vjp: gw needs the weight map w");
  solver::TvDims t;
  const int rc = tv_dims(ndim, shape, circular, t);
  if (rc) return rc;
  const int64_t n = static_cast<int64_t>(t.d[0]) * t.d[1] * t.d[2];
  const int64_t np = static_cast<int64_t>(t.P[0]) * t.P[1] * t.P[2];
  float* HB = ws;
  float* LB = ws + ndim * np;
  const float invK = 1.0f / static_cast<float>(2 * ndim);
  hipLaunchKernelGGL(solver::tv_vjp_coef_kernel, dim3(solver::elem_grid(np)), dim3(256), 0,
                     as_stream(stream), v, gout, w, t, tau, invK, isotropic, HB, LB, gw);
  SCXR_LAUNCHED("tv prox vjp coef");
  hipLaunchKernelGGL(solver::tv_vjp_v_kernel, dim3(solver::elem_grid(n)), dim3(256), 0,
                     as_stream(stream), HB, LB, t, gv);
  SCXR_LAUNCHED("tv prox vjp v");
  return SCXR_OK;
}

int scxr_tv_norm(int32_t ndim, const int32_t* shape, const float* x, int32_t isotropic,
                 int32_t circular, double* out, double* ws, const int32_t* pred, void* stream) {
  SCXR_REQUIRE(x && out && ws, SCXR_ERR_ARG, "bad tv_norm arguments");
  solver::TvDims t;
  const int rc = tv_dims(ndim, shape, circular, t);
  if (rc) return rc;
  const int64_t n = static_cast<int64_t>(t.d[0]) * t.d[1] * t.d[2];
  const unsigned g = solver::elem_grid(n);
  hipLaunchKernelGGL(solver::tv_norm_kernel, dim3(g), dim3(256), 0, as_stream(stream), x, t,
                     isotropic, ws, pred);
  SCXR_LAUNCHED("tv norm");
  return vec::sum_partials(static_cast<int>(g), ws, 1.0, vec::POST_NONE, out, pred,
                           as_stream(stream));
}

}  // extern "C"
