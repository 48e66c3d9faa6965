// 2D parallel-beam X-ray projector for gfx950: forward (pixel-driven,
// LDS-privatised accumulation, no global atomics) and adjoint (gather,
// sinogram windows staged in LDS).  Both evaluate the pixel-boxcar weights
// of scico/linop/xray/_xray2d.py:309-351 bit-exactly:
//   Px     = (Pxmin + P0 * i) + P1 * j          (no FMA contraction)
//   ind    = floor(Px)
//   weight = min(1 - (Px - ind), width) / width (IEEE float32 quotient)
// The per-view float32 geometry (Pxmin, P0, P1, width) is computed on the
// host at plan creation in the reference's operation order
// (_xray2d.py:326-347).
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"
#include "xray2d_impl.h"

using namespace scxr;

struct scxr_xr2d_plan {
  xr2d::Plan p;
};

namespace scxr {
namespace xr2d {

// ---------------------------------------------------------------- host --

void view_geometry(int A, const float* angles, const double* x0, const double* dx,
                   double y0, std::vector<float4>& geom, std::vector<double>& rw) {
  geom.resize(A);
  rw.resize(A);
  const float x00 = static_cast<float>(x0[0]), x01 = static_cast<float>(x0[1]);
  const float dx0 = static_cast<float>(dx[0]), dx1 = static_cast<float>(dx[1]);
  const float y0f = static_cast<float>(y0);
  for (int a = 0; a < A; ++a) {
    const double th = static_cast<double>(angles[a]);
    const float u0 = static_cast<float>(std::cos(th));  // _xray2d.py:326
    const float u1 = static_cast<float>(std::sin(th));
    const float t0 = x00 * u0;
    const float t1 = x01 * u1;
    const float Px0 = (t0 + t1) - y0f;  // :327
    const float P0 = dx0 * u0;          // :328
    const float P1 = dx1 * u1;
    const float c1 = Px0 + P0;
    const float c2 = Px0 + P1;
    const float c3 = c1 + P1;
    const float Pxmin = std::min(std::min(std::min(Px0, c1), c2), c3);  // :329
    const float d1 = std::fabs(P0 + P1);  // :342
    const float d2 = std::fabs(P0 - P1);  // :343
    const float w = std::max(d1, d2);
    const float f = std::min(d1, d2);
    const float width = (w + f) / 2.0f;  // :347
    geom[a] = make_float4(Pxmin, P0, P1, width);
    rw[a] = 1.0 / static_cast<double>(width);
  }
}

int create(Plan& p, int M, int N, int A, int D, const double* x0, const double* dx,
           double y0, const float* angles) {
  p.M = M;
  p.N = N;
  p.A = A;
  p.D = D;
  std::vector<float4> geom;
  std::vector<double> rw;
  view_geometry(A, angles, x0, dx, y0, geom, rw);
  p.host_geom = geom;
  for (int a = 0; a < A; ++a)
    SCXR_REQUIRE(std::isfinite(geom[a].x) && std::isfinite(geom[a].y) &&
                     std::isfinite(geom[a].z) && geom[a].w > 0.0f && std::isfinite(geom[a].w),
                 SCXR_ERR_GEOM, "view %d has a degenerate projected pixel (width %g)", a,
                 static_cast<double>(geom[a].w));

  // Forward: views whose detector coordinate varies fastest along image
  // columns (|P1| >= |P0|) stream the image row-major; the others stream
  // its transpose, so that the 64 lanes of a wave always spread over many
  // detector bins (few LDS atomic collisions).
  std::vector<int> rows, cols;
  for (int a = 0; a < A; ++a)
    (std::fabs(geom[a].z) >= std::fabs(geom[a].y) ? rows : cols).push_back(a);
  p.any_colfast = !cols.empty();
  // Views per workgroup: enough workgroups to fill 256 CUs at least twice.
  p.vpg = A >= 2048 ? 4 : (A >= 1024 ? 2 : 1);
  while (p.vpg > 1 && fwd_lds_bytes(p.vpg, D) > FWD_LDS_MAX) p.vpg /= 2;
  SCXR_REQUIRE(fwd_lds_bytes(1, D) <= FWD_LDS_MAX, SCXR_ERR_UNSUPPORTED,
               "det_count %d too large for the LDS-resident forward projector", D);
  std::vector<int> order;
  std::vector<int4> groups;
  for (int cls = 0; cls < 2; ++cls) {
    const std::vector<int>& list = cls ? cols : rows;
    for (size_t s = 0; s < list.size(); s += p.vpg) {
      const int cnt = static_cast<int>(std::min<size_t>(p.vpg, list.size() - s));
      groups.push_back(make_int4(static_cast<int>(order.size()), cnt, cls, 0));
      for (int k = 0; k < cnt; ++k) order.push_back(list[s + k]);
    }
  }
  p.n_groups = static_cast<int>(groups.size());

  // Adjoint: detector window per (tile, view) = all bins an ADJ_TI x ADJ_TJ
  // pixel tile can reach, plus margin for rounding.
  int wmax = 0;
  for (int a = 0; a < A; ++a) {
    const double ext = (ADJ_TI - 1) * std::fabs(static_cast<double>(geom[a].y)) +
                       (ADJ_TJ - 1) * std::fabs(static_cast<double>(geom[a].z));
    wmax = std::max(wmax, static_cast<int>(std::ceil(ext)) + 4);
  }
  p.adj_w = (wmax + 3) & ~3;
  SCXR_REQUIRE(p.adj_w <= ADJ_WMAX, SCXR_ERR_GEOM,
               "projected pixel footprint too large for the adjoint tile (%d bins > %d); "
               "pixel size dx is too large relative to the detector spacing",
               p.adj_w, ADJ_WMAX);

  size_t bytes_geom = sizeof(float4) * A, bytes_rw = sizeof(double) * A;
  size_t bytes_groups = sizeof(int4) * groups.size(), bytes_order = sizeof(int) * A;
  size_t total = bytes_geom + bytes_rw + bytes_groups + bytes_order + 256 * 4;
  if (hipMalloc(&p.dev_block, total) != hipSuccess)
    return fail(SCXR_ERR_NOMEM, "hipMalloc of %zu bytes for the plan tables failed", total);
  char* base = static_cast<char*>(p.dev_block);
  auto carve = [&](size_t bytes) {
    char* r = base;
    base += (bytes + 255) & ~static_cast<size_t>(255);
    return r;
  };
  p.geom = reinterpret_cast<float4*>(carve(bytes_geom));
  p.rw = reinterpret_cast<double*>(carve(bytes_rw));
  p.groups = reinterpret_cast<int4*>(carve(bytes_groups));
  p.order = reinterpret_cast<int*>(carve(bytes_order));
  SCXR_HIP(hipMemcpy(p.geom, geom.data(), bytes_geom, hipMemcpyHostToDevice));
  SCXR_HIP(hipMemcpy(p.rw, rw.data(), bytes_rw, hipMemcpyHostToDevice));
  SCXR_HIP(hipMemcpy(p.groups, groups.data(), bytes_groups, hipMemcpyHostToDevice));
  SCXR_HIP(hipMemcpy(p.order, order.data(), bytes_order, hipMemcpyHostToDevice));
  if (p.any_colfast) {
    if (hipMalloc(&p.work, sizeof(float) * static_cast<size_t>(M) * N) != hipSuccess)
      return fail(SCXR_ERR_NOMEM, "hipMalloc of the transpose workspace failed");
  }
  return SCXR_OK;
}

void destroy(Plan& p) {
  if (p.dev_block) (void)hipFree(p.dev_block);
  if (p.work) (void)hipFree(p.work);
  p.dev_block = nullptr;
  p.work = nullptr;
}

// -------------------------------------------------------------- device --

// Shared weight evaluation (bit-exact restatement of _xray2d.py:331-349).
struct Weight {
  float fl;  // floor(Px) as float
  float w;   // weight of bin fl; bin fl + 1 gets 1 - w
};

__device__ __forceinline__ Weight weight_at(float Pxmin, float P0, float P1, float width,
                                            double rw, float fi, float fj) {
  const float Px = __fadd_rn(__fadd_rn(Pxmin, __fmul_rn(P0, fi)), __fmul_rn(P1, fj));
  const float fl = floorf(Px);
  const float d = fminf(__fsub_rn(1.0f, __fsub_rn(Px, fl)), width);
  return Weight{fl, div_rn(d, rw)};
}

// Transpose (M, N) -> (N, M) through a padded LDS tile.
__global__ __launch_bounds__(256) void transpose_kernel(const float* __restrict__ in,
                                                        float* __restrict__ out, int M, int N,
                                                        const int32_t* pred) {
  if (pred_off(pred)) return;
  __shared__ float tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  const int i0 = blockIdx.y * 32, j0 = blockIdx.x * 32;
  for (int r = ty; r < 32; r += 8) {
    const int i = i0 + r, j = j0 + tx;
    if (i < M && j < N) tile[r][tx] = in[static_cast<size_t>(i) * N + j];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int j = j0 + r, i = i0 + tx;
    if (i < M && j < N) out[static_cast<size_t>(j) * M + i] = tile[tx][r];
  }
}

// Forward projection.  One workgroup owns up to V whole sinogram rows
// (views) and streams the entire image; each of its 4 waves accumulates
// into a private LDS copy of every owned row with LDS float atomics, and
// the copies are summed in a fixed order at the end, so the result is
// deterministic and needs no global atomics.  Row slot k holds detector
// bin k - 2; bins outside [0, D) land in the two guard slots on each side.
template <int V, bool COLFAST>
__device__ __forceinline__ void fwd_body(const float* __restrict__ src, float* acc, int L,
                                         int cnt, const float* Pxmin, const float* P0,
                                         const float* P1, const float* wd, const double* rw,
                                         int M, int N, int D) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int Lf = COLFAST ? M : N;  // fast (lane) axis length
  const int Ls = COLFAST ? N : M;
  const float Dmax = static_cast<float>(D);
  for (int s = wave; s < Ls; s += 4) {
    const float fs = static_cast<float>(s);
    const float* srow = src + static_cast<size_t>(s) * Lf;
    for (int f = lane; f < Lf; f += 64) {
      const float val = srow[f];
      const float ff = static_cast<float>(f);
      const float fi = COLFAST ? ff : fs;
      const float fj = COLFAST ? fs : ff;
#pragma unroll
      for (int v = 0; v < V; ++v) {
        if (v < cnt) {
          const Weight g = weight_at(Pxmin[v], P0[v], P1[v], wd[v], rw[v], fi, fj);
          const int k = static_cast<int>(fminf(fmaxf(g.fl, -2.0f), Dmax)) + 2;
          float* row = acc + (v * 4 + wave) * L;
          atomicAdd(row + k, __fmul_rn(val, g.w));
          atomicAdd(row + k + 1, __fmul_rn(val, __fsub_rn(1.0f, g.w)));
        }
      }
    }
  }
}

template <int V>
__global__ __launch_bounds__(256) void fwd_kernel(const float* __restrict__ im,
                                                  const float* __restrict__ imT,
                                                  float* __restrict__ sino,
                                                  const float4* __restrict__ geom,
                                                  const double* __restrict__ rwt,
                                                  const int4* __restrict__ groups,
                                                  const int* __restrict__ order, int M, int N,
                                                  int D, const int32_t* pred) {
  if (pred_off(pred)) return;
  extern __shared__ float acc[];  // [V][4 waves][L]
  const int L = D + 4;
  const int4 grp = groups[blockIdx.x];
  const int cnt = grp.y;
  float Pxmin[V], P0[V], P1[V], wd[V];
  double rw[V];
  int view[V];
#pragma unroll
  for (int v = 0; v < V; ++v) {
    const int a = order[grp.x + (v < cnt ? v : 0)];
    view[v] = a;
    const float4 g = geom[a];
    Pxmin[v] = g.x;
    P0[v] = g.y;
    P1[v] = g.z;
    wd[v] = g.w;
    rw[v] = rwt[a];
  }
  for (int e = threadIdx.x; e < V * 4 * L; e += 256) acc[e] = 0.0f;
  __syncthreads();
  if (grp.z)
    fwd_body<V, true>(imT, acc, L, cnt, Pxmin, P0, P1, wd, rw, M, N, D);
  else
    fwd_body<V, false>(im, acc, L, cnt, Pxmin, P0, P1, wd, rw, M, N, D);
  __syncthreads();
#pragma unroll
  for (int v = 0; v < V; ++v) {
    if (v < cnt) {
      const float* r = acc + v * 4 * L + 2;
      float* dst = sino + static_cast<size_t>(view[v]) * D;
      for (int b = threadIdx.x; b < D; b += 256)
        dst[b] = __fadd_rn(__fadd_rn(__fadd_rn(r[b], r[L + b]), r[2 * L + b]), r[3 * L + b]);
    }
  }
}

// Adjoint (back projection).  One workgroup per ADJ_TI x ADJ_TJ pixel
// tile; views are processed in chunks of ADJ_VC: for each view the
// detector window reachable from the tile is staged in LDS (zero outside
// [0, D)), then every thread gathers its pixel's two bins.  The two sums
// of _xray2d.py:299-304 are kept separately and added at the end.
__global__ __launch_bounds__(256) void adj_kernel(const float* __restrict__ sino,
                                                  float* __restrict__ out,
                                                  const float4* __restrict__ geom,
                                                  const double* __restrict__ rwt, int M, int N,
                                                  int A, int D, int W, const int32_t* pred) {
  if (pred_off(pred)) return;
  extern __shared__ float smem[];
  float* win = smem;                                    // [ADJ_VC][W]
  float* lo_s = smem + ADJ_VC * W;                      // [ADJ_VC]
  const int tx = threadIdx.x % ADJ_TJ, ty = threadIdx.x / ADJ_TJ;
  const int i0 = blockIdx.y * ADJ_TI, j0 = blockIdx.x * ADJ_TJ;
  const int i = i0 + ty, j = j0 + tx;
  const bool inside = (i < M) && (j < N);
  const float fi = static_cast<float>(i), fj = static_cast<float>(j);
  const float ci0 = static_cast<float>(i0), ci1 = static_cast<float>(min(i0 + ADJ_TI, M) - 1);
  const float cj0 = static_cast<float>(j0), cj1 = static_cast<float>(min(j0 + ADJ_TJ, N) - 1);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const float kmax = static_cast<float>(W - 2);
  float acc0 = 0.0f, acc1 = 0.0f;
  for (int a0 = 0; a0 < A; a0 += ADJ_VC) {
    const int nv = min(ADJ_VC, A - a0);
    __syncthreads();
    for (int v = wave; v < nv; v += 4) {
      const float4 g = geom[a0 + v];
      const float r0 = __fadd_rn(g.x, __fmul_rn(g.y, ci0));
      const float r1 = __fadd_rn(g.x, __fmul_rn(g.y, ci1));
      const float c0 = __fmul_rn(g.z, cj0), c1 = __fmul_rn(g.z, cj1);
      const float pmin = fminf(fminf(__fadd_rn(r0, c0), __fadd_rn(r0, c1)),
                               fminf(__fadd_rn(r1, c0), __fadd_rn(r1, c1)));
      const float lof = floorf(pmin);
      const int lo = static_cast<int>(lof);
      const float* srow = sino + static_cast<size_t>(a0 + v) * D;
      for (int k = lane; k < W; k += 64) {
        const int b = lo + k;
        win[v * W + k] = (b >= 0 && b < D) ? srow[b] : 0.0f;
      }
      if (lane == 0) lo_s[v] = lof;
    }
    __syncthreads();
    if (inside) {
      for (int v = 0; v < nv; ++v) {
        const float4 g = geom[a0 + v];
        const Weight wt = weight_at(g.x, g.y, g.z, g.w, rwt[a0 + v], fi, fj);
        const int k = static_cast<int>(fminf(fmaxf(__fsub_rn(wt.fl, lo_s[v]), 0.0f), kmax));
        const float* wrow = win + v * W + k;
        acc0 = fmaf(wrow[0], wt.w, acc0);
        acc1 = fmaf(wrow[1], __fsub_rn(1.0f, wt.w), acc1);
      }
    }
  }
  if (inside) out[static_cast<size_t>(i) * N + j] = __fadd_rn(acc0, acc1);
}

// Debug dump of (ind, weight) for every (view, pixel).
__global__ __launch_bounds__(256) void weights_kernel(int32_t* __restrict__ inds,
                                                      float* __restrict__ wts,
                                                      const float4* __restrict__ geom,
                                                      const double* __restrict__ rwt, int M,
                                                      int N, int A) {
  const size_t n = static_cast<size_t>(M) * N;
  const size_t e = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
  if (e >= n * A) return;
  const int a = static_cast<int>(e / n);
  const size_t p = e - static_cast<size_t>(a) * n;
  const int i = static_cast<int>(p / N), j = static_cast<int>(p - static_cast<size_t>(i) * N);
  const float4 g = geom[a];
  const Weight wt =
      weight_at(g.x, g.y, g.z, g.w, rwt[a], static_cast<float>(i), static_cast<float>(j));
  inds[e] = static_cast<int32_t>(wt.fl);
  wts[e] = wt.w;
}

int forward(const Plan& p, const float* im, float* sino, const int32_t* pred, hipStream_t s) {
  if (p.any_colfast) {
    dim3 g(ceil_div(p.N, 32), ceil_div(p.M, 32));
    hipLaunchKernelGGL(transpose_kernel, g, dim3(256), 0, s, im, p.work, p.M, p.N, pred);
    SCXR_LAUNCHED("xr2d transpose");
  }
  const size_t lds = fwd_lds_bytes(p.vpg, p.D);
  switch (p.vpg) {
    case 1:
      hipLaunchKernelGGL(fwd_kernel<1>, dim3(p.n_groups), dim3(256), lds, s, im, p.work, sino,
                         p.geom, p.rw, p.groups, p.order, p.M, p.N, p.D, pred);
      break;
    case 2:
      hipLaunchKernelGGL(fwd_kernel<2>, dim3(p.n_groups), dim3(256), lds, s, im, p.work, sino,
                         p.geom, p.rw, p.groups, p.order, p.M, p.N, p.D, pred);
      break;
    default:
      hipLaunchKernelGGL(fwd_kernel<4>, dim3(p.n_groups), dim3(256), lds, s, im, p.work, sino,
                         p.geom, p.rw, p.groups, p.order, p.M, p.N, p.D, pred);
      break;
  }
  SCXR_LAUNCHED("xr2d forward");
  return SCXR_OK;
}

int adjoint(const Plan& p, const float* sino, float* im, const int32_t* pred, hipStream_t s) {
  dim3 g(ceil_div(p.N, ADJ_TJ), ceil_div(p.M, ADJ_TI));
  const size_t lds = sizeof(float) * (static_cast<size_t>(ADJ_VC) * p.adj_w + ADJ_VC);
  hipLaunchKernelGGL(adj_kernel, g, dim3(256), lds, s, sino, im, p.geom, p.rw, p.M, p.N, p.A,
                     p.D, p.adj_w, pred);
  SCXR_LAUNCHED("xr2d adjoint");
  return SCXR_OK;
}

}  // namespace xr2d
}  // namespace scxr

// ------------------------------------------------------------- C ABI --

extern "C" {

int scxr_xr2d_plan_create(int32_t M, int32_t N, int32_t A, int32_t D, const double* x0,
                          const double* dx, double y0, const float* angles,
                          scxr_xr2d_plan** out) {
  SCXR_REQUIRE(out && x0 && dx && angles, SCXR_ERR_ARG, "null argument to plan_create");
  SCXR_REQUIRE(M > 0 && N > 0 && A > 0 && D > 0, SCXR_ERR_SHAPE,
               "non-positive size (M=%d N=%d A=%d D=%d)", M, N, A, D);
  SCXR_REQUIRE(M < (1 << 24) && N < (1 << 24), SCXR_ERR_SHAPE, "image too large");
  auto* plan = new scxr_xr2d_plan();
  int rc = xr2d::create(plan->p, M, N, A, D, x0, dx, y0, angles);
  if (rc != SCXR_OK) {
    xr2d::destroy(plan->p);
    delete plan;
    return rc;
  }
  *out = plan;
  return SCXR_OK;
}

int scxr_xr2d_plan_destroy(scxr_xr2d_plan* plan) {
  if (!plan) return SCXR_OK;
  xr2d::destroy(plan->p);
  delete plan;
  return SCXR_OK;
}

int scxr_xr2d_view_table(const scxr_xr2d_plan* plan, float* out) {
  SCXR_REQUIRE(plan && out, SCXR_ERR_ARG, "null argument to view_table");
  for (int a = 0; a < plan->p.A; ++a) {
    const float4 g = plan->p.host_geom[a];
    out[4 * a + 0] = g.x;
    out[4 * a + 1] = g.y;
    out[4 * a + 2] = g.z;
    out[4 * a + 3] = g.w;
  }
  return SCXR_OK;
}

int scxr_xr2d_forward(const scxr_xr2d_plan* plan, const float* im, float* sino,
                      const int32_t* pred, void* stream) {
  SCXR_REQUIRE(plan && im && sino, SCXR_ERR_ARG, "null argument to xr2d_forward");
  return xr2d::forward(plan->p, im, sino, pred, as_stream(stream));
}

int scxr_xr2d_adjoint(const scxr_xr2d_plan* plan, const float* sino, float* im,
                      const int32_t* pred, void* stream) {
  SCXR_REQUIRE(plan && im && sino, SCXR_ERR_ARG, "null argument to xr2d_adjoint");
  return xr2d::adjoint(plan->p, sino, im, pred, as_stream(stream));
}

int scxr_xr2d_weights(const scxr_xr2d_plan* plan, int32_t* inds, float* weights, void* stream) {
  SCXR_REQUIRE(plan && inds && weights, SCXR_ERR_ARG, "null argument to xr2d_weights");
  const auto& p = plan->p;
  const size_t total = static_cast<size_t>(p.M) * p.N * p.A;
  SCXR_REQUIRE(total < (static_cast<size_t>(1) << 40), SCXR_ERR_SHAPE, "too many entries");
  hipLaunchKernelGGL(xr2d::weights_kernel, dim3(ceil_div(static_cast<long long>(total), 256)),
                     dim3(256), 0, as_stream(stream), inds, weights, p.geom, p.rw, p.M, p.N,
                     p.A);
  SCXR_LAUNCHED("xr2d weights");
  return SCXR_OK;
}

}  // extern "C"
