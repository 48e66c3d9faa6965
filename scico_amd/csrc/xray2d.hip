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
#include "fd_impl.h"
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

// LDS bank model of one forward phase: lane l updates bins b and b + 1 with
// b = floor(p + d pos(l)), pos(l) = (l k) mod 64.  An LDS access serves 32
// lanes at a time over 32 banks, one word per bank per cycle, so each
// half-wave and word costs the largest number of its lanes that share a
// bank; the sum over both halves and both words is returned, averaged over
// fixed offsets p.  The cost depends only on |d| and k; the cheapest odd k
// and its cost are tabulated for |d| in steps of 1/256.
struct LaneChoice {
  int k;     // lane multiplier
  int cost;  // modelled bank cost of a phase (96: two lanes per bank)
};
static LaneChoice lane_choice(double d) {
  thread_local std::vector<int> memo;
  d = std::fabs(d);
  const int key = static_cast<int>(std::lround(d * 256.0));
  if (key < 0 || key >= 64 * 256) return {1, 96};
  if (memo.empty()) memo.assign(64 * 256, 0);
  if (memo[key]) return {memo[key] & 255, memo[key] >> 8};
  const double dq = key / 256.0;
  int best_k = 1;
  long best = -1;
  for (int k = 1; k < 64; k += 2) {
    long cost = 0;
    for (int j = 0; j < 12; ++j) {
      const double p0 = 0.37 + 5.31 * j;
      for (int word = 0; word < 2; ++word)
        for (int half = 0; half < 2; ++half) {
          int cnt[32] = {0}, mx = 0;
          for (int l = 32 * half; l < 32 * half + 32; ++l) {
            const long b = static_cast<long>(std::floor(p0 + dq * ((l * k) & 63))) + word;
            mx = std::max(mx, ++cnt[b & 31]);
          }
          cost += mx;
        }
    }
    if (best < 0 || cost < best) {
      best = cost;
      best_k = k;
    }
  }
  memo[key] = best_k | static_cast<int>(best << 8);
  return {best_k, static_cast<int>(best)};
}

// Workgroup placement of the forward groups.  The dispatcher hands a
// launch's workgroups out in index order, one per CU per round while every
// CU has room (block index i lands on CU slot i mod cus of round
// i / cus), and a forward group's time varies with its view (1.4x between
// a view whose lanes fall on distinct LDS banks and one whose bins alias).
// Each round, the groups still to place, most expensive first, go to the
// CU slots with the least estimated load so far (ties: higher slots
// first, since a last, partial round fills only the lowest slots).  Only
// the order in which groups are processed changes, not their arithmetic.
static std::vector<int> place_groups(const std::vector<float>& est, int cus) {
  const int G = static_cast<int>(est.size());
  std::vector<int> byc(G), block(G);
  for (int i = 0; i < G; ++i) byc[i] = i;
  std::stable_sort(byc.begin(), byc.end(), [&](int a, int b) { return est[a] > est[b]; });
  std::vector<double> load(cus, 0.0);
  std::vector<int> slot;
  for (int r = 0, next = 0; next < G; ++r) {
    const int n = std::min(cus, G - next);
    slot.resize(n);
    for (int i = 0; i < n; ++i) slot[i] = i;
    std::stable_sort(slot.begin(), slot.end(), [&](int a, int b) {
      return load[a] < load[b] || (load[a] == load[b] && a > b);
    });
    for (int j = 0; j < n; ++j) {
      load[slot[j]] += est[byc[next + j]];
      block[r * cus + slot[j]] = byc[next + j];
    }
    next += n;
  }
  return block;
}

int build_tables(HostTables& h, int M, int N, int A, int D, const double* x0, const double* dx,
                 double y0, const float* angles, int vpg, int ppt, int fwd_ch,
                 bool allow_generic) {
  view_geometry(A, angles, x0, dx, y0, h.geom, h.rw);
  const std::vector<float4>& geom = h.geom;
  for (int a = 0; a < A; ++a)
    SCXR_REQUIRE(std::isfinite(geom[a].x) && std::isfinite(geom[a].y) &&
                     std::isfinite(geom[a].z) && geom[a].w > 0.0f && std::isfinite(geom[a].w),
                 SCXR_ERR_GEOM, "view %d has a degenerate projected pixel (width %g)", a,
                 static_cast<double>(geom[a].w));

  // Forward: views whose detector coordinate varies fastest along image
  // columns (|P1| >= |P0|) stream the image row-major; the others stream
  // its transpose, so that the 64 lanes of a wave always spread over many
  // detector bins.
  std::vector<int> rows, cols;
  for (int a = 0; a < A; ++a)
    (std::fabs(geom[a].z) >= std::fabs(geom[a].y) ? rows : cols).push_back(a);
  h.any_colfast = !cols.empty();
  if (fwd_ch < 0) fwd_ch = fwd_chains(vpg, D);
  h.fwd_generic = false;
  if (fwd_ch == 0) {
    SCXR_REQUIRE(allow_generic, SCXR_ERR_UNSUPPORTED,
                 "det_count %d too large for the LDS-resident forward projector", D);
    h.fwd_generic = true;
  }
  h.order.clear();
  h.groups.clear();
  h.gcost.clear();
  for (int cls = 0; cls < 2 && !h.fwd_generic; ++cls) {
    const std::vector<int>& list = cls ? cols : rows;
    for (size_t s = 0; s < list.size(); s += vpg) {
      const int cnt = static_cast<int>(std::min<size_t>(vpg, list.size() - s));
      // Lane stride along the fast axis: S * |P_fast| >= 2 (+ margin for
      // the float32 rounding of Px) keeps the lanes' footprints disjoint.
      int S = 1;
      for (int k = 0; k < cnt; ++k) {
        const float4 g = geom[list[s + k]];
        const double pf = std::max(std::fabs(static_cast<double>(cls ? g.y : g.z)), 1e-30);
        S = std::max(S, static_cast<int>(std::ceil(2.0625 / pf)));
      }
      if (S > FWD_SMAX) {
        SCXR_REQUIRE(allow_generic, SCXR_ERR_GEOM,
                     "projected pixel spacing too small for the forward projector "
                     "(lane stride %d > %d)", S, FWD_SMAX);
        h.fwd_generic = true;
        break;
      }
      // A larger stride keeps the footprints disjoint too; among the
      // compiled strides take the one with the fewest phases per line,
      // ceil(Lf / 64 S) * S (e.g. 4 rather than 3 for 512-pixel lines: 8
      // phases instead of 9).
      if (S <= 5) {
        const int Lf = cls ? M : N;
        auto phases = [&](int s2) { return (Lf + 64 * s2 - 1) / (64 * s2) * s2; };
        int best = S;
        for (int s2 = S + 1; s2 <= 5; ++s2)
          if (phases(s2) < phases(best)) best = s2;
        S = best;
      }
      // Lane permutation (position (lane k) mod 64 for an odd k) that the
      // bank model prefers for the group's first view: the phase's set of
      // pixels, hence the disjointness of the footprints, is unchanged.
      const float4 g0 = geom[list[s]];
      const LaneChoice lc = lane_choice(S * static_cast<double>(cls ? g0.y : g0.z));
      const int kmul = lc.k;
      h.groups.push_back(make_int4(static_cast<int>(h.order.size()), cnt, cls | (kmul << 1), S));
      // Relative time of the group: phases per line times lines times a
      // per-phase cost, fitted to 512^2 plans of one repeated view on
      // MI355X (95-137 us over the angles; the phase time stops falling
      // below a bank cost of 92, where the VALU work sets it).
      {
        const int Lf = cls ? M : N, Ls = cls ? N : M;
        const double ph = static_cast<double>((Lf + 64 * S - 1) / (64 * S) * S);
        h.gcost.push_back(static_cast<float>(cnt * Ls * ph * (33.0 + std::max(lc.cost, 92))));
      }
      for (int k = 0; k < cnt; ++k) h.order.push_back(list[s + k]);
    }
  }

  // Px >= 1 for every pixel of every view (bounds of Px over the image in
  // float64, with a margin far above the float32 rounding of Px): then
  // Px - floor(Px) is the exact fraction and the adjoint may take it, and
  // the weight, from the fraction instruction (see weight_frac).
  h.px_ge1 = true;
  for (int a = 0; a < A; ++a) {
    const double lo = static_cast<double>(geom[a].x) + std::min<double>(geom[a].y, 0.0) * (M - 1) +
                      std::min<double>(geom[a].z, 0.0) * (N - 1);
    h.px_ge1 = h.px_ge1 && lo >= 1.01;
  }

  // Adjoint: window origin offset per view (see adj_tile): a lower bound
  // of floor(Px) - 1 over any tile; the window spans the projected tile
  // extent plus the lower margin, the rounding of Px and bin ind + 1.  Two
  // pixels per thread (16 x 32 tiles) when that window fits, else one.
  auto windows = [&](int pp, std::vector<float>& clo, int& wmax) {
    const int TJ = ADJ_TJ * pp;
    clo.assign(A, 0.0f);
    wmax = 0;
    for (int a = 0; a < A; ++a) {
      const double P0 = geom[a].y, P1 = geom[a].z;
      clo[a] = static_cast<float>(static_cast<double>(geom[a].x) +
                                  std::min(P0, 0.0) * (ADJ_TI - 1) +
                                  std::min(P1, 0.0) * (TJ - 1) - 1.0);
      const double ext = (ADJ_TI - 1) * std::fabs(P0) + (TJ - 1) * std::fabs(P1);
      wmax = std::max(wmax, static_cast<int>(std::ceil(ext)) + 5);
    }
  };
  if (ppt == 0) {
    windows(2, h.clo, h.adj_w);
    h.adj_ppt = 2;
    if (h.adj_w > ADJ_WFAST) {
      windows(1, h.clo, h.adj_w);
      h.adj_ppt = 1;
    }
  } else {
    windows(ppt, h.clo, h.adj_w);
    h.adj_ppt = ppt;
  }
  h.adj_generic = false;
  if (h.adj_w > ADJ_WFAST) {
    SCXR_REQUIRE(allow_generic, SCXR_ERR_GEOM,
                 "projected pixel footprint too large for the adjoint tile (%d bins > %d); "
                 "pixel size dx is too large relative to the detector spacing",
                 h.adj_w, ADJ_WFAST);
    h.adj_generic = true;
  }
  if (h.fwd_generic) {
    h.groups.clear();
    h.gcost.clear();
    h.order.assign(A, 0);
    for (int a = 0; a < A; ++a) h.order[a] = a;
  }
  return SCXR_OK;
}

static hipError_t fwd_allow_lds(int bytes);

int create(Plan& p, int M, int N, int A, int D, const double* x0, const double* dx,
           double y0, const float* angles) {
  p.M = M;
  p.N = N;
  p.A = A;
  p.D = D;
  // Views per workgroup: enough workgroups to fill 256 CUs at least twice.
  p.vpg = A >= 2048 ? 4 : (A >= 1024 ? 2 : 1);
  while (p.vpg > 1 && fwd_chains(p.vpg, D) == 0) p.vpg /= 2;
  size_t lds_cap = FWD_LDS_MAX;
  {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
        v > 0)
      lds_cap = std::max(lds_cap, static_cast<size_t>(v));
  }
  p.fwd_ch = fwd_chains_ext(p.vpg, D, lds_cap);
  if (p.fwd_ch > 0 && fwd_lds_bytes(p.vpg, D, p.fwd_ch) > FWD_LDS_MAX) {
    // rows beyond 64 KiB: raise the kernel's dynamic LDS limit, or take the
    // general-geometry forward if the runtime refuses
    const int bytes = static_cast<int>(fwd_lds_bytes(p.vpg, D, p.fwd_ch));
    if (fwd_allow_lds(bytes) != hipSuccess) {
      (void)hipGetLastError();
      p.fwd_ch = 0;
    }
  }
  HostTables h;
  const int rc = build_tables(h, M, N, A, D, x0, dx, y0, angles, p.vpg, 0, p.fwd_ch, true);
  if (rc) return rc;
  p.fwd_generic = h.fwd_generic;
  p.adj_generic = h.adj_generic;
  p.adj_ppt = h.adj_ppt;
  p.adj_fract = h.px_ge1;
  p.host_geom = h.geom;
  p.any_colfast = h.any_colfast;
  p.n_groups = static_cast<int>(h.groups.size());
  p.adj_w = h.adj_w;
  const std::vector<float4>& geom = h.geom;
  const std::vector<double>& rw = h.rw;
  const std::vector<int4>& groups = h.groups;
  const std::vector<int>& order = h.order;
  const std::vector<float>& clo = h.clo;

  size_t bytes_geom = sizeof(float4) * A, bytes_rw = sizeof(double) * A;
  size_t bytes_groups = sizeof(int4) * groups.size(), bytes_order = sizeof(int) * A;
  size_t bytes_clo = sizeof(float) * A;
  // block index -> forward group, balanced over the CUs
  std::vector<int> gblock;
  if (!groups.empty()) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
    gblock = place_groups(h.gcost, cus);
  }
  size_t bytes_gblock = sizeof(int) * gblock.size();
  size_t total =
      bytes_geom + bytes_rw + bytes_groups + bytes_order + bytes_clo + bytes_gblock + 256 * 6;
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
  p.clo = reinterpret_cast<float*>(carve(bytes_clo));
  p.gblock = reinterpret_cast<int*>(carve(bytes_gblock));
  if (bytes_gblock)
    SCXR_HIP(hipMemcpy(p.gblock, gblock.data(), bytes_gblock, hipMemcpyHostToDevice));
  SCXR_HIP(hipMemcpy(p.clo, clo.data(), bytes_clo, hipMemcpyHostToDevice));
  SCXR_HIP(hipMemcpy(p.geom, geom.data(), bytes_geom, hipMemcpyHostToDevice));
  SCXR_HIP(hipMemcpy(p.rw, rw.data(), bytes_rw, hipMemcpyHostToDevice));
  SCXR_HIP(hipMemcpy(p.groups, groups.data(), bytes_groups, hipMemcpyHostToDevice));
  SCXR_HIP(hipMemcpy(p.order, order.data(), bytes_order, hipMemcpyHostToDevice));
  if (p.any_colfast) {
    if (hipMalloc(&p.work, sizeof(float) * static_cast<size_t>(M) * N) != hipSuccess)
      return fail(SCXR_ERR_NOMEM, "hipMalloc of the transpose workspace failed");
  }
  p.fwd_sp = 1;
  (void)set_split(p, auto_split(p));  // no workspace: one workgroup per group
  return SCXR_OK;
}

// A forward group's lines can be split over several workgroups, whose
// partial rows are added in a fixed order afterwards: with one workgroup
// per view, a few hundred views leave the 256 CUs unevenly loaded.
// Automatic choice, from a sweep of 1..8 parts on MI355X (90..720 views,
// 128^2..512^2 pixels): 4 parts (within 10 % of the best count in every
// case measured; 512^2 with 360 views 94 -> 65 us, with 90 views 51 ->
// 35 us) for images of at least 256 x 256 pixels with fewer than two
// forward groups per CU.  Smaller images keep one workgroup per group (the
// extra launch costs more than the split saves: 128^2 with 60 views 12.5 ->
// 14.8 us), and so do plans with more groups: at 512^2 with 720 views the
// split would save 13 % of the forward, but its rounding took a 100-step
// ADMM-TV run with early-stopping CG outside the drift envelope of the
// reference-arithmetic runs at one iteration, which the unsplit projector
// stays well inside.
int auto_split(const Plan& p) {
  if (p.fwd_generic || static_cast<long long>(p.M) * p.N < 65536) return 1;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    return 1;
  return p.n_groups < 2 * cus ? 4 : 1;
}

// Sets the forward split (1..FWD_SP_MAX parts; 1 for the general-geometry
// forward), growing the partial-sinogram workspace when needed.
int set_split(Plan& p, int parts) {
  if (p.fwd_generic) parts = 1;
  if (parts > p.fwd_sp_cap) {
    float* w = nullptr;
    if (hipMalloc(&w, sizeof(float) * parts * static_cast<size_t>(p.A) * p.D) != hipSuccess) {
      (void)hipGetLastError();  // the plan keeps its current split; later launches stay clean
      return fail(SCXR_ERR_NOMEM, "hipMalloc of the %d-part forward workspace failed", parts);
    }
    if (p.parts) (void)hipFree(p.parts);
    p.parts = w;
    p.fwd_sp_cap = parts;
  }
  p.fwd_sp = parts;
  return SCXR_OK;
}

void destroy(Plan& p) {
  if (p.parts) (void)hipFree(p.parts);
  p.parts = nullptr;
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

// The reference computes RN(min(dd, width) / width) with dd = 1 - (Px - ind).
// Here the quotient is clamped instead: if dd < width both forms give
// RN(dd / width); if dd >= width, RN(dd * rw) >= 1 and both give 1.0.
__device__ __forceinline__ Weight weight_at(float Pxmin, float P0, float P1, float width,
                                            double rw, float fi, float fj) {
  (void)width;
  const float Px = __fadd_rn(__fadd_rn(Pxmin, __fmul_rn(P0, fi)), __fmul_rn(P1, fj));
  const float fl = floorf(Px);
  const float dd = __fsub_rn(1.0f, __fsub_rn(Px, fl));
  return Weight{fl, div_rn_sat(dd, rw)};
}

// The same weight from the fraction instruction, for Px >= 1: there
// f = fract(Px) = Px - floor(Px) exactly, and 1 - f is exact in float32 as
// well (a multiple of ulp(Px) >= 2^-23 in (0, 1]), so the float64
// multiply-add rw - f rw rounds the product (1 - f) rw once, exactly as
// div_rn(1 - f, rw) does (the adjoint's path when every
// projected coordinate of the plan is at least 1).
__device__ __forceinline__ float weight_frac(float Px, double rw) {
  const double f = static_cast<double>(__builtin_amdgcn_fractf(Px));
  return __builtin_amdgcn_fmed3f(static_cast<float>(fma(-f, rw, rw)), 0.0f, 1.0f);
}

// Transpose (M, N) -> (N, M) through a padded LDS tile (tile row by,
// column bx).
__device__ __forceinline__ void transpose_tile(const float* __restrict__ in,
                                               float* __restrict__ out, int M, int N, int by,
                                               int bx) {
  __shared__ float tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  const int i0 = by * 32, j0 = bx * 32;
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

__global__ __launch_bounds__(256) void transpose_kernel(const float* __restrict__ in,
                                                        float* __restrict__ out, int M, int N,
                                                        const int32_t* pred) {
  if (pred_off(pred)) return;
  transpose_tile(in, out, M, N, blockIdx.y, blockIdx.x);
}

// Forward projection.  One workgroup owns up to V whole sinogram rows
// (views) and streams the entire image.  Each of its FWD_WAVES waves
// accumulates into private LDS copies of every owned row (one per chain),
// and the copies are summed in
// a fixed order at the end: deterministic, no global atomics.  Row slot k
// holds detector bin k - 2; bins outside [0, D) land in two guard slots on
// each side.
//
// No LDS atomics either: within one instruction the 64 lanes take pixels S
// apart along the fast image axis, where S * |P_fast| >= 2 (S is chosen on
// the host per group of views).  Two such pixels project at least two bins
// apart, so the lanes' two-bin footprints {ind, ind + 1} are disjoint and a
// plain LDS read-modify-write is race free.  The S phases of a wave run in
// program order; LDS executes a wave's accesses in order, and a compiler
// fence keeps the phases' memory operations in that order.
// Each wave runs CH independent chains (lines s, s + FWD_WAVES, ..., each
// with its own private row copy) so that the chains' LDS round trips
// overlap each other's arithmetic; 8 waves of 2 chains measured 3 % faster
// than 4 waves of 4 for the same LDS.  Px = (Pxmin + P0 * i) + P1 * j is evaluated in the
// reference order with the line-constant product hoisted: for row-fast
// views t = Pxmin + P0 * i is constant along the line, for column-fast
// views t = P1 * j is.
// Accumulate one phase (lanes at f = c0 + S lane + r) of the CH lines.
template <int V, int CH, bool COLFAST, bool FR>
__device__ __forceinline__ void fwd_phase(float* acc, const int (&row)[CH][V],
                                          const float (&t)[CH][V], const float* val,
                                          float ff, int cnt, const float* Pxmin,
                                          const float* P0, const float* P1, const double* rw,
                                          float Dmax) {
#pragma unroll
  for (int v = 0; v < V; ++v) {
    if (V == 1 || v < cnt) {
      // the lane-axis term is shared by the chains
      const float tl = COLFAST ? __fadd_rn(Pxmin[v], __fmul_rn(P0[v], ff)) : __fmul_rn(P1[v], ff);
      int k[CH];
      float w[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const float P = COLFAST ? __fadd_rn(tl, t[c][v]) : __fadd_rn(t[c][v], tl);
        if (FR) {  // every pixel's Px >= 1: the fraction instruction (see weight_frac)
          w[c] = weight_frac(P, rw[v]);
          // lanes past the line end (value 0) can sit anywhere: the bin is
          // clamped to the padded row [-2, D] as on the floor path
          k[c] = row[c][v] + cvt_flr_i32(__builtin_amdgcn_fmed3f(P, -2.0f, Dmax));
        } else {
          const float fl = floorf(P);
          w[c] = div_rn_sat(__fsub_rn(1.0f, __fsub_rn(P, fl)), rw[v]);
          k[c] = row[c][v] + static_cast<int>(__builtin_amdgcn_fmed3f(fl, -2.0f, Dmax));
        }
      }
      float a0[CH], a1[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        a0[c] = acc[k[c]];
        a1[c] = acc[k[c] + 1];
      }
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        acc[k[c]] = fmaf(val[c], w[c], a0[c]);
        acc[k[c] + 1] = fmaf(val[c], __fsub_rn(1.0f, w[c]), a1[c]);
      }
    }
  }
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_wave_barrier();
}

// Forward projection of a view group by one wave.  Chain c of the wave
// streams lines s0 + FWD_WAVES c (s0 = wave + FWD_WAVES CH g) along the
// fast axis; lane l
// takes pixels f = c0 + S l + r in phase r, so the lanes of one phase hit
// bins at least two apart and each chain's LDS row is updated without
// atomics.  The walk over (line group, block c0, phase r) is flat and the
// image values of the next phase are loaded before the current phase is
// accumulated.
template <int V, int CH, bool COLFAST, bool FR>
__device__ __forceinline__ void fwd_body(const float* __restrict__ src, float* acc, int L,
                                         int cnt, int S, int kmul, const float* Pxmin,
                                         const float* P0, const float* P1, const double* rw,
                                         int M, int N, int D) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const unsigned Lf = COLFAST ? M : N;  // fast (lane) axis length
  const int Ls = COLFAST ? N : M;
  const int first = wave + FWD_WAVES * CH * static_cast<int>(blockIdx.y);
  if (first >= Ls) return;
  const float Dmax = static_cast<float>(D);
  const unsigned span = 64u * S;
  const unsigned loff = static_cast<unsigned>(S * ((lane * kmul) & 63));
  int row[CH][V];
#pragma unroll
  for (int v = 0; v < V; ++v)
#pragma unroll
    for (int c = 0; c < CH; ++c) row[c][v] = ((v * FWD_WAVES + wave) * CH + c) * L + 2;
  // Line group s0: chain c streams line s0 + 4c through its own buffer;
  // lanes past the line end (and lines past the image) read zeros, which
  // add nothing.
  __amdgpu_buffer_rsrc_t rs[CH];
  auto set_lines = [&](int s0) {
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int sc = s0 + FWD_WAVES * c;
      const bool has = sc < Ls;
      rs[c] = make_rsrc(src + static_cast<size_t>(has ? sc : s0) * Lf,
                        has ? static_cast<int>(Lf * 4u) : 0);
    }
  };
  // The walk over (line group, block, phase) is flat; the image values of
  // the next phase are loaded before the current phase is accumulated.
  int s0 = first, sn = first;
  unsigned c0 = 0, cn = 0;
  int r = 0, rn = 0;
  set_lines(sn);
  float nxt[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) nxt[c] = bload1(rs[c], (loff) * 4u, 0);
  float t[CH][V];
  int tline = -1;
  while (true) {
    float val[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) val[c] = nxt[c];
    s0 = sn;
    c0 = cn;
    r = rn;
    // advance to the next phase
    if (++rn == S) {
      rn = 0;
      cn += span;
      if (cn >= Lf) {
        cn = 0;
        sn += FWD_WAVES * CH * static_cast<int>(gridDim.y);
        if (sn < Ls) set_lines(sn);
      }
    }
    if (sn < Ls) {
      const unsigned fo = (cn + loff + rn) * 4u;
#pragma unroll
      for (int c = 0; c < CH; ++c) nxt[c] = bload1(rs[c], fo, 0);
    }
    if (s0 != tline) {
      tline = s0;
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const float fs = static_cast<float>(s0 + FWD_WAVES * c);
#pragma unroll
        for (int v = 0; v < V; ++v)
          t[c][v] = COLFAST ? __fmul_rn(P1[v], fs) : __fadd_rn(Pxmin[v], __fmul_rn(P0[v], fs));
      }
    }
    const float ff = static_cast<float>(c0 + loff + r);
    fwd_phase<V, CH, COLFAST, FR>(acc, row, t, val, ff, cnt, Pxmin, P0, P1, rw, Dmax);
    if (sn >= Ls) break;
  }
}

// The same walk for a lane stride S known at compile time: a block
// (line group, c0) is the S phases of every chain; lane l's S pixels of a
// block are contiguous (c0 + S l + r, r < S), loaded together one block
// ahead into one of two register buffers that alternate between blocks,
// so the per-phase loop bookkeeping and register copies of the runtime-S
// walk disappear.
template <int V, int CH, bool COLFAST, int S, bool FR>
__device__ __forceinline__ void fwd_body_s(const float* __restrict__ src, float* acc, int L,
                                           int cnt, int kmul, const float* Pxmin, const float* P0,
                                           const float* P1, const double* rw, int M, int N,
                                           int D) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const unsigned Lf = COLFAST ? M : N;
  const int Ls = COLFAST ? N : M;
  const int first = wave + FWD_WAVES * CH * static_cast<int>(blockIdx.y);
  if (first >= Ls) return;
  const float Dmax = static_cast<float>(D);
  constexpr unsigned span = 64u * S;
  const unsigned loff = static_cast<unsigned>(S * ((lane * kmul) & 63));
  int row[CH][V];
#pragma unroll
  for (int v = 0; v < V; ++v)
#pragma unroll
    for (int c = 0; c < CH; ++c) row[c][v] = ((v * FWD_WAVES + wave) * CH + c) * L + 2;
  __amdgpu_buffer_rsrc_t rs[CH];
  auto set_lines = [&](int s0) {
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int sc = s0 + FWD_WAVES * c;
      const bool has = sc < Ls;
      rs[c] = make_rsrc(src + static_cast<size_t>(has ? sc : s0) * Lf,
                        has ? static_cast<int>(Lf * 4u) : 0);
    }
  };
  float t[CH][V];
  auto set_t = [&](int s0) {
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const float fs = static_cast<float>(s0 + FWD_WAVES * c);
#pragma unroll
      for (int v = 0; v < V; ++v)
        t[c][v] = COLFAST ? __fmul_rn(P1[v], fs) : __fadd_rn(Pxmin[v], __fmul_rn(P0[v], fs));
    }
  };
  // Inside a line the compiler merges a lane's S dword loads into one or
  // two wide loads.  A block that runs past the line end keeps one
  // bounds-checked dword load per pixel, so that exactly the pixels past
  // the end (and every pixel of a line past the image) read zero, whatever
  // the range check of a partially out-of-range wide load returns: opaque
  // offsets prevent the merge, and the work-group coherence scope (still
  // served by the CU's vector cache) keeps the two load sequences from
  // being combined into one.
  auto load = [&](float (&b)[CH][S], unsigned c0) {
    const unsigned vo = (c0 + loff) * 4u;
    if (c0 + span <= Lf) {
#pragma unroll
      for (int c = 0; c < CH; ++c)
#pragma unroll
        for (int r = 0; r < S; ++r) b[c][r] = bload1(rs[c], vo + 4u * r, 0);
    } else {
      unsigned vb = vo, o[S];
      asm volatile("" : "+v"(vb));
#pragma unroll
      for (int r = 0; r < S; ++r) {
        o[r] = vb + 4u * r;
        asm volatile("" : "+v"(o[r]));
      }
#pragma unroll
      for (int c = 0; c < CH; ++c)
#pragma unroll
        for (int r = 0; r < S; ++r)
          b[c][r] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs[c], o[r], 0, 1));
    }
  };
  int s0 = first;
  unsigned c0 = 0;
  set_lines(s0);
  set_t(s0);
  // Accumulate block (s0, c0) from cur after issuing the loads of the next
  // block into nxt; false once the last block is done.
  // Lane l's pixel of phase r in block c0 is c0 + loff + r.  c0 and loff + r
  // are below 2^24 (plan creation checks the image size), so both convert
  // exactly and float(c0) + float(loff + r) rounds the same integer once, as
  // the conversion of the sum does: one add per phase instead of an integer
  // add and a conversion.
  float flr[S];
#pragma unroll
  for (int r = 0; r < S; ++r) flr[r] = static_cast<float>(loff + r);
  auto step = [&](const float (&cur)[CH][S], float (&nxt)[CH][S]) {
    int sn = s0;
    unsigned cn = c0 + span;
    if (cn >= Lf) {
      cn = 0;
      sn += FWD_WAVES * CH * static_cast<int>(gridDim.y);
    }
    const bool more = sn < Ls;
    if (more) {
      if (sn != s0) set_lines(sn);
      load(nxt, cn);
    }
    const float fc0 = static_cast<float>(c0);
#pragma unroll
    for (int r = 0; r < S; ++r) {
      float val[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) val[c] = cur[c][r];
      fwd_phase<V, CH, COLFAST, FR>(acc, row, t, val, __fadd_rn(fc0, flr[r]), cnt, Pxmin, P0, P1,
                                    rw, Dmax);
    }
    // a new line every few blocks: branch (hinted) rather than recompute the
    // line terms every block and select, which is what the compiler does
    // otherwise
    if (__builtin_expect(more && sn != s0, 0)) set_t(sn);
    s0 = sn;
    c0 = cn;
    return more;
  };
  float ba[CH][S], bb[CH][S];
  load(ba, 0);
  while (step(ba, bb) && step(bb, ba)) {
  }
}

// One forward group (gi indexes groups[]); acc is the dynamic LDS array.
template <int V, int CH, bool FR>
__device__ __forceinline__ void fwd_group(const float* __restrict__ im,
                                          const float* __restrict__ imT,
                                          float* __restrict__ sino,
                                          const float4* __restrict__ geom,
                                          const double* __restrict__ rwt,
                                          const int4* __restrict__ groups,
                                          const int* __restrict__ order, int M, int N, int D,
                                          const float* __restrict__ yres,
                                          double* __restrict__ res_ws, int gi, float* acc) {
  __shared__ double red[FWD_WAVES];
  const int L = D + 4;
  const int4 grp = groups[gi];
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
  for (int e = threadIdx.x; e < V * FWD_WAVES * CH * L; e += 64 * FWD_WAVES) acc[e] = 0.0f;
  __syncthreads();
  // lane strides 3 to 5 cover unit-spaced detectors with pixel sizes
  // sqrt(2)/2 to 1; other strides take the runtime-stride walk
  const bool colf = grp.z & 1;
  const int kmul = grp.z >> 1;  // lane permutation multiplier (odd)
  const float* src = colf ? imT : im;
#define SCXR_FWD_S(SS)                                                                 \
  if (colf)                                                                            \
    fwd_body_s<V, CH, true, SS, FR>(src, acc, L, cnt, kmul, Pxmin, P0, P1, rw, M, N, D);   \
  else                                                                                 \
    fwd_body_s<V, CH, false, SS, FR>(src, acc, L, cnt, kmul, Pxmin, P0, P1, rw, M, N, D);
  switch (grp.w) {
    case 3:
      SCXR_FWD_S(3)
      break;
    case 4:
      SCXR_FWD_S(4)
      break;
    case 5:
      SCXR_FWD_S(5)
      break;
    default:
      if (colf)
        fwd_body<V, CH, true, FR>(src, acc, L, cnt, grp.w, kmul, Pxmin, P0, P1, rw, M, N, D);
      else
        fwd_body<V, CH, false, FR>(src, acc, L, cnt, grp.w, kmul, Pxmin, P0, P1, rw, M, N, D);
  }
#undef SCXR_FWD_S
  __syncthreads();
  double res = 0.0;
#pragma unroll
  for (int v = 0; v < V; ++v) {
    if (v < cnt) {
      const float* r = acc + v * FWD_WAVES * CH * L + 2;
      float* dst = sino + static_cast<size_t>(view[v]) * D;
      const float* yrow = yres ? yres + static_cast<size_t>(view[v]) * D : nullptr;
      for (int b = threadIdx.x; b < D; b += 64 * FWD_WAVES) {
        float val = r[b];
        for (int c = 1; c < FWD_WAVES * CH; ++c) val = __fadd_rn(val, r[c * L + b]);
        dst[b] = val;
        if (yrow) {
          const double d = static_cast<double>(__fsub_rn(yrow[b], val));
          res += d * d;
        }
      }
    }
  }
  if (res_ws) {  // per-group partial of sum (y - A x)^2
    // fixed order: wave shuffles, then the wave sums left to right
    for (int off = 32; off > 0; off >>= 1) res += __shfl_down(res, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = res;
    __syncthreads();
    if (threadIdx.x == 0) {
      double q = red[0];
      for (int w = 1; w < FWD_WAVES; ++w) q += red[w];
      res_ws[gi] = q;
    }
  }
}

// Workgroup (b, part) of a grid (groups, SP) runs forward group gblock[b]
// (the CU-balanced placement of the plan), taking every SP-th set of
// FWD_WAVES * CH lines of it, from set `part`; with SP > 1 it writes
// its rows to partial sinogram `part` of `parts` (A x D each), which
// fwd_parts_kernel adds up.
template <int V, int CH, bool FR>
__global__ __launch_bounds__(64 * FWD_WAVES) void fwd_kernel(const float* __restrict__ im,
                                                  const float* __restrict__ imT,
                                                  float* __restrict__ sino,
                                                  const float4* __restrict__ geom,
                                                  const double* __restrict__ rwt,
                                                  const int4* __restrict__ groups,
                                                  const int* __restrict__ order, int M, int N,
                                                  int D, const float* __restrict__ yres,
                                                  double* __restrict__ res_ws,
                                                  float* __restrict__ parts, int A,
                                                  const int* __restrict__ gblock,
                                                  const int32_t* pred) {
  if (pred_off(pred)) return;
  extern __shared__ float acc[];  // [V][FWD_WAVES waves][CH chains][L]
  const int gi = gblock[blockIdx.x];
  if (gridDim.y > 1)
    fwd_group<V, CH, FR>(im, imT, parts + static_cast<size_t>(blockIdx.y) * A * D, geom, rwt,
                         groups, order, M, N, D, nullptr, nullptr, gi, acc);
  else
    fwd_group<V, CH, FR>(im, imT, sino, geom, rwt, groups, order, M, N, D, yres, res_ws, gi,
                         acc);
}

// sino = parts[0] + parts[1] + ... (in part order) for the views of group
// blockIdx.x, with that group's residual partial as fwd_group writes it.
__global__ __launch_bounds__(256) void fwd_parts_kernel(const float* __restrict__ parts, int sp,
                                                        int A, int D, float* __restrict__ sino,
                                                        const int4* __restrict__ groups,
                                                        const int* __restrict__ order,
                                                        const float* __restrict__ yres,
                                                        double* __restrict__ res_ws,
                                                        const int32_t* pred) {
  if (pred_off(pred)) return;
  __shared__ double red[4];
  const int gi = blockIdx.x;
  const int4 grp = groups[gi];
  const size_t plane = static_cast<size_t>(A) * D;
  double res = 0.0;
  for (int v = 0; v < grp.y; ++v) {
    const size_t row = static_cast<size_t>(order[grp.x + v]) * D;
    for (int b = threadIdx.x; b < D; b += 256) {
      float val = parts[row + b];
      for (int q = 1; q < sp; ++q) val = __fadd_rn(val, parts[q * plane + row + b]);
      sino[row + b] = val;
      if (yres) {
        const double d = static_cast<double>(__fsub_rn(yres[row + b], val));
        res += d * d;
      }
    }
  }
  if (res_ws) {
    res = block_sum_f64(res, red);
    if (threadIdx.x == 0) res_ws[gi] = res;
  }
}

// Dynamic LDS beyond 64 KiB for the one-view, one-chain forward (wide
// detectors).
static hipError_t fwd_allow_lds(int bytes) {
  const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&fwd_kernel<1, 1, false>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(&fwd_kernel<1, 1, true>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

// Adjoint (back projection).  One workgroup per pixel tile of ADJ_TI rows
// and ADJ_TJ * PPT columns (each thread owns PPT pixels of one row, ADJ_TJ
// apart); views are processed in chunks of ADJ_VC.  For each view the
// detector window reachable from the tile (W <= 64 bins, zero outside
// [0, D)) and the view's geometry are staged in LDS, then every thread
// gathers its pixels' two bins.  The windows of the next chunk are
// prefetched into registers while the current chunk is computed, so the
// global-load latency is hidden behind the gather arithmetic.  The two sums
// of _xray2d.py:299-304 are kept separately and added at the end.
// Px = (Pxmin + P0 i) + P1 j keeps the reference order; the row term
// Pxmin + P0 i is shared by a thread's pixels.
//
// Window origin of view a for the tile at (i0, j0):
//   lo = floor(P0 * i0 + P1 * j0 + clo[a]),
//   clo[a] = Pxmin + min(P0, 0) (TI - 1) + min(P1, 0) (TJ PPT - 1) - 1,
// a lower bound of floor(Px) over the tile with a one-bin margin for the
// float32 rounding of Px; W covers the tile's projected extent plus the
// margins, so every pixel's bins ind and ind + 1 lie inside the window.
struct __align__(16) ViewLds {
  float Pxmin, P0, P1;
  // byte offset in win_all of detector bin 0 for this view: an integer
  // (float32 bits when the weights use floor, so that 4 fl + base4 is one
  // exact float multiply-add; int32 when they use the fraction instruction)
  uint32_t base4;
  double rw;
  double pad;
};

// One adjoint tile (tile row by, column bx; dot partial written at index
// dot_i).  The workgroup is VS groups of 256 threads over the same pixels
// (ADJ_VS for single slices; batches may take 2); group g takes the g-th
// contiguous share of the views, and the groups' sums are added in group
// order at the end, so VS decides each pixel's summation order.
template <int PPT, bool FR, int VS = ADJ_VS>
__device__ __forceinline__ void adj_tile(const float* __restrict__ sino, float* __restrict__ out,
                                         const float4* __restrict__ geom,
                                         const double* __restrict__ rwt,
                                         const float* __restrict__ clo, int M, int N, int A,
                                         int D, int W, const AdjEpi& epi, int bx, int by,
                                         int dot_i) {
  constexpr int VPW = ADJ_VC / 4;  // views staged per wave
  __shared__ float win_all[VS][ADJ_VC * ADJ_WFAST];
  __shared__ ViewLds gv_all[VS][ADJ_VC];
  __shared__ double red[4 * VS];
  const int grp = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8), t = threadIdx.x & 255;
  float* win = win_all[grp];
  ViewLds* gv = gv_all[grp];
  const int tx = t % ADJ_TJ, ty = t / ADJ_TJ;
  const int i0 = by * ADJ_TI, j0 = bx * (ADJ_TJ * PPT);
  const int i = i0 + ty;
  int j[PPT];
  bool inside[PPT];
  float fj[PPT];
  bool any = false;
#pragma unroll
  for (int q = 0; q < PPT; ++q) {
    j[q] = j0 + tx + ADJ_TJ * q;
    inside[q] = (i < M) && (j[q] < N);
    fj[q] = static_cast<float>(j[q]);
    any = any || inside[q];
  }
  const float fi = static_cast<float>(i);
  const float fi0 = static_cast<float>(i0), fj0 = static_cast<float>(j0);
  const int wave = t >> 6, lane = t & 63;
  const char* winb = reinterpret_cast<const char*>(&win_all[0][0]);
  const int gbase4 = grp * (ADJ_VC * ADJ_WFAST * 4);
  // this group's views [vb, ve); every group runs the same number of chunks
  const int share = (A + VS - 1) / VS;
  const int vb = min(A, grp * share), ve = min(A, vb + share);
  const int nchunk = (share + ADJ_VC - 1) / ADJ_VC;

  // Staging is spread over the lanes: lane 8 sq + sw of a wave handles
  // view slot sv = wave + 4 sq of a chunk; it computes that view's window
  // origin and writes word sw of its LDS record.  The windows themselves
  // are read by the whole wave, one view at a time, through a buffer
  // resource bounded to the view's sinogram row, so bins outside [0, D)
  // read zero.
  const int wv = __builtin_amdgcn_readfirstlane(wave);
  const int sq = lane >> 3, sw = lane & 7, sv = wv + 4 * sq;
  float pf[VPW];
  int lo_nx = 0;  // window origin of this lane's view slot in the prefetched chunk
  auto prefetch = [&](int a0) {
    const int a = a0 + sv;
    const int as = a < ve ? a : 0;
    const float4 g = geom[as];
    const int lo = static_cast<int>(floorf(fmaf(g.y, fi0, fmaf(g.z, fj0, clo[as]))));
    lo_nx = a < ve ? lo : 0;
#pragma unroll
    for (int q = 0; q < VPW; ++q) {
      const int aq = a0 + wv + 4 * q;
      float val = 0.0f;
      if (aq < ve) {
        const int loq = __builtin_amdgcn_readlane(lo, 8 * q);
        const __amdgpu_buffer_rsrc_t rs = make_rsrc(sino + static_cast<size_t>(aq) * D, D * 4);
        val = bload1(rs, static_cast<unsigned>(loq + lane) * 4u, 0);
      }
      pf[q] = val;
    }
  };

  float acc0[PPT], acc1[PPT];
#pragma unroll
  for (int q = 0; q < PPT; ++q) acc0[q] = acc1[q] = 0.0f;
  prefetch(vb);
  for (int c = 0; c < nchunk; ++c) {
    const int a0 = vb + c * ADJ_VC;
    const int nv = max(0, min(ADJ_VC, ve - a0));
    const int nv4 = (nv + 3) & ~3;  // padded views: zero window, benign geometry
    __syncthreads();  // the previous chunk's gathers are done
#pragma unroll
    for (int q = 0; q < VPW; ++q) {
      const int v = wv + 4 * q;
      if (v < nv4 && lane < W) win[v * W + lane] = v < nv ? pf[q] : 0.0f;
    }
    if (sv < nv4) {
      // record words: Pxmin, P0, P1, base4, rw (2 words), pad (2 words);
      // padded views get zero geometry and rw = 1
      const bool real = sv < nv;
      const int a = real ? a0 + sv : 0;
      uint32_t word = 0;
      if (sw < 3) {
        word = real ? reinterpret_cast<const uint32_t*>(geom)[4 * a + sw] : 0u;
      } else if (sw == 3) {
        const int b4 = gbase4 + 4 * (sv * W - (real ? lo_nx : 0));
        word = FR ? static_cast<uint32_t>(b4) : __float_as_uint(static_cast<float>(b4));
      } else if (sw < 6) {
        word = real ? reinterpret_cast<const uint32_t*>(rwt)[2 * a + (sw - 4)]
                    : (sw == 5 ? 0x3ff00000u : 0u);
      }
      reinterpret_cast<uint32_t*>(gv)[8 * sv + sw] = word;
    }
    __syncthreads();
    if (a0 + ADJ_VC < ve) prefetch(a0 + ADJ_VC);
    if (any) {
      for (int v = 0; v < nv4; v += 4) {
        float w[4][PPT];
        int k[4][PPT];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const ViewLds tv = gv[v + u];
          const float ti = __fadd_rn(tv.Pxmin, __fmul_rn(tv.P0, fi));
#pragma unroll
          for (int q = 0; q < PPT; ++q) {
            const float Px = __fadd_rn(ti, __fmul_rn(tv.P1, fj[q]));
            if (FR) {
              w[u][q] = weight_frac(Px, tv.rw);
              k[u][q] = (cvt_flr_i32(Px) << 2) + static_cast<int>(tv.base4);
            } else {
              const float fl = floorf(Px);
              const float dd = __fsub_rn(1.0f, __fsub_rn(Px, fl));
              w[u][q] = div_rn_sat(dd, tv.rw);
              k[u][q] = static_cast<int>(
                  static_cast<unsigned>(__fmaf_rn(fl, 4.0f, __uint_as_float(tv.base4))));
            }
          }
        }
        float y0[4][PPT], y1[4][PPT];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int q = 0; q < PPT; ++q) {
            y0[u][q] = *reinterpret_cast<const float*>(winb + k[u][q]);
            y1[u][q] = *reinterpret_cast<const float*>(winb + k[u][q] + 4);
          }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int q = 0; q < PPT; ++q) {
            acc0[q] = fmaf(y0[u][q], w[u][q], acc0[q]);
            acc1[q] = fmaf(y1[u][q], __fsub_rn(1.0f, w[u][q]), acc1[q]);
          }
      }
    }
  }
  // add the other groups' sums, in group order, into group 0
  if (VS > 1) {
    float* xch = win_all[0];  // 2 PPT floats per thread of each group > 0
    static_assert((VS - 1) * 2 * PPT * 256 <= ADJ_VC * ADJ_WFAST,
                  "group exchange must fit one window buffer");
    __syncthreads();
    if (grp > 0) {
#pragma unroll
      for (int q = 0; q < PPT; ++q) {
        xch[((grp - 1) * 2 * PPT + 2 * q) * 256 + t] = acc0[q];
        xch[((grp - 1) * 2 * PPT + 2 * q + 1) * 256 + t] = acc1[q];
      }
    }
    __syncthreads();
    if (grp == 0) {
      for (int g2 = 1; g2 < VS; ++g2)
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
          acc0[q] = __fadd_rn(acc0[q], xch[((g2 - 1) * 2 * PPT + 2 * q) * 256 + t]);
          acc1[q] = __fadd_rn(acc1[q], xch[((g2 - 1) * 2 * PPT + 2 * q + 1) * 256 + t]);
        }
    }
  }
  double dot = 0.0;
  const bool circ = epi.circular != 0;
  if (grp == 0) {
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
      if (!inside[q]) continue;
      float v = __fadd_rn(acc0[q], acc1[q]);
      if (epi.kind == EPI_GRAM) {  // (rho * C^T C)(p) + A^T A p   (_admmaux.py:222-229)
        v = __fadd_rn(__fmul_rn(epi.rho, fd::gram2d(epi.p, M, N, i, j[q], circ)), v);
      } else if (epi.kind == EPI_RHS) {  // A^T y + rho * C^T (z - u)   (_admmaux.py:245-256)
        v = __fadd_rn(v, __fmul_rn(epi.rho, fd::adj2d_diff(epi.z, epi.u, M, N, i, j[q], circ)));
      }
      out[static_cast<size_t>(i) * N + j[q]] = v;
      if (epi.dot_ws) dot += static_cast<double>(epi.p[static_cast<size_t>(i) * N + j[q]]) * v;
    }
  }
  if (epi.dot_ws) {  // per-tile partial of sum p * out (CG curvature p^T A p)
    for (int off = 32; off > 0; off >>= 1) dot += __shfl_down(dot, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = dot;
    __syncthreads();
    if (threadIdx.x == 0) epi.dot_ws[dot_i] = ((red[0] + red[1]) + red[2]) + red[3];
  }
}

template <int PPT, bool FR>
__global__ __launch_bounds__(256 * ADJ_VS) __attribute__((amdgpu_waves_per_eu(4))) void adj_kernel(const float* __restrict__ sino,
                                                  float* __restrict__ out,
                                                  const float4* __restrict__ geom,
                                                  const double* __restrict__ rwt,
                                                  const float* __restrict__ clo, int M, int N,
                                                  int A, int D, int W, AdjEpi epi,
                                                  const int32_t* pred) {
  if (pred_off(pred)) return;
  adj_tile<PPT, FR>(sino, out, geom, rwt, clo, M, N, A, D, W, epi, blockIdx.x, blockIdx.y,
                blockIdx.y * gridDim.x + blockIdx.x);
}

// FBP row filter (_xray2d.py:187-194): the reference convolves each row
// with the ramp taps h[0..D) by zero-padded FFT of length 2D - 1 and keeps
// [s, 2D - 1 - s), s = (D - 1) / 2; here the same linear convolution is
// summed directly: out[a, t] = sum_m h[m] y[a, t + s - m], 0 <= t + s - m < D.
__global__ __launch_bounds__(256) void ramp_kernel(const float* __restrict__ y,
                                                   float* __restrict__ out,
                                                   const float* __restrict__ taps, int D, int Dout,
                                                   int s, const int32_t* pred) {
  if (pred_off(pred)) return;
  __shared__ float row[4096];
  __shared__ float h[4096];
  const int a = blockIdx.x;
  for (int k = threadIdx.x; k < D; k += 256) {
    row[k] = y[static_cast<size_t>(a) * D + k];
    h[k] = taps[k];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < Dout; t += 256) {
    const int c = t + s;
    const int m0 = max(0, c - D + 1), m1 = min(D - 1, c);
    float acc = 0.0f;
    for (int m = m0; m <= m1; ++m) acc = fmaf(h[m], row[c - m], acc);
    out[static_cast<size_t>(a) * Dout + t] = acc;
  }
}

// ------------------------------------------------ general geometries --
// Plans whose geometry the tiled kernels cannot hold (a detector row wider
// than the LDS, projected pixel spacing below 1/32 bin, or a pixel tile
// projecting over more than 64 bins) run these two gathers instead: the
// same bins and weights (weight_at), deterministic, no LDS staging.

// Adjoint: one thread per pixel, views in order, the two sums of
// _xray2d.py:299-304 kept apart and added at the end.
__global__ __launch_bounds__(256) void adj_generic_kernel(const float* __restrict__ sino,
                                                          float* __restrict__ im,
                                                          const float4* __restrict__ geom,
                                                          const double* __restrict__ rwt, int M,
                                                          int N, int A, int D,
                                                          const int32_t* pred) {
  if (pred_off(pred)) return;
  const long long e = blockIdx.x * 256ll + threadIdx.x;
  if (e >= static_cast<long long>(M) * N) return;
  const int i = static_cast<int>(e / N), j = static_cast<int>(e - static_cast<long long>(i) * N);
  const float fi = static_cast<float>(i), fj = static_cast<float>(j);
  const float Dmax = static_cast<float>(D);
  float s0 = 0.0f, s1 = 0.0f;
  for (int a = 0; a < A; ++a) {
    const float4 g = geom[a];
    const Weight wt = weight_at(g.x, g.y, g.z, g.w, rwt[a], fi, fj);
    const int k = static_cast<int>(__builtin_amdgcn_fmed3f(wt.fl, -2.0f, Dmax));
    const float* y = sino + static_cast<size_t>(a) * D;
    if (k >= 0 && k < D) s0 = fmaf(y[k], wt.w, s0);
    if (k + 1 >= 0 && k + 1 < D) s1 = fmaf(y[k + 1], __fsub_rn(1.0f, wt.w), s1);
  }
  im[e] = __fadd_rn(s0, s1);
}

// Forward: one thread per (bin b, view a).  Along the slow image axis s the
// projected coordinate is monotone in the fast index f, so the pixels of a
// line whose bin is b or b - 1 lie in a short run of f; the run is located
// in float64 (widened by one pixel each side) and every candidate's bin is
// then decided by the exact float32 arithmetic.
__global__ __launch_bounds__(256) void fwd_generic_kernel(const float* __restrict__ im,
                                                          float* __restrict__ sino,
                                                          const float4* __restrict__ geom,
                                                          const double* __restrict__ rwt, int M,
                                                          int N, int D, const int32_t* pred) {
  if (pred_off(pred)) return;
  const int a = blockIdx.y;
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= D) return;
  const float4 g = geom[a];
  const double rw = rwt[a];
  const bool rowfast = fabsf(g.z) >= fabsf(g.y);
  const int Ls = rowfast ? M : N, Lf = rowfast ? N : M;
  const double Pf = rowfast ? g.z : g.y, Ps = rowfast ? g.y : g.z;
  const float Dmax = static_cast<float>(D);
  // bound of the float32 rounding of Px (three roundings of terms at most
  // |Pxmin| + |Ps| Ls + |Pf| Lf), added to the bin interval [b - 1, b + 1)
  const double del = 4e-7 * (fabs(static_cast<double>(g.x)) + fabs(Ps) * Ls + fabs(Pf) * Lf) + 1e-6;
  float acc = 0.0f;
  for (int s = 0; s < Ls; ++s) {
    const double c = static_cast<double>(g.x) + Ps * s;
    const double f0 = (b - 1 - del - c) / Pf, f1 = (b + 1 + del - c) / Pf;
    const int lo = max(0, static_cast<int>(floor(fmin(f0, f1))) - 1);
    const int hi = min(Lf - 1, static_cast<int>(ceil(fmax(f0, f1))) + 1);
    for (int f = lo; f <= hi; ++f) {
      const int i = rowfast ? s : f, j = rowfast ? f : s;
      const Weight wt = weight_at(g.x, g.y, g.z, g.w, rw, static_cast<float>(i),
                                  static_cast<float>(j));
      const int k = static_cast<int>(__builtin_amdgcn_fmed3f(wt.fl, -2.0f, Dmax));
      if (k == b || k == b - 1) {
        const float v = im[static_cast<size_t>(i) * N + j];
        acc = fmaf(v, k == b ? wt.w : __fsub_rn(1.0f, wt.w), acc);
      }
    }
  }
  sino[static_cast<size_t>(a) * D + b] = acc;
}

// Debug dump of (ind, weight) for every (view, pixel).
__global__ __launch_bounds__(256) void weights_kernel(int32_t* __restrict__ inds,
                                                      float* __restrict__ wts,
                                                      const float4* __restrict__ geom,
                                                      const double* __restrict__ rwt, int M,
                                                      int N, int A, int fract) {
  const size_t n = static_cast<size_t>(M) * N;
  const size_t e = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
  if (e >= n * A) return;
  const int a = static_cast<int>(e / n);
  const size_t p = e - static_cast<size_t>(a) * n;
  const int i = static_cast<int>(p / N), j = static_cast<int>(p - static_cast<size_t>(i) * N);
  const float4 g = geom[a];
  if (fract) {  // the adjoint's arithmetic for plans with Px >= 1 everywhere
    const float Px = __fadd_rn(__fadd_rn(g.x, __fmul_rn(g.y, static_cast<float>(i))),
                               __fmul_rn(g.z, static_cast<float>(j)));
    inds[e] = cvt_flr_i32(Px);
    wts[e] = weight_frac(Px, rwt[a]);
    return;
  }
  const Weight wt =
      weight_at(g.x, g.y, g.z, g.w, rwt[a], static_cast<float>(i), static_cast<float>(j));
  inds[e] = static_cast<int32_t>(wt.fl);
  wts[e] = wt.w;
}

// ------------------------------------------------------------- batched --
// Ragged batches of slices in one launch: a persistent grid walks the work
// items (slice, group / tile) and runs exactly the single-slice bodies
// with that slice's tables, so every slice's result is bitwise the
// single-slice result.  A slice whose predicate is 0 skips its items.

__device__ __forceinline__ bool slice_off(const int32_t* pred, int pstride, int s) {
  return pred != nullptr && pred[static_cast<int64_t>(s) * pstride] == 0;
}

__global__ __launch_bounds__(256) void transpose_batch_kernel(BatchTabs t,
                                                              const float* __restrict__ im,
                                                              float* __restrict__ imT,
                                                              const int32_t* pred, int pstride) {
  for (int it = blockIdx.x; it < t.nt; it += gridDim.x) {
    const int4 w = t.titems[it];
    if (slice_off(pred, pstride, w.x)) continue;
    const int4 sh = t.shape[w.x];
    const int64_t io = t.ioff[w.x];
    transpose_tile(im + io, imT + io, sh.x, sh.y, w.y, w.z);
    __syncthreads();
  }
}

template <int V, int CH, bool FR>
__global__ __launch_bounds__(64 * FWD_WAVES) void fwd_batch_kernel(BatchTabs t, const float* __restrict__ im,
                                                        const float* __restrict__ imT,
                                                        float* __restrict__ sino,
                                                        const float* __restrict__ yres,
                                                        double* __restrict__ res_ws,
                                                        const int32_t* pred, int pstride) {
  extern __shared__ float acc[];
  for (int it = blockIdx.x; it < t.nf; it += gridDim.x) {
    const int2 w = t.fitems[it];
    const int s = w.x;
    if (slice_off(pred, pstride, s)) continue;
    const int4 sh = t.shape[s];
    const int64_t io = t.ioff[s], so = t.soff[s];
    const int vo = t.voff[s], go = t.goff[s];
    fwd_group<V, CH, FR>(im + io, imT + io, sino + so, t.geom + vo, t.rw + vo, t.groups + go,
                 t.order + vo, sh.x, sh.y, sh.w, yres ? yres + so : nullptr,
                 res_ws ? res_ws + go : nullptr, w.y, acc);
    __syncthreads();  // the next item re-zeroes acc
  }
}

template <int PPT, bool FR, int VS>
__global__ __launch_bounds__(256 * VS) __attribute__((amdgpu_waves_per_eu(4))) void adj_batch_kernel(BatchTabs t, const float* __restrict__ sino,
                                                        float* __restrict__ out, BatchEpi be,
                                                        const int32_t* pred, int pstride) {
  for (int it = blockIdx.x; it < t.na; it += gridDim.x) {
    const int2 w = t.aitems[it];
    const int s = w.x;
    if (slice_off(pred, pstride, s)) continue;
    const int4 sh = t.shape[s], me = t.meta[s];
    const int64_t io = t.ioff[s], so = t.soff[s];
    const int vo = t.voff[s];
    AdjEpi e;
    e.kind = be.kind;
    e.circular = be.circular;
    e.rho = be.rho;
    e.p = be.p ? be.p + io : nullptr;
    e.z = be.z ? be.z + 2 * io : nullptr;
    e.u = be.u ? be.u + 2 * io : nullptr;
    e.dot_ws = be.dot_ws ? be.dot_ws + t.toff[s] : nullptr;
    adj_tile<PPT, FR, VS>(sino + so, out + io, t.geom + vo, t.rw + vo, t.clo + vo, sh.x, sh.y,
                          sh.z, sh.w, me.x, e, w.y % me.z, w.y / me.z, w.y);
    __syncthreads();
  }
}

int batch_forward(const BatchTabs& t, int V, int Dmax, int grid, const float* im, float* imT,
                  float* sino, const float* yres, double* res_ws, const int32_t* pred,
                  int pstride, hipStream_t s) {
  if (t.nt > 0) {
    const unsigned g = grid > 0 ? std::min(grid, t.nt) : t.nt;
    hipLaunchKernelGGL(transpose_batch_kernel, dim3(g), dim3(256), 0, s, t, im, imT, pred,
                       pstride);
    SCXR_LAUNCHED("xr2d batch transpose");
  }
  const unsigned g = grid > 0 ? std::min(grid, t.nf) : t.nf;
  const size_t lds = fwd_lds_bytes(V, Dmax, t.ch);
#define SCXR_FWDB(VV, CC)                                                                    \
  do {                                                                                       \
    if (t.fract)                                                                             \
      hipLaunchKernelGGL((fwd_batch_kernel<VV, CC, true>), dim3(g), dim3(64 * FWD_WAVES), lds, s, \
                         t, im, imT, sino, yres, res_ws, pred, pstride);                     \
    else                                                                                     \
      hipLaunchKernelGGL((fwd_batch_kernel<VV, CC, false>), dim3(g), dim3(64 * FWD_WAVES), lds, \
                         s, t, im, imT, sino, yres, res_ws, pred, pstride);                  \
  } while (0)
  if (t.ch == 2) {
    if (V == 1)
      SCXR_FWDB(1, 2);
    else if (V == 2)
      SCXR_FWDB(2, 2);
    else
      SCXR_FWDB(4, 2);
  } else {
    if (V == 1)
      SCXR_FWDB(1, 1);
    else if (V == 2)
      SCXR_FWDB(2, 1);
    else
      SCXR_FWDB(4, 1);
  }
#undef SCXR_FWDB
  SCXR_LAUNCHED("xr2d batch forward");
  return SCXR_OK;
}

int batch_adjoint(const BatchTabs& t, int grid, const float* sino, float* im,
                  const BatchEpi& epi, const int32_t* pred, int pstride, hipStream_t s) {
  const unsigned g = grid > 0 ? std::min(grid, t.na) : t.na;
#define SCXR_ADJB(PP, FF)                                                                  \
  do {                                                                                     \
    if (t.avs == 2)                                                                        \
      hipLaunchKernelGGL((adj_batch_kernel<PP, FF, 2>), dim3(g), dim3(512), 0, s, t, sino, im, \
                         epi, pred, pstride);                                              \
    else                                                                                   \
      hipLaunchKernelGGL((adj_batch_kernel<PP, FF, ADJ_VS>), dim3(g), dim3(256 * ADJ_VS), 0, s, \
                         t, sino, im, epi, pred, pstride);                                 \
  } while (0)
  if (t.ppt == 2) {
    if (t.fract)
      SCXR_ADJB(2, true);
    else
      SCXR_ADJB(2, false);
  } else {
    if (t.fract)
      SCXR_ADJB(1, true);
    else
      SCXR_ADJB(1, false);
  }
#undef SCXR_ADJB
  SCXR_LAUNCHED("xr2d batch adjoint");
  return SCXR_OK;
}

int forward(const Plan& p, const float* im, float* sino, const int32_t* pred, hipStream_t s,
            const float* yres, double* res_ws, const float* imT) {
  if (p.fwd_generic) {
    SCXR_REQUIRE(!res_ws, SCXR_ERR_UNSUPPORTED,
                 "residual partials are not available for this geometry (general-geometry "
                 "forward)");
    hipLaunchKernelGGL(fwd_generic_kernel, dim3(ceil_div(p.D, 256), p.A), dim3(256), 0, s, im,
                       sino, p.geom, p.rw, p.M, p.N, p.D, pred);
    SCXR_LAUNCHED("xr2d general forward");
    return SCXR_OK;
  }
  // column-fast views stream the transposed image: the caller's, or the
  // plan's workspace filled here
  const float* srcT = imT ? imT : p.work;
  if (p.any_colfast && !imT) {
    dim3 g(ceil_div(p.N, 32), ceil_div(p.M, 32));
    hipLaunchKernelGGL(transpose_kernel, g, dim3(256), 0, s, im, p.work, p.M, p.N, pred);
    SCXR_LAUNCHED("xr2d transpose");
  }
  const size_t lds = fwd_lds_bytes(p.vpg, p.D, p.fwd_ch);
  const dim3 fgrid(p.n_groups, p.fwd_sp);
#define SCXR_FWD(VV, CC)                                                                      \
  do {                                                                                         \
    if (p.adj_fract)                                                                           \
      hipLaunchKernelGGL((fwd_kernel<VV, CC, true>), fgrid, dim3(64 * FWD_WAVES), lds, s, im,  \
                         srcT, sino, p.geom, p.rw, p.groups, p.order, p.M, p.N, p.D, yres,    \
                         res_ws, p.parts, p.A, p.gblock, pred);                               \
    else                                                                                       \
      hipLaunchKernelGGL((fwd_kernel<VV, CC, false>), fgrid, dim3(64 * FWD_WAVES), lds, s, im, \
                         srcT, sino, p.geom, p.rw, p.groups, p.order, p.M, p.N, p.D, yres,    \
                         res_ws, p.parts, p.A, p.gblock, pred);                               \
  } while (0)
  if (p.fwd_ch == 2) {
    if (p.vpg == 1)
      SCXR_FWD(1, 2);
    else if (p.vpg == 2)
      SCXR_FWD(2, 2);
    else
      SCXR_FWD(4, 2);
  } else {
    if (p.vpg == 1)
      SCXR_FWD(1, 1);
    else if (p.vpg == 2)
      SCXR_FWD(2, 1);
    else
      SCXR_FWD(4, 1);
  }
#undef SCXR_FWD
  SCXR_LAUNCHED("xr2d forward");
  if (p.fwd_sp > 1) {
    hipLaunchKernelGGL(fwd_parts_kernel, dim3(p.n_groups), dim3(256), 0, s, p.parts, p.fwd_sp,
                       p.A, p.D, sino, p.groups, p.order, yres, res_ws, pred);
    SCXR_LAUNCHED("xr2d forward part sums");
  }
  return SCXR_OK;
}

int adjoint(const Plan& p, const float* sino, float* im, const int32_t* pred, hipStream_t s,
            const AdjEpi& epi) {
  if (p.adj_generic) {
    SCXR_REQUIRE(epi.kind == EPI_NONE && !epi.dot_ws, SCXR_ERR_UNSUPPORTED,
                 "adjoint epilogues are not available for this geometry (general-geometry "
                 "adjoint)");
    const long long n = static_cast<long long>(p.M) * p.N;
    hipLaunchKernelGGL(adj_generic_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, sino, im,
                       p.geom, p.rw, p.M, p.N, p.A, p.D, pred);
    SCXR_LAUNCHED("xr2d general adjoint");
    return SCXR_OK;
  }
  dim3 g(ceil_div(p.N, ADJ_TJ * p.adj_ppt), ceil_div(p.M, ADJ_TI));
#define SCXR_ADJ(PP, FF)                                                                    \
  hipLaunchKernelGGL((adj_kernel<PP, FF>), g, dim3(256 * ADJ_VS), 0, s, sino, im, p.geom, p.rw,  \
                     p.clo, p.M, p.N, p.A, p.D, p.adj_w, epi, pred)
  if (p.adj_ppt == 2) {
    if (p.adj_fract)
      SCXR_ADJ(2, true);
    else
      SCXR_ADJ(2, false);
  } else {
    if (p.adj_fract)
      SCXR_ADJ(1, true);
    else
      SCXR_ADJ(1, false);
  }
#undef SCXR_ADJ
  SCXR_LAUNCHED("xr2d adjoint");
  return SCXR_OK;
}

}  // namespace xr2d
}  // namespace scxr

// ------------------------------------------------------------- C ABI --

using namespace scxr::xr2d;

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
  return xr2d::forward(plan->p, im, sino, pred, as_stream(stream), nullptr, nullptr);
}

int scxr_xr2d_forward_t(const scxr_xr2d_plan* plan, const float* im, const float* imT,
                        float* sino, const int32_t* pred, void* stream) {
  SCXR_REQUIRE(plan && im && imT && sino, SCXR_ERR_ARG, "null argument to xr2d_forward_t");
  return xr2d::forward(plan->p, im, sino, pred, as_stream(stream), nullptr, nullptr, imT);
}

int scxr_xr2d_forward_res(const scxr_xr2d_plan* plan, const float* im, float* sino,
                          const float* y, double* res_ws, const int32_t* pred, void* stream) {
  SCXR_REQUIRE(plan && im && sino && y && res_ws, SCXR_ERR_ARG,
               "null argument to xr2d_forward_res");
  return xr2d::forward(plan->p, im, sino, pred, as_stream(stream), y, res_ws);
}

int scxr_xr2d_adjoint(const scxr_xr2d_plan* plan, const float* sino, float* im,
                      const int32_t* pred, void* stream) {
  SCXR_REQUIRE(plan && im && sino, SCXR_ERR_ARG, "null argument to xr2d_adjoint");
  return xr2d::adjoint(plan->p, sino, im, pred, as_stream(stream), AdjEpi{});
}

int scxr_xr2d_adjoint_epi(const scxr_xr2d_plan* plan, const float* sino, float* im,
                          int32_t kind, float rho, int32_t circular, const float* p,
                          const float* z, const float* u, double* dot_ws, const int32_t* pred,
                          void* stream) {
  SCXR_REQUIRE(plan && im && sino, SCXR_ERR_ARG, "null argument to xr2d_adjoint_epi");
  SCXR_REQUIRE(kind >= EPI_NONE && kind <= EPI_RHS, SCXR_ERR_ARG, "bad epilogue kind %d", kind);
  SCXR_REQUIRE(kind != EPI_GRAM || p, SCXR_ERR_ARG, "gram epilogue needs p");
  SCXR_REQUIRE(kind != EPI_RHS || z, SCXR_ERR_ARG, "rhs epilogue needs z");
  SCXR_REQUIRE(!dot_ws || p, SCXR_ERR_ARG, "dot partials need p");
  AdjEpi e;
  e.kind = kind;
  e.circular = circular;
  e.rho = rho;
  e.p = p;
  e.z = z;
  e.u = u;
  e.dot_ws = dot_ws;
  return xr2d::adjoint(plan->p, sino, im, pred, as_stream(stream), e);
}

int scxr_xr2d_set_forward_split(scxr_xr2d_plan* plan, int32_t parts, int32_t* effective) {
  SCXR_REQUIRE(plan && parts >= 0 && parts <= FWD_SP_MAX, SCXR_ERR_ARG,
               "forward split must be 0 (automatic) or 1..%d parts", FWD_SP_MAX);
  auto& p = plan->p;
  const int rc = xr2d::set_split(p, parts == 0 ? xr2d::auto_split(p) : parts);
  if (effective) *effective = p.fwd_sp;
  return rc;
}

int scxr_xr2d_info(const scxr_xr2d_plan* plan, int32_t* out) {
  SCXR_REQUIRE(plan && out, SCXR_ERR_ARG, "null argument to xr2d_info");
  const auto& p = plan->p;
  out[0] = p.fwd_generic ? 0 : p.n_groups;
  out[1] = p.adj_generic
               ? 0
               : static_cast<int32_t>(ceil_div(p.N, ADJ_TJ * p.adj_ppt) * ceil_div(p.M, ADJ_TI));
  out[2] = p.vpg;
  out[3] = p.adj_w;
  return SCXR_OK;
}

int scxr_ramp_filter(int32_t A, int32_t D, int32_t Dout, int32_t shift, const float* taps,
                     const float* y, float* out, const int32_t* pred, void* stream) {
  SCXR_REQUIRE(A > 0 && D > 0 && D <= 4096 && Dout > 0 && taps && y && out, SCXR_ERR_ARG,
               "bad ramp_filter arguments (D must be in [1, 4096])");
  hipLaunchKernelGGL(ramp_kernel, dim3(A), dim3(256), 0, as_stream(stream), y, out, taps, D, Dout,
                     shift, pred);
  SCXR_LAUNCHED("ramp filter");
  return SCXR_OK;
}

int scxr_xr2d_weights(const scxr_xr2d_plan* plan, int32_t* inds, float* weights, void* stream) {
  SCXR_REQUIRE(plan && inds && weights, SCXR_ERR_ARG, "null argument to xr2d_weights");
  const auto& p = plan->p;
  const size_t total = static_cast<size_t>(p.M) * p.N * p.A;
  SCXR_REQUIRE(total < (static_cast<size_t>(1) << 40), SCXR_ERR_SHAPE, "too many entries");
  hipLaunchKernelGGL(xr2d::weights_kernel, dim3(ceil_div(static_cast<long long>(total), 256)),
                     dim3(256), 0, as_stream(stream), inds, weights, p.geom, p.rw, p.M, p.N,
                     p.A, static_cast<int>(p.adj_fract));
  SCXR_LAUNCHED("xr2d weights");
  return SCXR_OK;
}

}  // extern "C"
