// Parallel-beam 3D X-ray projector (scico.linop.xray.XRayTransform3D) for
// gfx950.
//
// Model (scico/linop/xray/_xray3d.py:206-266): voxel (i, j, k) has centre
// x = (i + 1/2, j + 1/2, k + 1/2); for view v with float32 matrix m (2 x 4)
//   Px_r = ((m[r][0] x0 + m[r][1] x1) + m[r][2] x2) + m[r][3]
//   left = Px - 1/4,  t = min(ceil(left) - left, 1/2),  ul = floor(left)
// and the voxel splats onto the 2 x 2 detector pixels at ul with weights
// (t0 t1) 4, ((1/2 - t0) t1) 4, (t0 (1/2 - t1)) 4, ((1/2 - t0)(1/2 - t1)) 4,
// each zeroed when its pixel is off the detector.
//
// Slice mode.  When one detector axis d depends only on one volume axis s
// (the rotation axis) through the same coefficient and offset for every
// view, and that axis maps slice n exactly onto detector column n + delta
// with t_d = 1/2, the other two weights of every voxel are exactly zero and
// the transform splits into independent 2D slices that share one in-plane
// weight table.  The kernels then run with the SLICE index across the
// lanes: a wave owns 64 lanes x 4 consecutive slices, computes each
// (pixel, view) weight once in one lane, broadcasts it with readlane and
// applies it to 256 slices with float4 loads of contiguous slice rows.
// Volume and projections are used in a canonical layout with the slice
// index fastest: vol_c[p][q][s] and proj_c[v][u][s] (p < q the in-plane
// volume axes, u the in-plane detector axis); a tiled transpose converts
// to and from the caller's layout when it differs.
//
// General mode (any other geometry): the adjoint gathers per voxel in the
// reference's per-view order; the forward scatters with float atomics.

#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"

namespace scxr {
namespace xr3d {

enum { MODE_GENERAL = 0, MODE_SLICE = 1 };

constexpr int FB = 8;   // forward: detector bins per wave
constexpr int AR = 8;   // adjoint: pixels per wave (2 x 4 block)

struct Plan {
  int X = 0, Y = 0, Z = 0, V = 0, D0 = 0, D1 = 0;
  int mode = MODE_GENERAL;
  int sax = -1, dax = -1, delta = 0, pax = -1, qax = -1;
  int P = 0, Q = 0, S = 0, SQ = 0, U = 0;  // SQ = slice quads (slices padded to 4)
  int cpr = 32;                            // forward candidates per row
  int spg = 2;                             // forward slots per bin group
  bool vol_direct = false, proj_direct = false;
  bool adj_staged = false;                 // adjoint rows staged in LDS (adj_slice_staged_kernel)
  float* mats = nullptr;   // (V, 2, 4) float32
  float4* geom = nullptr;  // slice mode: (a, b, t, 0) in-plane row per view
  float* volc = nullptr;   // canonical volume workspace (P Q rows of 4 SQ floats)
  float* proja = nullptr;  // canonical projection workspace (V U 4 SQ)
  struct GenView* gview = nullptr;  // general mode: per-view gather set-up
};

// General-mode forward set-up of one view.  The detector pixel (d0, d1)
// receives from the voxels whose projected centre P satisfies
// floor(P_r - 1/4) in {d_r - 1, d_r}, i.e. P_r in [d_r - 3/4, d_r + 5/4).
// The kernel marches along volume axis a (the one whose complementary
// 2 x 2 minor of the matrix is largest) and, per plane, solves for the
// other two coordinates (b, c) of the box centre with the inverse minor;
// the candidates are the nb x nc integer points of the box's bounding
// rectangle (half-extents hb, hc plus a rounding margin).
struct GenView {
  int a, nb, nc, pad;
  float i00, i01, i10, i11, hb, hc, pad2, pad3;
};

// ------------------------------------------------------------ weights --

struct Tap {
  int ul;
  float wl, wr;
};

// In-plane weights of pixel (p, q) for a slice-mode view with in-plane row
// (a, b, t): P = (a (p + 1/2) + b (q + 1/2)) + t, the two taps at ul and
// ul + 1 carry (t 1/2) 4 and ((1/2 - t) 1/2) 4 (the slice axis has
// t_d = 1/2), zeroed off the detector.
__device__ __forceinline__ Tap slice_tap(float a, float b, float t, int p, int q, int U) {
  const float xp = __fadd_rn(static_cast<float>(p), 0.5f);
  const float xq = __fadd_rn(static_cast<float>(q), 0.5f);
  const float P = __fadd_rn(__fadd_rn(__fmul_rn(a, xp), __fmul_rn(b, xq)), t);
  const float left = __fsub_rn(P, 0.25f);
  const float tt = fminf(__fsub_rn(ceilf(left), left), 0.5f);
  Tap r;
  r.ul = static_cast<int>(floorf(left));
  r.wl = __fmul_rn(__fmul_rn(tt, 0.5f), 4.0f);
  r.wr = __fmul_rn(__fmul_rn(__fsub_rn(0.5f, tt), 0.5f), 4.0f);
  if (!(r.ul >= 0 && r.ul < U)) r.wl = 0.0f;
  if (!(r.ul + 1 >= 0 && r.ul + 1 < U)) r.wr = 0.0f;
  return r;
}

// Full reference weights of voxel (i, j, k) for matrix m (general mode).
__device__ __forceinline__ void gen_weights(const float* __restrict__ m, int i, int j, int k,
                                            int D0, int D1, int ul[2], float w[4]) {
  const float x0 = __fadd_rn(static_cast<float>(i), 0.5f);
  const float x1 = __fadd_rn(static_cast<float>(j), 0.5f);
  const float x2 = __fadd_rn(static_cast<float>(k), 0.5f);
  float t[2];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const float* mr = m + 4 * r;
    const float P = __fadd_rn(
        __fadd_rn(__fadd_rn(__fmul_rn(mr[0], x0), __fmul_rn(mr[1], x1)), __fmul_rn(mr[2], x2)),
        mr[3]);
    const float left = __fsub_rn(P, 0.25f);
    t[r] = fminf(__fsub_rn(ceilf(left), left), 0.5f);
    ul[r] = static_cast<int>(floorf(left));
  }
  const float s0 = __fsub_rn(0.5f, t[0]), s1 = __fsub_rn(0.5f, t[1]);
  w[0] = __fmul_rn(__fmul_rn(t[0], t[1]), 4.0f);
  w[1] = __fmul_rn(__fmul_rn(s0, t[1]), 4.0f);
  w[2] = __fmul_rn(__fmul_rn(t[0], s1), 4.0f);
  w[3] = __fmul_rn(__fmul_rn(s0, s1), 4.0f);
  const bool a0 = ul[0] >= 0 && ul[0] < D0, b0 = ul[0] + 1 >= 0 && ul[0] + 1 < D0;
  const bool a1 = ul[1] >= 0 && ul[1] < D1, b1 = ul[1] + 1 >= 0 && ul[1] + 1 < D1;
  if (!(a0 && a1)) w[0] = 0.0f;
  if (!(b0 && a1)) w[1] = 0.0f;
  if (!(a0 && b1)) w[2] = 0.0f;
  if (!(b0 && b1)) w[3] = 0.0f;
}

__device__ __forceinline__ float4 fma4(float w, float4 x, float4 acc) {
  return make_float4(__builtin_fmaf(w, x.x, acc.x), __builtin_fmaf(w, x.y, acc.y),
                     __builtin_fmaf(w, x.z, acc.z), __builtin_fmaf(w, x.w, acc.w));
}

__device__ __forceinline__ int rl(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }


__device__ __forceinline__ float rlf(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// ------------------------------------------------------ slice kernels --

// Forward: each wave owns FB consecutive bins of one view for 256 slices
// and gathers the strip of pixels that touch them, row by row along the
// slow in-plane axis.  Lane kk of a row pass holds candidate kk of the
// row, ordered so that its bin is non-decreasing; ballots give the first
// candidate of every bin group g (ul = u0 + g, g = -1 .. FB - 1).  Each
// group has SPG static slots (a group holds at most SPG pixels: pixel
// centres are at least |f| apart along the fast axis), so the whole row is
// straight-line code apart from wave-uniform branches: every occupied
// slot's row of slices is loaded independently and accumulated into
// static registers.
template <int CPR, int SPG>
__global__ __launch_bounds__(256) void fwd_slice_kernel(const float4* __restrict__ vol,
                                                        float4* __restrict__ proj,
                                                        const float4* __restrict__ geom, int P,
                                                        int Q, int SQ, int U, int V,
                                                        const int32_t* __restrict__ pred) {
  if (pred_off(pred)) return;
  constexpr int RPP = 64 / CPR;
  constexpr int BIG = INT_MAX / 2;
  const int lane = threadIdx.x & 63;
  // Workgroup order: consecutive workgroups go to the 8 XCDs in turn; each
  // XCD is given a contiguous run of an order that takes the view groups
  // in blocks of about an eighth of them, every bin group of a block with
  // its views adjacent.  Each XCD then works on its own range of views
  // (neighbouring views' strips share most of their pixel rows, so they
  // share its L2) across all bin groups (whose strips differ in length,
  // so every XCD gets a similar share of work).  A bijection of the grid.
  int bg, vg;
  {
    const int GX = gridDim.x, GY = gridDim.y, G = GX * GY, G8 = (G / 8) * 8;
    int L = blockIdx.y * GX + blockIdx.x;
    if (L < G8) L = (L % 8) * (G8 / 8) + L / 8;
    const int VB = GY >= 8 ? GY / 8 : 1;  // view groups per block
    const int blk = L / (GX * VB);
    const int nb = min(VB, GY - blk * VB);  // the last block may be short
    const int within = L - blk * GX * VB;
    bg = within / nb;
    vg = blk * VB + (within - bg * nb);
  }
  const int v = vg * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (v >= V) return;
  const int u0 = bg * FB;
  const unsigned quad = blockIdx.z * 64 + lane;
  const bool qok = quad < static_cast<unsigned>(SQ);
  const unsigned qoff = (qok ? quad : 0u) * 16u;
  // vol holds P Q rows of SQ slice quads
  const int rowb = SQ * 16;
  const __amdgpu_buffer_rsrc_t vr = make_rsrc(vol, P * Q * rowb);
  const float4 g = geom[v];
  const float a = g.x, b = g.y, t = g.z;
  const bool qfast = fabsf(b) >= fabsf(a);
  const float f = qfast ? b : a, sc = qfast ? a : b;
  const int nslow = qfast ? P : Q, nfast = qfast ? Q : P;
  // pixel index step along the fast axis, in candidate order
  const int kstep = (f > 0.0f ? 1 : -1) * (qfast ? 1 : Q);
  const float plo = static_cast<float>(u0) - 0.75f;
  const float phi = static_cast<float>(u0 + FB) + 0.25f;
  const float rf = 1.0f / f;
  const int kk = lane % CPR;

  float4 acc[FB];
#pragma unroll
  for (int j = 0; j < FB; ++j) acc[j] = make_float4(0.f, 0.f, 0.f, 0.f);

  for (int r0 = 0; r0 < nslow; r0 += RPP) {
    const int row = r0 + lane / CPR;
    // Approximate fast-index interval of the strip in this row, widened by
    // one pixel each side; the exact bin of each candidate decides.
    const float c = sc * (static_cast<float>(row) + 0.5f) + t;
    const float ka = (plo - c) * rf - 0.5f, kb = (phi - c) * rf - 0.5f;
    const int klo = max(0, static_cast<int>(floorf(fminf(ka, kb))) - 1);
    const int khi = min(nfast - 1, static_cast<int>(ceilf(fmaxf(ka, kb))) + 1);
    const int k0 = (f > 0.0f) ? klo : khi;  // fast index of candidate 0
    const int pix0 = qfast ? row * Q + k0 : k0 * Q + row;
    int gi = BIG;
    float wl = 0.0f, wr = 0.0f;
    if (row < nslow && kk <= khi - klo) {
      const int k = (f > 0.0f) ? klo + kk : khi - kk;
      const int p = qfast ? row : k, q = qfast ? k : row;
      const Tap tp = slice_tap(a, b, t, p, q, U);
      gi = tp.ul - u0;
      wl = tp.wl;
      wr = tp.wr;
    }
#pragma unroll
    for (int rr = 0; rr < RPP; ++rr) {
      if (r0 + rr >= nslow) break;
      const int base = rr * CPR;
      const int rpix = rl(pix0, base);
      int first[FB + 2];
#pragma unroll
      for (int gg = -1; gg <= FB; ++gg) {
        const unsigned long long bal = __ballot(gi < gg);
        int cnt;
        if (CPR == 64) {
          cnt = __popcll(bal);
        } else {
          const unsigned part = static_cast<unsigned>(bal >> (rr * CPR)) &
                                (CPR == 32 ? 0xffffffffu : ((1u << CPR) - 1u));
          cnt = __popc(part);
        }
        first[gg + 1] = base + cnt;
      }
      // Groups in chunks of GC: the loads of a chunk's occupied slots are
      // issued first, then those slots are accumulated.  Emptiness is
      // wave-uniform, so empty slots are skipped by scalar branches: they
      // issue no load (an out-of-range load would still take its cycles
      // in the texture addresser, which bounds this kernel).
      constexpr int NG = FB + 1, GC = 5;
      // byte offset of slot L: (rpix + (L - base) kstep) rowb, formed once
      // per group and stepped by kstep rowb from slot to slot
      const int rbyte = rpix * rowb, kbyte = kstep * rowb;
#pragma unroll
      for (int g0 = 0; g0 < NG; g0 += GC) {
        float4 xs[GC][SPG];
#pragma unroll
        for (int c = 0; c < GC; ++c) {
          const int gg = g0 + c - 1;
          if (gg < FB) {
            const int gbyte = rbyte + (first[gg + 1] - base) * kbyte;
#pragma unroll
            for (int k = 0; k < SPG; ++k) {
              const int L = first[gg + 1] + k;
              // an empty slot's register is never read
              if (L < first[gg + 2]) xs[c][k] = bload4(vr, qoff, gbyte + k * kbyte);
            }
          }
        }
#pragma unroll
        for (int c = 0; c < GC; ++c) {
          const int gg = g0 + c - 1;
          if (gg < FB) {
#pragma unroll
            for (int k = 0; k < SPG; ++k) {
              const int L = first[gg + 1] + k;
              if (L < first[gg + 2]) {
                const float cl = rlf(wl, L), cr = rlf(wr, L);
                if (gg >= 0) acc[gg] = fma4(cl, xs[c][k], acc[gg]);
                if (gg + 1 < FB) acc[gg + 1] = fma4(cr, xs[c][k], acc[gg + 1]);
              }
            }
          }
        }
      }
    }
  }
  if (!qok) return;
#pragma unroll
  for (int j = 0; j < FB; ++j)
    if (u0 + j < U) proj[(static_cast<size_t>(v) * U + u0 + j) * SQ + quad] = acc[j];
}

// Adjoint: each wave owns a 2 x 4 block of pixels for 256 slices and walks
// the views in order, 8 at a time: lane (view j, pixel r) computes that
// pair's taps, then every (view, pixel) pair is broadcast and applied as
// acc += wl * proj_c[v][ul] + wr * proj_c[v][ul + 1] (float4 per lane).
__global__ __launch_bounds__(256) void adj_slice_kernel(const float4* __restrict__ proj,
                                                        float4* __restrict__ vol,
                                                        const float4* __restrict__ geom, int P,
                                                        int Q, int SQ, int U, int V,
                                                        const int32_t* __restrict__ pred) {
  if (pred_off(pred)) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int p0 = blockIdx.y * 4 + 2 * (w >> 1), q0 = blockIdx.x * 8 + 4 * (w & 1);
  const int quad = blockIdx.z * 64 + lane;
  const bool qok = quad < SQ;
  const unsigned qoff = (qok ? quad : 0) * 16u;
  const int rowb = SQ * 16;
  const __amdgpu_buffer_rsrc_t yr = make_rsrc(proj, V * U * rowb);
  const int r = lane & (AR - 1), jv = lane / AR;
  const int pp = p0 + (r >> 2), qq = q0 + (r & 3);
  const bool pok = pp < P && qq < Q;

  float4 acc[AR];
#pragma unroll
  for (int i = 0; i < AR; ++i) acc[i] = make_float4(0.f, 0.f, 0.f, 0.f);

  for (int v0 = 0; v0 < V; v0 += 64 / AR) {
    const int vv = v0 + jv;
    int ul = 0;
    float wl = 0.0f, wr = 0.0f;
    if (vv < V && pok) {
      const float4 g = geom[vv];
      const Tap tp = slice_tap(g.x, g.y, g.z, pp, qq, U);
      ul = tp.ul;
      wl = tp.wl;
      wr = tp.wr;
    }
    const int nv = min(64 / AR, V - v0);
    for (int j = 0; j < nv; ++j) {
      const int vb = (v0 + j) * U;
      // The bin pair (u, u + 1) is uniform across the wave; consecutive
      // pixels of a row move it by at most one bin, so rows already loaded
      // for the previous pixel are reused (scalar branches).
      auto row = [&](int u) { return bload4(yr, qoff, (vb + min(max(u, 0), U - 1)) * rowb); };
      int ucur = INT_MIN / 2;
      float4 ya = make_float4(0.f, 0.f, 0.f, 0.f), yb = ya;
#pragma unroll
      for (int i = 0; i < AR; ++i) {
        const int L = j * AR + i;
        const int u = rl(ul, L);
        const float cl = rlf(wl, L), cr = rlf(wr, L);
        if (u == ucur + 1) {
          ya = yb;
          yb = row(u + 1);
        } else if (u == ucur - 1) {
          yb = ya;
          ya = row(u);
        } else if (u != ucur) {
          ya = row(u);
          yb = row(u + 1);
        }
        ucur = u;
        acc[i] = fma4(cr, yb, fma4(cl, ya, acc[i]));
      }
    }
  }
  if (!qok) return;
#pragma unroll
  for (int i = 0; i < AR; ++i) {
    const int p = p0 + (i >> 2), q = q0 + (i & 3);
    if (p < P && q < Q) vol[(static_cast<size_t>(p) * Q + q) * SQ + quad] = acc[i];
  }
}

// Copy one 64-quad row of slices (1 KiB per wave) from global memory into
// LDS at dst without passing through registers (lane i writes dst + 16 i).
__device__ __forceinline__ void lds_row_load(__amdgpu_buffer_rsrc_t r, float4* dst, unsigned voff,
                                             int soff) {
  __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)dst, 16,
                                           voff, soff, 0, 0);
}
// Adjoint with the detector rows staged in LDS.  The workgroup's four
// waves cover a 4 x 8 pixel tile; for each view the tile's bins lie
// between those of its corner pixels (the projected centre is monotone in
// p and q, in float32 too), a range of at most ceil(3 |a| + 7 |b|) + 2
// rows.  That range is copied into LDS once per view (buffer loads that
// write LDS directly, one view ahead), and each wave reads its pixels'
// rows from there with the same reuse and accumulation order as
// adj_slice_kernel, so the results are bitwise the same; the texture data
// path then moves each row once per workgroup instead of once per wave.
// The host takes this kernel when every view's range fits ADJ_NW rows.
constexpr int ADJ_NW = 12;  // staged detector rows per view

__global__ __launch_bounds__(256) void adj_slice_staged_kernel(
    const float4* __restrict__ proj, float4* __restrict__ vol, const float4* __restrict__ geom,
    int P, int Q, int SQ, int U, int V, const int32_t* __restrict__ pred) {
  if (pred_off(pred)) return;
  __shared__ float4 stage[2][ADJ_NW][64];
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int pt = blockIdx.y * 4, qt = blockIdx.x * 8;  // tile origin
  const int p0 = pt + 2 * (w >> 1), q0 = qt + 4 * (w & 1);
  const int quad = blockIdx.z * 64 + lane;
  const bool qok = quad < SQ;
  const unsigned qoff = (qok ? quad : 0) * 16u;
  const int rowb = SQ * 16;
  const __amdgpu_buffer_rsrc_t yr = make_rsrc(proj, V * U * rowb);
  const int r = lane & (AR - 1), jv = lane / AR;
  const int pp = p0 + (r >> 2), qq = q0 + (r & 3);
  const bool pok = pp < P && qq < Q;

  // Staged range (first row, row count) of views c0 .. c0 + 15: lane
  // 4 k + corner computes corner `corner` of view c0 + k; a quad reduction
  // leaves the range in lane 4 k.
  auto ranges = [&](int c0, int& lo, int& n) {
    const int vv = min(c0 + (lane >> 2), V - 1), c = lane & 3;
    const float4 g = geom[vv];
    const int ul = slice_tap(g.x, g.y, g.z, pt + 3 * (c >> 1), qt + 7 * (c & 1), U).ul;
    int mn = ul, mx = ul;
    mn = min(mn, __builtin_amdgcn_mov_dpp(mn, 0xB1, 0xF, 0xF, true));  // quad_perm 1 0 3 2
    mx = max(mx, __builtin_amdgcn_mov_dpp(mx, 0xB1, 0xF, 0xF, true));
    mn = min(mn, __builtin_amdgcn_mov_dpp(mn, 0x4E, 0xF, 0xF, true));  // quad_perm 2 3 0 1
    mx = max(mx, __builtin_amdgcn_mov_dpp(mx, 0x4E, 0xF, 0xF, true));
    lo = min(max(mn, 0), U - 1);
    n = min(max(mx + 1, 0), U - 1) - lo + 1;
  };
  int lo_c, n_c, lo_n = 0, n_n = 0;  // ranges of the current and the next 16 views
  ranges(0, lo_c, n_c);
  if (V > 16) ranges(16, lo_n, n_n);
  // stage view v's rows into buffer b
  auto stage_view = [&](int v, int lo, int n, int b) {
    for (int k = w; k < n; k += 4)
      lds_row_load(yr, &stage[b][k][0], qoff, (v * U + lo + k) * rowb);
  };

  float4 acc[AR];
#pragma unroll
  for (int i = 0; i < AR; ++i) acc[i] = make_float4(0.f, 0.f, 0.f, 0.f);

  stage_view(0, rl(lo_c, 0), rl(n_c, 0), 0);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  int cab = 0;  // the pixel's two detector rows, clamped as adj_slice_kernel
                // clamps them, packed as two 16-bit fields (U < 65536)
  float wl = 0.0f, wr = 0.0f;
#pragma unroll 1
  for (int v = 0; v < V; ++v) {
    const int j = v & 7;
    if (j == 0) {  // taps of the wave's 8 pixels for views v .. v + 7
      const int vv = v + jv;
      int ul = 0;
      wl = wr = 0.0f;
      if (vv < V) {
        const float4 g = geom[vv];
        const Tap tp = slice_tap(g.x, g.y, g.z, pp, qq, U);
        ul = tp.ul;
        wl = pok ? tp.wl : 0.0f;
        wr = pok ? tp.wr : 0.0f;
      }
      cab = min(max(ul, 0), U - 1) | (min(max(ul + 1, 0), U - 1) << 16);
    }
    // this view's range, then the next view's, from the 16-view range sets
    const int lo = rl(lo_c, 4 * (v & 15));
    if (v + 1 < V) {
      const int vn = v + 1, kn = 4 * (vn & 15);
      const bool nxt_set = (vn & 15) == 0;
      const int lon = nxt_set ? rl(lo_n, kn) : rl(lo_c, kn);
      const int nn = nxt_set ? rl(n_n, kn) : rl(n_c, kn);
      stage_view(vn, lon, nn, vn & 1);
    }
    // both rows of all eight pixels first (no reuse branches: LDS reads
    // are cheap), then the accumulation in adj_slice_kernel's order
    const char* sb = reinterpret_cast<const char*>(&stage[v & 1][0][0] + lane);
    // byte offsets of both staged rows (64 float4 = 1 KiB apart): every
    // pixel's clamped rows lie in [lo, lo + n), n <= ADJ_NW (the range is
    // taken over the tile's corner bins), so neither 16-bit field borrows
    // and each offset stays below 2^14
    const unsigned rab = static_cast<unsigned>(cab - (lo | (lo << 16))) << 10;
    float4 ya[AR], yb[AR];
#pragma unroll
    for (int i = 0; i < AR; ++i) {
      const unsigned x = static_cast<unsigned>(rl(static_cast<int>(rab), j * AR + i));
      ya[i] = *reinterpret_cast<const float4*>(sb + (x & 0xffffu));
      yb[i] = *reinterpret_cast<const float4*>(sb + (x >> 16));
    }
#pragma unroll
    for (int i = 0; i < AR; ++i) {
      const int L = j * AR + i;
      const float cl = rlf(wl, L), cr = rlf(wr, L);
      acc[i] = fma4(cr, yb[i], fma4(cl, ya[i], acc[i]));
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if ((v & 15) == 15) {  // the next 16 views become current
      lo_c = lo_n;
      n_c = n_n;
      if (v + 17 < V) ranges(v + 17, lo_n, n_n);
    }
  }
  if (!qok) return;
#pragma unroll
  for (int i = 0; i < AR; ++i) {
    const int p = p0 + (i >> 2), q = q0 + (i & 3);
    if (p < P && q < Q) vol[(static_cast<size_t>(p) * Q + q) * SQ + quad] = acc[i];
  }
}

// ----------------------------------------------------- layout kernels --

// Batched 2D transpose with offsets and zero fill:
//   out[b ob + c oc + r] = in[b ib + (r + roff) ir + (c + coff)]
// when 0 <= r + roff < Rin and 0 <= c + coff < Cin, else 0, for
// r < Rout, c < Cout.  32 x 32 tiles through LDS.
__global__ __launch_bounds__(256) void transpose_kernel(const float* __restrict__ in,
                                                        float* __restrict__ out, int Rout,
                                                        int Cout, int Rin, int Cin, int roff,
                                                        int coff, long long ib, long long ir,
                                                        long long ob, long long oc,
                                                        const int32_t* __restrict__ pred) {
  if (pred_off(pred)) return;
  __shared__ float tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  const long long b = blockIdx.z;
  for (int rr = ty; rr < 32; rr += 8) {
    const int ri = r0 + rr + roff, ci = c0 + tx + coff;
    float val = 0.0f;
    if (ri >= 0 && ri < Rin && ci >= 0 && ci < Cin) val = in[b * ib + ri * ir + ci];
    tile[rr][tx] = val;
  }
  __syncthreads();
  for (int cc = ty; cc < 32; cc += 8) {
    const int c = c0 + cc, r = r0 + tx;
    if (c < Cout && r < Rout) out[b * ob + c * oc + r] = tile[tx][cc];
  }
}

// Row copy with a column shift and zero fill:
//   out[row os + s] = in[row is + s + off] if 0 <= s + off < Cin else 0.
__global__ __launch_bounds__(256) void shift_rows_kernel(const float* __restrict__ in,
                                                         float* __restrict__ out, long long rows,
                                                         int Cout, int Cin, int off, long long is,
                                                         long long os,
                                                         const int32_t* __restrict__ pred) {
  if (pred_off(pred)) return;
  const long long n = rows * Cout;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < n; e += 256LL * gridDim.x) {
    const long long row = e / Cout;
    const int s = static_cast<int>(e - row * Cout);
    const int c = s + off;
    out[row * os + s] = (c >= 0 && c < Cin) ? in[row * is + c] : 0.0f;
  }
}

// ------------------------------------------------- general-mode kernels --

__global__ __launch_bounds__(256) void gen_adj_kernel(const float* __restrict__ proj,
                                                      float* __restrict__ vol,
                                                      const float* __restrict__ mats, int X,
                                                      int Y, int Z, int V, int D0, int D1,
                                                      const int32_t* __restrict__ pred) {
  if (pred_off(pred)) return;
  const long long n = static_cast<long long>(X) * Y * Z;
  const long long e = blockIdx.x * 256LL + threadIdx.x;
  if (e >= n) return;
  const int k = static_cast<int>(e % Z), j = static_cast<int>((e / Z) % Y),
            i = static_cast<int>(e / (static_cast<long long>(Y) * Z));
  float acc = 0.0f;
  for (int v = 0; v < V; ++v) {
    int ul[2];
    float w[4];
    gen_weights(mats + 8 * v, i, j, k, D0, D1, ul, w);
    const float* y = proj + static_cast<size_t>(v) * D0 * D1;
    const int a0 = min(max(ul[0], 0), D0 - 1), b0 = min(max(ul[0] + 1, 0), D0 - 1);
    const int a1 = min(max(ul[1], 0), D1 - 1), b1 = min(max(ul[1] + 1, 0), D1 - 1);
    // _xray3d.py:200-203, in order
    acc = __fadd_rn(acc, __fmul_rn(y[a0 * D1 + a1], w[0]));
    acc = __fadd_rn(acc, __fmul_rn(y[b0 * D1 + a1], w[1]));
    acc = __fadd_rn(acc, __fmul_rn(y[a0 * D1 + b1], w[2]));
    acc = __fadd_rn(acc, __fmul_rn(y[b0 * D1 + b1], w[3]));
  }
  vol[e] = acc;
}

// General-mode forward as a gather: one thread per (view, detector pixel)
// sums w * vol over the voxels whose reference taps land on the pixel, the
// weights evaluated per voxel exactly as gen_weights does.  Deterministic,
// no atomics.
template <int A>
__device__ __forceinline__ float gen_gather(const float* __restrict__ vol,
                                            const float* __restrict__ m, const GenView& g,
                                            int X, int Y, int Z, int d0, int d1, int t0,
                                            int t1) {
  constexpr int B = A == 0 ? 1 : 0, C = A == 2 ? 1 : 2;
  const int dims[3] = {X, Y, Z};
  const float r0 = __fsub_rn(__fadd_rn(static_cast<float>(d0), 0.25f), m[3]);
  const float r1 = __fsub_rn(__fadd_rn(static_cast<float>(d1), 0.25f), m[7]);
  float acc = 0.0f;
  for (int t = t0; t < t1; ++t) {
    const float xa = static_cast<float>(t) + 0.5f;
    const float q0 = r0 - m[A] * xa, q1 = r1 - m[4 + A] * xa;
    const float cb = g.i00 * q0 + g.i01 * q1, cc = g.i10 * q0 + g.i11 * q1;
    const int lob = static_cast<int>(ceilf(cb - g.hb - 0.5f));
    const int loc = static_cast<int>(ceilf(cc - g.hc - 0.5f));
    for (int kb = 0; kb < g.nb; ++kb) {
      const int ib = lob + kb;
      if (ib < 0 || ib >= dims[B]) continue;
      // Within the row the contributing voxels satisfy, for both detector
      // axes r, P_r - (d_r + 1/4) = m_r[B] xb + m_r[C] xc - q_r in [-1, 1):
      // an interval of xc per axis.  Their intersection, widened by 0.01 in
      // P (far above this estimate's float32 error), bounds the
      // candidates; the exact taps below still decide.
      const float xb = static_cast<float>(ib) + 0.5f;
      int c0 = max(loc, 0), c1 = min(loc + g.nc - 1, dims[C] - 1);
      {
        const float s0 = m[B] * xb - q0, s1 = m[4 + B] * xb - q1;
        const float e0 = m[C], e1 = m[4 + C];
        const float cmax = static_cast<float>(dims[C]);
        if (fabsf(e0) > 1e-6f) {
          const float u = (-1.01f - s0) / e0, w = (1.01f - s0) / e0;
          // clamped to the volume before the conversions (|u|, |w| can be huge)
          c0 = max(c0, static_cast<int>(ceilf(fmaxf(fminf(u, w) - 0.5f, -1.0f))));
          c1 = min(c1, static_cast<int>(floorf(fminf(fmaxf(u, w) - 0.5f, cmax))));
        } else if (!(fabsf(s0) <= 1.01f)) {
          continue;
        }
        if (fabsf(e1) > 1e-6f) {
          const float u = (-1.01f - s1) / e1, w = (1.01f - s1) / e1;
          // clamped to the volume before the conversions (|u|, |w| can be huge)
          c0 = max(c0, static_cast<int>(ceilf(fmaxf(fminf(u, w) - 0.5f, -1.0f))));
          c1 = min(c1, static_cast<int>(floorf(fminf(fmaxf(u, w) - 0.5f, cmax))));
        } else if (!(fabsf(s1) <= 1.01f)) {
          continue;
        }
      }
      for (int ic = c0; ic <= c1; ++ic) {
        const int i = A == 0 ? t : (B == 0 ? ib : ic);
        const int j = A == 1 ? t : (B == 1 ? ib : ic);
        const int k = A == 2 ? t : (C == 2 ? ic : ib);
        const float x0 = __fadd_rn(static_cast<float>(i), 0.5f);
        const float x1 = __fadd_rn(static_cast<float>(j), 0.5f);
        const float x2 = __fadd_rn(static_cast<float>(k), 0.5f);
        float tr[2];
        int ul[2];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          const float* mr = m + 4 * r;
          const float P = __fadd_rn(__fadd_rn(__fadd_rn(__fmul_rn(mr[0], x0),
                                                        __fmul_rn(mr[1], x1)),
                                              __fmul_rn(mr[2], x2)),
                                    mr[3]);
          const float left = __fsub_rn(P, 0.25f);
          tr[r] = fminf(__fsub_rn(ceilf(left), left), 0.5f);
          ul[r] = static_cast<int>(floorf(left));
        }
        const bool h0 = ul[0] == d0, l0 = ul[0] == d0 - 1;
        const bool h1 = ul[1] == d1, l1 = ul[1] == d1 - 1;
        if ((h0 || l0) && (h1 || l1)) {
          const float f0 = h0 ? tr[0] : __fsub_rn(0.5f, tr[0]);
          const float f1 = h1 ? tr[1] : __fsub_rn(0.5f, tr[1]);
          const float w = __fmul_rn(__fmul_rn(f0, f1), 4.0f);
          acc = __builtin_fmaf(w, vol[(static_cast<size_t>(i) * Y + j) * Z + k], acc);
        }
      }
    }
  }
  return acc;
}

// One workgroup per (view, 64 consecutive pixels): wave w of GEN_W takes
// planes [w n / GEN_W, (w + 1) n / GEN_W) of the march axis and the
// partial sums are added in wave order.  Few views give few pixel groups,
// so the planes are split over 16 waves to fill the chip (4 waves: 148 us,
// 16: 120 us on the three-view 129 x 130 x 131 comparison geometry).
constexpr int GEN_W = 16;
__global__ __launch_bounds__(64 * GEN_W) void gen_fwd_kernel(const float* __restrict__ vol,
                                                          float* __restrict__ proj,
                                                          const float* __restrict__ mats,
                                                          const GenView* __restrict__ gv,
                                                          int X, int Y, int Z, int D0, int D1,
                                                          const int32_t* __restrict__ pred) {
  if (pred_off(pred)) return;
  __shared__ float part[GEN_W][64];
  const int v = blockIdx.y;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long e = blockIdx.x * 64LL + lane;
  const bool ok = e < static_cast<long long>(D0) * D1;
  const int d0 = ok ? static_cast<int>(e / D1) : 0, d1 = ok ? static_cast<int>(e % D1) : 0;
  const float* m = mats + 8 * v;
  const GenView g = gv[v];
  const int n = g.a == 0 ? X : (g.a == 1 ? Y : Z);
  const int t0 = (wave * n) / GEN_W, t1 = ((wave + 1) * n) / GEN_W;
  float acc = 0.0f;
  if (ok) {
    if (g.a == 0)
      acc = gen_gather<0>(vol, m, g, X, Y, Z, d0, d1, t0, t1);
    else if (g.a == 1)
      acc = gen_gather<1>(vol, m, g, X, Y, Z, d0, d1, t0, t1);
    else
      acc = gen_gather<2>(vol, m, g, X, Y, Z, d0, d1, t0, t1);
  }
  part[wave][lane] = acc;
  __syncthreads();
  if (wave == 0 && ok) {
    float sum = part[0][lane];
    for (int w = 1; w < GEN_W; ++w) sum = __fadd_rn(sum, part[w][lane]);
    proj[static_cast<size_t>(v) * D0 * D1 + e] = sum;
  }
}

// Reference weight arrays of one view: ul (2, X, Y, Z) int32 and
// w (4, X, Y, Z) float32 (_xray3d.py:206-266), computed by the same device
// functions that the projection kernels use.
__global__ __launch_bounds__(256) void weights_kernel(Plan p, int view, int32_t* __restrict__ ul,
                                                      float* __restrict__ w) {
  const long long n = static_cast<long long>(p.X) * p.Y * p.Z;
  const long long e = blockIdx.x * 256LL + threadIdx.x;
  if (e >= n) return;
  const int idx[3] = {static_cast<int>(e / (static_cast<long long>(p.Y) * p.Z)),
                      static_cast<int>((e / p.Z) % p.Y), static_cast<int>(e % p.Z)};
  int u[2];
  float ww[4];
  if (p.mode == MODE_SLICE) {
    const float4 g = p.geom[view];
    const Tap tp = slice_tap(g.x, g.y, g.z, idx[p.pax], idx[p.qax], p.U);
    const int col = idx[p.sax] + p.delta;
    const int Dd = p.dax == 0 ? p.D0 : p.D1;
    const bool cok = col >= 0 && col < Dd;
    u[p.dax] = col;
    u[1 - p.dax] = tp.ul;
    // taps (ul0, ul1), (ul0 + 1, ul1), (ul0, ul1 + 1), (ul0 + 1, ul1 + 1)
    ww[0] = cok ? tp.wl : 0.0f;
    ww[3] = 0.0f;
    if (p.dax == 1) {
      ww[1] = cok ? tp.wr : 0.0f;
      ww[2] = 0.0f;
    } else {
      ww[1] = 0.0f;
      ww[2] = cok ? tp.wr : 0.0f;
    }
  } else {
    gen_weights(p.mats + 8 * view, idx[0], idx[1], idx[2], p.D0, p.D1, u, ww);
  }
  ul[e] = u[0];
  ul[n + e] = u[1];
  for (int t = 0; t < 4; ++t) w[t * n + e] = ww[t];
}

// ------------------------------------------------------------- host --

static float mat(const std::vector<float>& m, int v, int r, int c) { return m[8 * v + 4 * r + c]; }

// Slice decomposition test for slice axis sax and detector axis dax, all in
// the reference's float32 arithmetic.
static bool try_slice(Plan& p, const std::vector<float>& m, int sax, int dax) {
  const int e = 1 - dax;
  const int shape[3] = {p.X, p.Y, p.Z};
  const int S = shape[sax];
  int delta = 0;
  for (int v = 0; v < p.V; ++v) {
    for (int a = 0; a < 3; ++a)
      if (a != sax && mat(m, v, dax, a) != 0.0f) return false;
    if (mat(m, v, e, sax) != 0.0f) return false;
    // Every view must map slice s to column s + delta with t_d = 1/2,
    // evaluated as the reference does: ((0 + m x) + 0) + t up to exact zeros.
    const float ms = mat(m, v, dax, sax), td = mat(m, v, dax, 3);
    if (!(ms != 0.0f)) return false;
    for (int s = 0; s < S; ++s) {
      const float x = static_cast<float>(s) + 0.5f;
      const float prod = ms * x;
      const float P = prod + td;
      const float left = P - 0.25f;
      const float ce = std::ceil(left);
      const float tt = std::fmin(ce - left, 0.5f);
      if (tt != 0.5f) return false;
      const int d = static_cast<int>(std::floor(left)) - s;
      if (v == 0 && s == 0)
        delta = d;
      else if (d != delta)
        return false;
    }
  }
  int pax = -1, qax = -1;
  for (int a = 0; a < 3; ++a)
    if (a != sax) (pax < 0 ? pax : qax) = a;
  float fmin = INFINITY;
  for (int v = 0; v < p.V; ++v) {
    const float a = mat(m, v, e, pax), b = mat(m, v, e, qax), t = mat(m, v, e, 3);
    if (!std::isfinite(a) || !std::isfinite(b) || !std::isfinite(t)) return false;
    fmin = std::fmin(fmin, std::fmax(std::fabs(a), std::fabs(b)));
  }
  if (!(fmin > 0.0f)) return false;
  const double need = std::ceil((FB + 1) / static_cast<double>(fmin)) + 6.0;
  if (need > 64.0) return false;
  p.cpr = need <= 16.0 ? 16 : (need <= 32.0 ? 32 : 64);
  // pixels of one bin group (left in [g, g + 1)) along the fast axis
  const double spg = std::ceil(1.0 / static_cast<double>(fmin) + 0.01);
  if (spg > 3.0) return false;
  p.spg = spg <= 2.0 ? 2 : 3;
  p.mode = MODE_SLICE;
  p.sax = sax;
  p.dax = dax;
  p.delta = delta;
  p.pax = pax;
  p.qax = qax;
  p.P = shape[pax];
  p.Q = shape[qax];
  p.S = S;
  p.SQ = (S + 3) / 4;
  p.U = e == 0 ? p.D0 : p.D1;
  // The forward reads the canonical (slice-fastest) volume through one
  // buffer resource; buffer offsets are 32-bit.
  p.vol_direct = false;
  const double rowb = 16.0 * ((S + 3) / 4);
  if (static_cast<double>(p.P) * p.Q * rowb >= 2147483647.0) return false;
  const int Ue = e == 0 ? p.D0 : p.D1;
  if (static_cast<double>(p.V) * Ue * rowb >= 2147483647.0) return false;
  p.proj_direct = dax == 1 && delta == 0 && p.D1 == S && S % 4 == 0;
  return true;
}

void destroy(Plan& p) {
  if (p.mats) (void)hipFree(p.mats);
  if (p.geom) (void)hipFree(p.geom);
  if (p.volc) (void)hipFree(p.volc);
  if (p.proja) (void)hipFree(p.proja);
  if (p.gview) (void)hipFree(p.gview);
  p.gview = nullptr;
  p.mats = nullptr;
  p.geom = nullptr;
  p.volc = p.proja = nullptr;
}

static int dev_alloc(void** ptr, size_t bytes) {
  if (hipMalloc(ptr, bytes) != hipSuccess) {
    (void)hipGetLastError();
    *ptr = nullptr;
    return fail(SCXR_ERR_NOMEM, "hipMalloc of %zu bytes failed", bytes);
  }
  return SCXR_OK;
}

int create(Plan& p, int X, int Y, int Z, int V, const float* mats, int D0, int D1,
           int allow_slice) {
  p.X = X;
  p.Y = Y;
  p.Z = Z;
  p.V = V;
  p.D0 = D0;
  p.D1 = D1;
  std::vector<float> m(mats, mats + 8 * static_cast<size_t>(V));
  for (float f : m)
    SCXR_REQUIRE(std::isfinite(f), SCXR_ERR_GEOM, "non-finite projection matrix entry");
  if (allow_slice & 1) {
    for (int dax = 1; dax >= 0 && p.mode != MODE_SLICE; --dax)
      for (int sax = 0; sax < 3 && p.mode != MODE_SLICE; ++sax) try_slice(p, m, sax, dax);
  }
  int rc = dev_alloc(reinterpret_cast<void**>(&p.mats), m.size() * sizeof(float));
  if (rc) return rc;
  SCXR_HIP(hipMemcpy(p.mats, m.data(), m.size() * sizeof(float), hipMemcpyHostToDevice));
  if (p.mode == MODE_GENERAL) {
    std::vector<GenView> gv(V);
    for (int v = 0; v < V; ++v) {
      double best = 0.0;
      int ba = -1;
      for (int a = 0; a < 3; ++a) {
        const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
        const double det = static_cast<double>(mat(m, v, 0, b)) * mat(m, v, 1, c) -
                           static_cast<double>(mat(m, v, 0, c)) * mat(m, v, 1, b);
        if (std::fabs(det) > std::fabs(best)) {
          best = det;
          ba = a;
        }
      }
      SCXR_REQUIRE(ba >= 0 && std::fabs(best) > 1e-6, SCXR_ERR_GEOM,
                   "projection matrix of view %d has rank < 2", v);
      const int b = ba == 0 ? 1 : 0, c = ba == 2 ? 1 : 2;
      GenView g{};
      g.a = ba;
      const double i00 = mat(m, v, 1, c) / best, i01 = -mat(m, v, 0, c) / best;
      const double i10 = -mat(m, v, 1, b) / best, i11 = mat(m, v, 0, b) / best;
      // box half-width 1 in P; margin for the float32 solve
      const double hb = std::fabs(i00) + std::fabs(i01) + 1e-2;
      const double hc = std::fabs(i10) + std::fabs(i11) + 1e-2;
      g.i00 = static_cast<float>(i00);
      g.i01 = static_cast<float>(i01);
      g.i10 = static_cast<float>(i10);
      g.i11 = static_cast<float>(i11);
      g.hb = static_cast<float>(hb);
      g.hc = static_cast<float>(hc);
      g.nb = static_cast<int>(std::floor(2.0 * hb)) + 2;
      g.nc = static_cast<int>(std::floor(2.0 * hc)) + 2;
      gv[v] = g;
    }
    rc = dev_alloc(reinterpret_cast<void**>(&p.gview), V * sizeof(GenView));
    if (rc) return rc;
    SCXR_HIP(hipMemcpy(p.gview, gv.data(), V * sizeof(GenView), hipMemcpyHostToDevice));
  }
  if (p.mode == MODE_SLICE) {
    const int e = 1 - p.dax;
    std::vector<float4> g(V);
    for (int v = 0; v < V; ++v)
      g[v] = make_float4(mat(m, v, e, p.pax), mat(m, v, e, p.qax), mat(m, v, e, 3), 0.0f);
    // Staged adjoint: a 4 x 8 tile's rows per view, at most
    // ceil(3 |a| + 7 |b| + rounding) + 2 (the corner bins' spread plus the
    // right-hand tap), must fit ADJ_NW.  The rounding margin covers the
    // float32 evaluation of the corner centres (a few ulps of |P|).  The
    // kernel packs two detector-row indices into 16-bit fields: U < 65536.
    p.adj_staged = !(allow_slice & 2) && p.U < 65536;
    for (int v = 0; v < V && p.adj_staged; ++v) {
      const double a = std::fabs(static_cast<double>(g[v].x)),
                   b = std::fabs(static_cast<double>(g[v].y));
      const double pmax = a * (p.P + 1) + b * (p.Q + 1) + std::fabs(static_cast<double>(g[v].z));
      const double spread = 3.0 * a + 7.0 * b + 1e-3 + 1e-6 * (pmax + 2.0);
      if (std::ceil(spread) + 2.0 > ADJ_NW) p.adj_staged = false;
    }
    rc = dev_alloc(reinterpret_cast<void**>(&p.geom), V * sizeof(float4));
    if (rc) return rc;
    SCXR_HIP(hipMemcpy(p.geom, g.data(), V * sizeof(float4), hipMemcpyHostToDevice));
    const size_t s4 = 4 * static_cast<size_t>(p.SQ);
    {
      const size_t rows = static_cast<size_t>(p.P) * p.Q;
      rc = dev_alloc(reinterpret_cast<void**>(&p.volc), rows * s4 * sizeof(float));
      if (rc) return rc;
    }
    if (!p.proj_direct) {
      rc = dev_alloc(reinterpret_cast<void**>(&p.proja),
                     static_cast<size_t>(V) * p.U * s4 * sizeof(float));
      if (rc) return rc;
    }
  }
  return SCXR_OK;
}

static int transpose(const float* in, float* out, int B, int Rout, int Cout, int Rin, int Cin,
                     int roff, int coff, long long ib, long long ir, long long ob, long long oc,
                     const int32_t* pred, hipStream_t s) {
  dim3 grid(ceil_div(Cout, 32), ceil_div(Rout, 32), B);
  hipLaunchKernelGGL(transpose_kernel, grid, dim3(256), 0, s, in, out, Rout, Cout, Rin, Cin, roff,
                     coff, ib, ir, ob, oc, pred);
  SCXR_LAUNCHED("xr3d transpose");
  return SCXR_OK;
}

static int shift_rows(const float* in, float* out, long long rows, int Cout, int Cin, int off,
                      long long is, long long os, const int32_t* pred, hipStream_t s) {
  const long long n = rows * Cout;
  const unsigned blocks = static_cast<unsigned>(std::min<long long>((n + 255) / 256, 65536));
  hipLaunchKernelGGL(shift_rows_kernel, dim3(blocks), dim3(256), 0, s, in, out, rows, Cout, Cin,
                     off, is, os, pred);
  SCXR_LAUNCHED("xr3d shift_rows");
  return SCXR_OK;
}

// vol (X, Y, Z) -> vol_c[p][q][s] (slices padded to 4 SQ with zeros)
static int vol_to_canon(const Plan& p, const float* vol, float* vc, const int32_t* pred,
                        hipStream_t s) {
  const int S4 = 4 * p.SQ;
  switch (p.sax) {
    case 2:
      return shift_rows(vol, vc, static_cast<long long>(p.X) * p.Y, S4, p.S, 0, p.S, S4, pred,
                        s);
    case 1:  // vol[p][s][q] -> vc[p][q][s]
      return transpose(vol, vc, p.X, S4, p.Z, p.S, p.Z, 0, 0,
                       static_cast<long long>(p.Y) * p.Z, p.Z,
                       static_cast<long long>(p.Z) * S4, S4, pred, s);
    default: {  // vol[s][pq] -> vc[pq][s]
      const int PQ = p.Y * p.Z;
      return transpose(vol, vc, 1, S4, PQ, p.S, PQ, 0, 0, 0, PQ, 0, S4, pred, s);
    }
  }
}

static int canon_to_vol(const Plan& p, const float* vc, float* vol, const int32_t* pred,
                        hipStream_t s) {
  const int S4 = 4 * p.SQ;
  switch (p.sax) {
    case 2:
      return shift_rows(vc, vol, static_cast<long long>(p.X) * p.Y, p.S, S4, 0, S4, p.S, pred,
                        s);
    case 1:  // vol[p][s][q] = vc[p][q][s]
      return transpose(vc, vol, p.X, p.Z, p.S, p.Z, S4, 0, 0,
                       static_cast<long long>(p.Z) * S4, S4,
                       static_cast<long long>(p.Y) * p.Z, p.Z, pred, s);
    default: {  // vol[s][pq] = vc[pq][s]
      const int PQ = p.Y * p.Z;
      return transpose(vc, vol, 1, PQ, p.S, PQ, S4, 0, 0, 0, S4, 0, PQ, pred, s);
    }
  }
}

// proj (V, D0, D1) -> proj_c[v][u][s]
static int proj_to_canon(const Plan& p, const float* y, float* yc, const int32_t* pred,
                         hipStream_t s) {
  const int S4 = 4 * p.SQ;
  if (p.dax == 1)  // proj[v][u][s + delta]
    return shift_rows(y, yc, static_cast<long long>(p.V) * p.D0, S4, p.D1, p.delta, p.D1, S4,
                      pred, s);
  // proj[v][s + delta][u] -> yc[v][u][s]
  return transpose(y, yc, p.V, S4, p.U, p.D0, p.D1, p.delta, 0,
                   static_cast<long long>(p.D0) * p.D1, p.D1,
                   static_cast<long long>(p.U) * S4, S4, pred, s);
}

static int canon_to_proj(const Plan& p, const float* yc, float* y, const int32_t* pred,
                         hipStream_t s) {
  const int S4 = 4 * p.SQ;
  if (p.dax == 1)  // proj[v][u][c] = yc[v][u][c - delta] for c - delta in [0, S)
    return shift_rows(yc, y, static_cast<long long>(p.V) * p.D0, p.D1, p.S, -p.delta, S4, p.D1,
                      pred, s);
  // proj[v][c][u] = yc[v][u][c - delta]
  return transpose(yc, y, p.V, p.U, p.D0, p.U, p.S, 0, -p.delta,
                   static_cast<long long>(p.U) * S4, S4,
                   static_cast<long long>(p.D0) * p.D1, p.D1, pred, s);
}

int forward(const Plan& p, const float* vol, float* proj, const int32_t* pred, hipStream_t s) {
  if (p.mode == MODE_GENERAL) {
    const long long npix = static_cast<long long>(p.D0) * p.D1;
    hipLaunchKernelGGL(gen_fwd_kernel, dim3(ceil_div(npix, 64), p.V), dim3(64 * GEN_W), 0, s, vol,
                       proj, p.mats, p.gview, p.X, p.Y, p.Z, p.D0, p.D1, pred);
    SCXR_LAUNCHED("xr3d general forward");
    return SCXR_OK;
  }
  const float* vc = vol;
  if (!p.vol_direct) {
    int rc = vol_to_canon(p, vol, p.volc, pred, s);
    if (rc) return rc;
    vc = p.volc;
  }
  float* yc = p.proj_direct ? proj : p.proja;
  dim3 grid(ceil_div(p.U, FB), ceil_div(p.V, 4), ceil_div(p.SQ, 64));
  const float4* v4 = reinterpret_cast<const float4*>(vc);
  float4* y4 = reinterpret_cast<float4*>(yc);
#define SCXR_FWD3(C, G)                                                                       \
  hipLaunchKernelGGL((fwd_slice_kernel<C, G>), grid, dim3(256), 0, s, v4, y4, p.geom, p.P, p.Q, \
                     p.SQ, p.U, p.V, pred)
  if (p.spg == 2) {
    if (p.cpr == 16)
      SCXR_FWD3(16, 2);
    else if (p.cpr == 32)
      SCXR_FWD3(32, 2);
    else
      SCXR_FWD3(64, 2);
  } else {
    if (p.cpr == 16)
      SCXR_FWD3(16, 3);
    else if (p.cpr == 32)
      SCXR_FWD3(32, 3);
    else
      SCXR_FWD3(64, 3);
  }
#undef SCXR_FWD3
  SCXR_LAUNCHED("xr3d slice forward");
  if (!p.proj_direct) return canon_to_proj(p, yc, proj, pred, s);
  return SCXR_OK;
}

int adjoint(const Plan& p, const float* proj, float* vol, const int32_t* pred, hipStream_t s) {
  if (p.mode == MODE_GENERAL) {
    const long long n = static_cast<long long>(p.X) * p.Y * p.Z;
    hipLaunchKernelGGL(gen_adj_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, proj, vol, p.mats,
                       p.X, p.Y, p.Z, p.V, p.D0, p.D1, pred);
    SCXR_LAUNCHED("xr3d general adjoint");
    return SCXR_OK;
  }
  const float* yc = proj;
  if (!p.proj_direct) {
    int rc = proj_to_canon(p, proj, p.proja, pred, s);
    if (rc) return rc;
    yc = p.proja;
  }
  float* vc = p.vol_direct ? vol : p.volc;
  dim3 grid(ceil_div(p.Q, 8), ceil_div(p.P, 4), ceil_div(p.SQ, 64));
  hipLaunchKernelGGL(p.adj_staged ? adj_slice_staged_kernel : adj_slice_kernel, grid, dim3(256), 0,
                     s, reinterpret_cast<const float4*>(yc), reinterpret_cast<float4*>(vc),
                     p.geom, p.P, p.Q, p.SQ, p.U, p.V, pred);
  SCXR_LAUNCHED("xr3d slice adjoint");
  if (!p.vol_direct) return canon_to_vol(p, vc, vol, pred, s);
  return SCXR_OK;
}

}  // namespace xr3d
}  // namespace scxr

struct scxr_xr3d_plan {
  scxr::xr3d::Plan p;
};

using namespace scxr;

extern "C" {

int scxr_xr3d_plan_create(int32_t X, int32_t Y, int32_t Z, int32_t V, const float* matrices,
                          int32_t D0, int32_t D1, int32_t allow_slice, scxr_xr3d_plan** out) {
  SCXR_REQUIRE(out && matrices, SCXR_ERR_ARG, "null argument to xr3d_plan_create");
  SCXR_REQUIRE(X > 0 && Y > 0 && Z > 0 && V > 0 && D0 > 0 && D1 > 0, SCXR_ERR_SHAPE,
               "non-positive size (X=%d Y=%d Z=%d V=%d D0=%d D1=%d)", X, Y, Z, V, D0, D1);
  SCXR_REQUIRE(static_cast<long long>(X) * Y * Z < (1LL << 31) &&
                   static_cast<long long>(D0) * D1 < (1LL << 31),
               SCXR_ERR_SHAPE, "volume or detector too large");
  auto* plan = new scxr_xr3d_plan();
  int rc = xr3d::create(plan->p, X, Y, Z, V, matrices, D0, D1, allow_slice);
  if (rc != SCXR_OK) {
    xr3d::destroy(plan->p);
    delete plan;
    return rc;
  }
  *out = plan;
  return SCXR_OK;
}

int scxr_xr3d_plan_destroy(scxr_xr3d_plan* plan) {
  if (!plan) return SCXR_OK;
  xr3d::destroy(plan->p);
  delete plan;
  return SCXR_OK;
}

int scxr_xr3d_info(const scxr_xr3d_plan* plan, int32_t* out) {
  SCXR_REQUIRE(plan && out, SCXR_ERR_ARG, "null argument to xr3d_info");
  const auto& p = plan->p;
  out[0] = p.mode;
  out[1] = p.sax;
  out[2] = p.dax;
  out[3] = p.delta;
  out[4] = p.cpr;
  out[5] = p.vol_direct;
  out[6] = p.proj_direct;
  out[7] = p.U;
  out[8] = p.adj_staged;
  return SCXR_OK;
}

int scxr_xr3d_forward(const scxr_xr3d_plan* plan, const float* vol, float* proj,
                      const int32_t* pred, void* stream) {
  SCXR_REQUIRE(plan && vol && proj, SCXR_ERR_ARG, "null argument to xr3d_forward");
  return xr3d::forward(plan->p, vol, proj, pred, as_stream(stream));
}

int scxr_xr3d_adjoint(const scxr_xr3d_plan* plan, const float* proj, float* vol,
                      const int32_t* pred, void* stream) {
  SCXR_REQUIRE(plan && vol && proj, SCXR_ERR_ARG, "null argument to xr3d_adjoint");
  return xr3d::adjoint(plan->p, proj, vol, pred, as_stream(stream));
}

int scxr_xr3d_weights(const scxr_xr3d_plan* plan, int32_t view, int32_t* ul, float* w,
                      void* stream) {
  SCXR_REQUIRE(plan && ul && w, SCXR_ERR_ARG, "null argument to xr3d_weights");
  const auto& p = plan->p;
  SCXR_REQUIRE(view >= 0 && view < p.V, SCXR_ERR_ARG, "view %d out of range", view);
  const long long n = static_cast<long long>(p.X) * p.Y * p.Z;
  hipLaunchKernelGGL(xr3d::weights_kernel, dim3(ceil_div(n, 256)), dim3(256), 0,
                     as_stream(stream), p, view, ul, w);
  SCXR_LAUNCHED("xr3d weights");
  return SCXR_OK;
}

}  // extern "C"
