// Multi-dimensional complex FFT and DFT-domain circular convolution
// (scico/linop/_circconv.py:79-209) for gfx950.
//
// Layout: data (B, F0, ..., F{d-1}) row-major, B = product of the leading
// non-transformed axes, d <= 3 transformed axes.  One kernel performs one
// pass of 1D transforms along one axis; each workgroup loads LPB lines
// into LDS, transforms them and writes them back:
//   * power-of-two lengths <= 4096: iterative radix-2 decimation in time
//     (bit-reversed load, log2(n) butterfly stages);
//   * other lengths <= 512: direct DFT (out-of-place in LDS).
// Twiddles are float32 roundings of float64 exp(-2 pi i k / n).
// Inverse passes scale by 1/n, like numpy / jax ifft.
//
// Convolution y = real(ifftn(H * fftn(x))): forward passes on F{d-1} ..
// F1, one fused pass on F0 (forward, multiply by H or conj(H), inverse),
// inverse passes on F1 .. F{d-1}; the first pass reads the real input and
// the last one writes the real part.  With d == 1 the single fused pass
// does both.
//
// Real convolution over two power-of-two axes (the deconvolution workload)
// runs on half spectra instead: rows transformed two real rows per complex
// FFT into (F1/2 + 1)-wide half spectra, one fused pass on those columns
// with the Hermitian part of H, rows inverted back to real.  At 1024 x 1024
// each line is one 256-thread workgroup's registers (four values per
// thread, radix-4 Stockham stages exchanged through LDS).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "apgm_stats.h"
#include "common.h"

using namespace scxr;

namespace scxr {
namespace fft {

constexpr int MAX_POW2 = 4096;
constexpr int MAX_DFT = 512;
constexpr int LDS_ELEMS = 4096;  // float2 of LDS per workgroup (32 KiB)

struct Pass {
  int n, log2n;          // line length, log2 (pow2) or -1 (direct DFT)
  long long stride;      // element stride along the line
  long long inner;       // lines per outer index (= stride)
  long long lines;       // total lines
  int lpb;               // lines per workgroup
  int in_real;           // input is real float32
  int out_real;          // output real part (float32)
  int inverse;           // direction of the (non-fused) transform
  int fused;             // forward, multiply by h (conj if conj_h), inverse
  int conj_h;
  const float2* tw;      // pow2: n/2 twiddles; DFT: n twiddles
  const float2* h;       // fused: spectrum, indexed by k * stride + inner
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(__fsub_rn(__fmul_rn(a.x, b.x), __fmul_rn(a.y, b.y)),
                     __fadd_rn(__fmul_rn(a.x, b.y), __fmul_rn(a.y, b.x)));
}

__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

// In-place radix-2 DIT stages on lpb lines of length n stored at buf[l*n + k]
// (input in bit-reversed order, output in natural order).
__device__ void dit_stages(float2* buf, int n, int log2n, int lpb, const float2* __restrict__ tw,
                           bool inv) {
  const int half_total = lpb * (n >> 1);
  for (int s = 1; s <= log2n; ++s) {
    const int half = 1 << (s - 1);
    const int tstep = n >> s;  // twiddle index step
    __syncthreads();
    for (int b = threadIdx.x; b < half_total; b += blockDim.x) {
      const int l = b >> (log2n - 1);
      const int j = b & ((n >> 1) - 1);
      const int pos = j & (half - 1);
      const int i0 = ((j >> (s - 1)) << s) + pos;
      float2* line = buf + l * n;
      float2 w = tw[pos * tstep];
      if (inv) w = conjf2(w);
      const float2 a = line[i0];
      const float2 t = cmul(w, line[i0 + half]);
      line[i0] = make_float2(__fadd_rn(a.x, t.x), __fadd_rn(a.y, t.y));
      line[i0 + half] = make_float2(__fsub_rn(a.x, t.x), __fsub_rn(a.y, t.y));
    }
  }
  __syncthreads();
}

// Direct DFT of lpb lines from src to dst (both [l*n + k]).
__device__ void direct_dft(const float2* src, float2* dst, int n, int lpb,
                           const float2* __restrict__ tw, bool inv) {
  __syncthreads();
  for (int e = threadIdx.x; e < lpb * n; e += blockDim.x) {
    const int l = e / n, k = e - l * n;
    const float2* x = src + l * n;
    float re = 0.0f, im = 0.0f;
    int idx = 0;
    for (int m = 0; m < n; ++m) {
      float2 w = tw[idx];
      if (inv) w = conjf2(w);
      const float2 p = cmul(x[m], w);
      re = __fadd_rn(re, p.x);
      im = __fadd_rn(im, p.y);
      idx += k;
      if (idx >= n) idx -= n;
    }
    dst[e] = make_float2(re, im);
  }
  __syncthreads();
}

__device__ __forceinline__ int bitrev(int k, int log2n) {
  return static_cast<int>(__builtin_bitreverse32(static_cast<unsigned>(k)) >> (32 - log2n));
}

__global__ __launch_bounds__(256) void pass_kernel(const void* __restrict__ in,
                                                   void* __restrict__ out, Pass P, float scale,
                                                   const int32_t* pred) {
  if (pred_off(pred)) return;
  extern __shared__ float2 lds[];  // two buffers of lpb * n
  __shared__ long long base[64];
  const int n = P.n, lpb = P.lpb, tot = lpb * n;
  float2* A = lds;
  float2* B = lds + tot;
  const long long line0 = static_cast<long long>(blockIdx.x) * lpb;
  const int nl = static_cast<int>(min(static_cast<long long>(lpb), P.lines - line0));
  if (threadIdx.x < lpb) {
    const long long L = line0 + threadIdx.x;
    const long long outer = L / P.inner, inr = L - outer * P.inner;
    base[threadIdx.x] = outer * n * P.stride + inr;
  }
  __syncthreads();
  const bool pow2 = P.log2n >= 0;
  const bool contiguous = P.stride == 1;
  auto split = [&](int e, int& l, int& k) {
    if (contiguous) {
      l = e / n;
      k = e - l * n;
    } else {
      k = e / lpb;
      l = e - k * lpb;
    }
  };
  // load (bit-reversed for the radix-2 path)
  for (int e = threadIdx.x; e < tot; e += blockDim.x) {
    int l, k;
    split(e, l, k);
    float2 v = make_float2(0.0f, 0.0f);
    if (l < nl) {
      const long long g = base[l] + static_cast<long long>(k) * P.stride;
      v = P.in_real ? make_float2(static_cast<const float*>(in)[g], 0.0f)
                    : static_cast<const float2*>(in)[g];
    }
    A[l * n + (pow2 ? bitrev(k, P.log2n) : k)] = v;
  }
  // forward (or the only) transform: result in natural order in R
  const bool inv0 = P.fused ? false : (P.inverse != 0);
  float2* R;
  if (pow2) {
    dit_stages(A, n, P.log2n, lpb, P.tw, inv0);
    R = A;
  } else {
    direct_dft(A, B, n, lpb, P.tw, inv0);
    R = B;
  }
  if (P.fused) {
    // multiply by the spectrum into the other buffer, then inverse transform
    float2* Q = (R == A) ? B : A;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
      const int l = e / n, k = e - l * n;
      float2 pr = make_float2(0.0f, 0.0f);
      if (l < nl) {
        const long long inr = (line0 + l) % P.inner;
        float2 hv = P.h[static_cast<long long>(k) * P.stride + inr];
        if (P.conj_h) hv = conjf2(hv);
        pr = cmul(hv, R[e]);
      }
      Q[l * n + (pow2 ? bitrev(k, P.log2n) : k)] = pr;
    }
    if (pow2) {
      dit_stages(Q, n, P.log2n, lpb, P.tw, true);
      R = Q;
    } else {
      direct_dft(Q, R, n, lpb, P.tw, true);  // R is free again
    }
  }
  // store
  for (int e = threadIdx.x; e < tot; e += blockDim.x) {
    int l, k;
    split(e, l, k);
    if (l < nl) {
      const long long g = base[l] + static_cast<long long>(k) * P.stride;
      const float2 v = R[l * n + k];
      if (P.out_real)
        static_cast<float*>(out)[g] = __fmul_rn(v.x, scale);
      else
        static_cast<float2*>(out)[g] = make_float2(__fmul_rn(v.x, scale), __fmul_rn(v.y, scale));
    }
  }
}

// Power-of-two lengths: Stockham auto-sort passes of radix 16 (after one
// pass of radix 2, 4 or 8 for log2 n mod 4) on lpb lines held in one LDS
// buffer.  Each
// thread takes whole butterflies: it reads its R inputs x[j + r n / R] into
// registers, twiddles them by w^(k r) (k = j mod Ns, w = exp(-2 pi i /
// (R Ns)), taken from the n-entry table), transforms them and, after a
// barrier, writes them to x[(j / Ns) R Ns + k + r Ns]; the output is in
// natural order, so the fused convolution pass multiplies by the spectrum
// and runs the inverse stages on the same buffer.
constexpr int SH_THREADS = 256;
constexpr int SH_MAXV = 32;  // values per thread per stage (lpb n <= 8192)

__device__ __forceinline__ float2 cadd(float2 a, float2 b) {
  return make_float2(__fadd_rn(a.x, b.x), __fadd_rn(a.y, b.y));
}
__device__ __forceinline__ float2 csub(float2 a, float2 b) {
  return make_float2(__fsub_rn(a.x, b.x), __fsub_rn(a.y, b.y));
}

// 4-point DFT of a[0], a[s], a[2 s], a[3 s] in place (natural order).
template <int S, int N>
__device__ __forceinline__ void dft4_strided(float2 (&v)[N], int o, bool inv) {
  const float2 s02 = cadd(v[o], v[o + 2 * S]), d02 = csub(v[o], v[o + 2 * S]);
  const float2 s13 = cadd(v[o + S], v[o + 3 * S]), d13 = csub(v[o + S], v[o + 3 * S]);
  // -i d13 (forward) or +i d13 (inverse)
  const float2 t = inv ? make_float2(-d13.y, d13.x) : make_float2(d13.y, -d13.x);
  v[o] = cadd(s02, s13);
  v[o + 2 * S] = csub(s02, s13);
  v[o + S] = cadd(d02, t);
  v[o + 3 * S] = csub(d02, t);
}

// exp(-2 pi i m / 16) = (cos(pi m / 8), -sin(pi m / 8)) for m = 0 .. 15,
// conjugated for the inverse: the internal twiddles of the 8- and 16-point
// butterflies (m is a compile-time constant there, so the tables fold).
__device__ __forceinline__ float2 w16(int m, bool inv) {
  constexpr float c1 = 0.92387953251128674f, s1 = 0.38268343236508978f;
  constexpr float h = 0.70710678118654752f;
  constexpr float cs[16] = {1.0f, c1, h, s1, 0.0f, -s1, -h, -c1,
                            -1.0f, -c1, -h, -s1, 0.0f, s1, h, c1};
  constexpr float sn[16] = {0.0f, s1, h, c1, 1.0f, c1, h, s1,
                            0.0f, -s1, -h, -c1, -1.0f, -c1, -h, -s1};
  const float2 w = make_float2(cs[m & 15], -sn[m & 15]);
  return inv ? make_float2(w.x, -w.y) : w;
}

template <int R>
__device__ __forceinline__ void dft_small(float2 (&v)[R], bool inv) {
  if constexpr (R == 16 || R == 8) {
    // R = 4 Q: inputs n1 + 4 n2 ... as R = Q x 4 with Q = R / 4:
    // DFT over the Q-strided groups, twiddle, then 4- (or 2-) point DFTs.
    constexpr int Q = R / 4;  // 4 (R = 16) or 2 (R = 8)
    // step 1: for each n1 < Q, 4-point DFT over n2 of v[n1 + Q n2]
#pragma unroll
    for (int n1 = 0; n1 < Q; ++n1) dft4_strided<Q, R>(v, n1, inv);
    // step 2: twiddle v[n1 + Q k2] by W_R^(n1 k2)
#pragma unroll
    for (int n1 = 1; n1 < Q; ++n1)
#pragma unroll
      for (int k2 = 1; k2 < 4; ++k2) v[n1 + Q * k2] = cmul(v[n1 + Q * k2], w16(n1 * k2 * (16 / R), inv));
    // step 3: for each k2, Q-point DFT over n1; output k = k2 + 4 k1
    float2 o[R];
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) {
      if constexpr (Q == 4) {
        float2 g[4] = {v[Q * k2], v[1 + Q * k2], v[2 + Q * k2], v[3 + Q * k2]};
        dft4_strided<1, 4>(g, 0, inv);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) o[k2 + 4 * k1] = g[k1];
      } else {
        const float2 a = v[Q * k2], b = v[1 + Q * k2];
        o[k2] = cadd(a, b);
        o[k2 + 4] = csub(a, b);
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = o[r];
  } else if constexpr (R == 2) {
    const float2 a = v[0], b = v[1];
    v[0] = make_float2(__fadd_rn(a.x, b.x), __fadd_rn(a.y, b.y));
    v[1] = make_float2(__fsub_rn(a.x, b.x), __fsub_rn(a.y, b.y));
  } else {
    const float2 s02 = make_float2(__fadd_rn(v[0].x, v[2].x), __fadd_rn(v[0].y, v[2].y));
    const float2 d02 = make_float2(__fsub_rn(v[0].x, v[2].x), __fsub_rn(v[0].y, v[2].y));
    const float2 s13 = make_float2(__fadd_rn(v[1].x, v[3].x), __fadd_rn(v[1].y, v[3].y));
    const float2 d13 = make_float2(__fsub_rn(v[1].x, v[3].x), __fsub_rn(v[1].y, v[3].y));
    // -i d13 (forward) or +i d13 (inverse)
    const float2 t = inv ? make_float2(-d13.y, d13.x) : make_float2(d13.y, -d13.x);
    v[0] = make_float2(__fadd_rn(s02.x, s13.x), __fadd_rn(s02.y, s13.y));
    v[2] = make_float2(__fsub_rn(s02.x, s13.x), __fsub_rn(s02.y, s13.y));
    v[1] = make_float2(__fadd_rn(d02.x, t.x), __fadd_rn(d02.y, t.y));
    v[3] = make_float2(__fsub_rn(d02.x, t.x), __fsub_rn(d02.y, t.y));
  }
}

template <int R>
__device__ __forceinline__ void stockham_stage(float2* X, int n, int lpb, int Ns,
                                               const float2* tw, bool inv) {
  const int nb = n / R, tot = lpb * nb;  // butterflies
  const int tstep = n / (R * Ns);
  float2 v[SH_MAXV / R][R];
  int dst[SH_MAXV / R];
#pragma unroll
  for (int q = 0; q < SH_MAXV / R; ++q) {
    const int b = threadIdx.x + q * SH_THREADS;
    dst[q] = -1;
    if (b < tot) {
      const int l = b / nb, j = b - l * nb;
      float2* x = X + l * n;
      const int k = j & (Ns - 1);
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = x[j + r * nb];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 w = tw[(k * r * tstep) & (n - 1)];
        if (inv) w = conjf2(w);
        v[q][r] = cmul(v[q][r], w);
      }
      dft_small<R>(v[q], inv);
      dst[q] = l * n + (j - k) * R + k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < SH_MAXV / R; ++q) {
    if (dst[q] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) X[dst[q] + r * Ns] = v[q][r];
    }
  }
  __syncthreads();
}

__device__ __forceinline__ void stockham(float2* X, int n, int log2n, int lpb,
                                         const float2* tw, bool inv) {
  // one stage of radix 2, 4 or 8 for log2 n mod 4, then radix-16 stages
  int Ns = 1;
  switch (log2n & 3) {
    case 1: stockham_stage<2>(X, n, lpb, Ns, tw, inv); Ns = 2; break;
    case 2: stockham_stage<4>(X, n, lpb, Ns, tw, inv); Ns = 4; break;
    case 3: stockham_stage<8>(X, n, lpb, Ns, tw, inv); Ns = 8; break;
    default: break;
  }
  for (; Ns < n; Ns *= 16) stockham_stage<16>(X, n, lpb, Ns, tw, inv);
}

__global__ __launch_bounds__(SH_THREADS) void stockham_kernel(const void* __restrict__ in,
                                                              void* __restrict__ out, Pass P,
                                                              float scale,
                                                              const int32_t* pred) {
  if (pred_off(pred)) return;
  extern __shared__ float2 lds[];  // lpb * n values
  __shared__ long long base[64];
  __shared__ long long hoff[64];  // spectrum offset of each line (fused pass)
  const int n = P.n, lpb = P.lpb, tot = lpb * n;
  float2* X = lds;
  // twiddles are read from the n-entry table in global memory (it stays in
  // the vector cache; staging it in LDS per workgroup cost more)
  const float2* twl = P.tw;
  const bool contiguous = P.stride == 1;
  // Strided passes: workgroups dispatched to one XCD (consecutive ids go to
  // the 8 XCDs in turn) take consecutive line groups, so the cache lines
  // that neighbouring columns share stay in that XCD's L2.
  long long blk = blockIdx.x;
  // (a bijection of the first 8 * floor(G / 8) workgroups; the rest keep
  // their own line group)
  if (!contiguous) {
    const long long G8 = (static_cast<long long>(gridDim.x) / 8) * 8;
    if (blk < G8) blk = (blk % 8) * (G8 / 8) + blk / 8;
  }
  const long long line0 = blk * lpb;
  const int nl = static_cast<int>(min(static_cast<long long>(lpb), P.lines - line0));
  if (threadIdx.x < lpb) {
    const long long L = line0 + threadIdx.x;
    const long long outer = L / P.inner, inr = L - outer * P.inner;
    base[threadIdx.x] = outer * n * P.stride + inr;
    hoff[threadIdx.x] = inr;
  }
  __syncthreads();
  auto split = [&](int e, int& l, int& k) {
    if (contiguous) {
      l = e / n;
      k = e - l * n;
    } else {
      k = e / lpb;
      l = e - k * lpb;
    }
  };
  {
    // every load of the thread is issued before the first LDS store
    float2 vv[SH_MAXV];
#pragma unroll
    for (int q = 0; q < SH_MAXV; ++q) {
      const int e = threadIdx.x + q * SH_THREADS;
      if (e < tot) {
        int l, k;
        split(e, l, k);
        float2 v = make_float2(0.0f, 0.0f);
        if (l < nl) {
          const long long g = base[l] + static_cast<long long>(k) * P.stride;
          v = P.in_real ? make_float2(static_cast<const float*>(in)[g], 0.0f)
                        : static_cast<const float2*>(in)[g];
        }
        vv[q] = v;
      }
    }
#pragma unroll
    for (int q = 0; q < SH_MAXV; ++q) {
      const int e = threadIdx.x + q * SH_THREADS;
      if (e < tot) {
        int l, k;
        split(e, l, k);
        X[l * n + k] = vv[q];
      }
    }
  }
  __syncthreads();
  stockham(X, n, P.log2n, lpb, twl, P.fused ? false : (P.inverse != 0));
  if (P.fused) {
    for (int e = threadIdx.x; e < tot; e += SH_THREADS) {
      const int l = e / n, k = e - l * n;
      if (l < nl) {
        float2 hv = P.h[static_cast<long long>(k) * P.stride + hoff[l]];
        if (P.conj_h) hv = conjf2(hv);
        X[e] = cmul(hv, X[e]);
      }
    }
    __syncthreads();
    stockham(X, n, P.log2n, lpb, twl, true);
  }
  for (int e = threadIdx.x; e < tot; e += SH_THREADS) {
    int l, k;
    split(e, l, k);
    if (l < nl) {
      const long long g = base[l] + static_cast<long long>(k) * P.stride;
      const float2 v = X[l * n + k];
      if (P.out_real)
        static_cast<float*>(out)[g] = __fmul_rn(v.x, scale);
      else
        static_cast<float2*>(out)[g] = make_float2(__fmul_rn(v.x, scale), __fmul_rn(v.y, scale));
    }
  }
}

// ------------------------------------------- real 2D convolution path --
//
// For two power-of-two FFT axes (F0, F1) the convolution of real data runs
// on half spectra.  Rows are transformed two at a time: z = x_a + i x_b
// has the DFT Z = X_a + i X_b with X_a, X_b Hermitian, so
//   X_a[k] = (Z[k] + conj(Z[-k])) / 2,   X_b[k] = (Z[k] - conj(Z[-k])) / (2 i),
// and only k = 0 .. F1/2 of each row is kept (H1 = F1/2 + 1 columns).  The
// columns of that half spectrum take the fused pass (forward, times the
// spectrum, inverse).  real(ifft2(H X)) = irfft2(Hs X) with the Hermitian
// part Hs[k] = (H[k] + conj(H[-k])) / 2 of H (equal to H for a real
// kernel), so the plan keeps Hs on the half grid.  The inverse rows rebuild
// Z = X_a + i X_b on the full row from the Hermitian extensions (the
// imaginary parts of the k = 0 and k = F1/2 bins dropped, as irfft does)
// and write x_a = Re z, x_b = Im z.

// Forward rows: real x (R rows of n) -> half spectra (R rows of n/2 + 1).
__global__ __launch_bounds__(SH_THREADS) void r2c_rows_kernel(const float* __restrict__ x,
                                                              float2* __restrict__ half, int n,
                                                              int log2n, long long R, int lpb,
                                                              const float2* __restrict__ tw,
                                                              const int32_t* pred) {
  if (pred_off(pred)) return;
  extern __shared__ float2 lds[];  // lpb * n values
  float2* X = lds;
  const int tot = lpb * n, H1 = n / 2 + 1;
  const long long row0 = 2ll * blockIdx.x * lpb;  // first real row of the workgroup
  {
    float2 vv[SH_MAXV];
#pragma unroll
    for (int q = 0; q < SH_MAXV; ++q) {
      const int e = threadIdx.x + q * SH_THREADS;
      if (e < tot) {
        const int l = e / n, k = e - l * n;
        const long long ra = row0 + 2 * l, rb = ra + 1;
        const float a = ra < R ? x[ra * n + k] : 0.0f;
        const float b = rb < R ? x[rb * n + k] : 0.0f;
        vv[q] = make_float2(a, b);
      }
    }
#pragma unroll
    for (int q = 0; q < SH_MAXV; ++q) {
      const int e = threadIdx.x + q * SH_THREADS;
      if (e < tot) X[e] = vv[q];
    }
  }
  __syncthreads();
  stockham(X, n, log2n, lpb, tw, false);
  for (int e = threadIdx.x; e < lpb * H1; e += SH_THREADS) {
    const int l = e / H1, k = e - l * H1;
    const long long ra = row0 + 2 * l, rb = ra + 1;
    const float2 zk = X[l * n + k], zm = X[l * n + ((n - k) & (n - 1))];
    if (ra < R)
      half[ra * H1 + k] = make_float2(__fmul_rn(0.5f, __fadd_rn(zk.x, zm.x)),
                                      __fmul_rn(0.5f, __fsub_rn(zk.y, zm.y)));
    if (rb < R)
      half[rb * H1 + k] = make_float2(__fmul_rn(0.5f, __fadd_rn(zk.y, zm.y)),
                                      __fmul_rn(0.5f, __fsub_rn(zm.x, zk.x)));
  }
}

// Inverse rows: half spectra (R rows of n/2 + 1) -> real y (R rows of n),
// times scale.
__global__ __launch_bounds__(SH_THREADS) void c2r_rows_kernel(const float2* __restrict__ half,
                                                              float* __restrict__ y, int n,
                                                              int log2n, long long R, int lpb,
                                                              const float2* __restrict__ tw,
                                                              float scale, const int32_t* pred) {
  if (pred_off(pred)) return;
  extern __shared__ float2 lds[];
  float2* X = lds;
  const int H1 = n / 2 + 1, nh = n / 2;
  const long long row0 = 2ll * blockIdx.x * lpb;
  for (int e = threadIdx.x; e < lpb * H1; e += SH_THREADS) {
    const int l = e / H1, k = e - l * H1;
    const long long ra = row0 + 2 * l, rb = ra + 1;
    float2 a = ra < R ? half[ra * H1 + k] : make_float2(0.0f, 0.0f);
    float2 b = rb < R ? half[rb * H1 + k] : make_float2(0.0f, 0.0f);
    if (k == 0 || k == nh) a.y = b.y = 0.0f;
    // Z[k] = A[k] + i B[k];  Z[n - k] = conj(A[k]) + i conj(B[k])
    X[l * n + k] = make_float2(__fsub_rn(a.x, b.y), __fadd_rn(a.y, b.x));
    if (k > 0 && k < nh) X[l * n + n - k] = make_float2(__fadd_rn(a.x, b.y), __fsub_rn(b.x, a.y));
  }
  __syncthreads();
  stockham(X, n, log2n, lpb, tw, true);
  for (int e = threadIdx.x; e < lpb * n; e += SH_THREADS) {
    const int l = e / n, k = e - l * n;
    const long long ra = row0 + 2 * l, rb = ra + 1;
    const float2 v = X[e];
    if (ra < R) y[ra * n + k] = __fmul_rn(v.x, scale);
    if (rb < R) y[rb * n + k] = __fmul_rn(v.y, scale);
  }
}

// ------------------------------- 1024-point lines on 256 threads --------
//
// Radix-4 Stockham transform of one 1024-point line by a 256-thread
// workgroup, four values per thread: thread j holds x[j + 256 r], r < 4,
// on entry and X[j + 256 r] on exit (natural order, no reordering pass).
// Stage s (Ns = 4^s, s < 5): twiddle a[r] by W_{4 Ns}^{r k}, k = j mod Ns
// (W = exp(-2 pi i / ...)), 4-point DFT over r, write a[r] to
// (j / Ns) 4 Ns + k + r Ns; the next stage reads positions j + 256 r.
// Stage 0 takes the registers as loaded and stage 4's outputs stay in
// registers, so a transform makes four LDS exchanges.  Four waves per line
// (2052 waves for the 513 columns of a 1024 x 1024 half spectrum) hide the
// latency that one wave per line leaves exposed.
constexpr int Q1K = 1024, QT = 256;
constexpr int Q_LDS = Q1K + Q1K / 16;  // padded line (float2)
__device__ __forceinline__ int q_pad(int i) { return i + (i >> 4); }

struct QTw {
  float2 w[4][3];  // stages 1..4, r = 1..3
};

__device__ __forceinline__ void q_twiddles(QTw& t, const float2* __restrict__ tw, int j) {
#pragma unroll
  for (int s = 1; s < 5; ++s) {
    const int Ns = 1 << (2 * s);
    const int k = j & (Ns - 1);
#pragma unroll
    for (int r = 1; r < 4; ++r) t.w[s - 1][r - 1] = tw[(r * k * (QT / Ns)) & (Q1K - 1)];
  }
}

// Transform of thread j's four values on the line buffer lds (every thread
// of the workgroup calls it: the barriers are workgroup wide).
template <bool INV>
__device__ __forceinline__ void q_fft(float2 (&a)[4], float2* lds, const QTw& t, int j) {
#pragma unroll
  for (int s = 0; s < 5; ++s) {
    const int Ns = 1 << (2 * s);
    if (s > 0) {
      __syncthreads();  // the previous stage's writes are visible
#pragma unroll
      for (int r = 0; r < 4; ++r) a[r] = lds[q_pad(j + QT * r)];
#pragma unroll
      for (int r = 1; r < 4; ++r)
        a[r] = cmul(a[r], INV ? conjf2(t.w[s - 1][r - 1]) : t.w[s - 1][r - 1]);
    }
    dft_small<4>(a, INV);
    if (s < 4) {
      const int base = (j >> (2 * s)) * 4 * Ns + (j & (Ns - 1));
      __syncthreads();  // every read of the buffer is done before it is overwritten
#pragma unroll
      for (int r = 0; r < 4; ++r) lds[q_pad(base + r * Ns)] = a[r];
    }
  }
}

// The 1024 x 1024 half spectrum is kept transposed, column k1 of every item
// contiguous: halfT[(item H1 + k1) 1024 + row], so that the column
// transforms read and write whole lines.  The row kernels take QR = 4 rows
// (two complex transforms, 512 threads) per workgroup and move each
// column's four values of those rows as two 16-byte pairs.
constexpr int QR = 4;
constexpr int QH1 = Q1K / 2 + 1;

// Row block of a row-kernel workgroup whose grid holds `first` other
// workgroups before its G row blocks: consecutive workgroups go to the 8
// XCDs in turn, so each XCD takes a run of neighbouring row blocks, whose
// pieces of one transposed column share cache lines (a bijection of the
// first 8 * floor(G / 8) row workgroups; the rest keep their own block).
__device__ __forceinline__ long long q_row_block(int first = 0) {
  long long b = static_cast<long long>(blockIdx.x) - first;
  const long long G8 = ((static_cast<long long>(gridDim.x) - first) / 8) * 8;
  if (b < G8) b = (b % 8) * (G8 / 8) + b / 8;
  return b;
}

// Forward rows: two real rows per complex transform (z = x_a + i x_b;
// X_a[k] = (Z[k] + conj Z[-k]) / 2, X_b[k] = (Z[k] - conj Z[-k]) / 2i).
__global__ __launch_bounds__(2 * QT) void q1k_r2c_kernel(const float* __restrict__ x,
                                                         float2* __restrict__ halfT,
                                                         const float2* __restrict__ tw,
                                                         const int32_t* pred) {
  if (pred_off(pred)) return;
  __shared__ float2 lds[2][Q_LDS];
  const int g = threadIdx.x >> 8, j = threadIdx.x & (QT - 1);
  const long long r0 = q_row_block() * QR;  // first row (all in one item)
  const long long ra = r0 + 2 * g;
  QTw t;
  q_twiddles(t, tw, j);
  float2 a[4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
    a[r] = make_float2(x[ra * Q1K + j + QT * r], x[(ra + 1) * Q1K + j + QT * r]);
  q_fft<false>(a, lds[g], t, j);
  __syncthreads();  // stage 4's reads are done
#pragma unroll
  for (int r = 0; r < 4; ++r) lds[g][q_pad(j + QT * r)] = a[r];
  __syncthreads();
  const long long item = r0 / Q1K;
  const int rl = static_cast<int>(r0 - item * Q1K);
  float4* out = reinterpret_cast<float4*>(halfT + item * QH1 * Q1K + rl);
  for (int e = threadIdx.x; e < 2 * QH1; e += 2 * QT) {
    const int k = e >> 1, h = e & 1;
    const float2 zk = lds[h][q_pad(k)], zm = lds[h][q_pad((Q1K - k) & (Q1K - 1))];
    out[(static_cast<long long>(k) * Q1K >> 1) + h] =
        make_float4(__fmul_rn(0.5f, __fadd_rn(zk.x, zm.x)), __fmul_rn(0.5f, __fsub_rn(zk.y, zm.y)),
                    __fmul_rn(0.5f, __fadd_rn(zk.y, zm.y)), __fmul_rn(0.5f, __fsub_rn(zm.x, zk.x)));
  }
}

// Columns (F0 = 1024) of the transposed half spectrum: forward, times hsT
// (conj for the adjoint), inverse, times 1/1024; one column per workgroup.
__global__ __launch_bounds__(QT) void q1k_cols_kernel(float2* __restrict__ halfT,
                                                      const float2* __restrict__ hsT, int conj_h,
                                                      const float2* __restrict__ tw,
                                                      const int32_t* pred) {
  if (pred_off(pred)) return;
  __shared__ float2 lds[Q_LDS];
  const int j = threadIdx.x;
  const int c = static_cast<int>(blockIdx.x % QH1);
  float2* col = halfT + static_cast<long long>(blockIdx.x) * Q1K;
  const float2* hc = hsT + static_cast<long long>(c) * Q1K;
  QTw t;
  q_twiddles(t, tw, j);
  float2 h[4], a[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const float2 hv = hc[j + QT * r];
    h[r] = conj_h ? conjf2(hv) : hv;
    a[r] = col[j + QT * r];
  }
  q_fft<false>(a, lds, t, j);
#pragma unroll
  for (int r = 0; r < 4; ++r) a[r] = cmul(h[r], a[r]);
  q_fft<true>(a, lds, t, j);
  constexpr float sc = 1.0f / Q1K;
#pragma unroll
  for (int r = 0; r < 4; ++r)
    col[j + QT * r] = make_float2(__fmul_rn(a[r].x, sc), __fmul_rn(a[r].y, sc));
}

// Inverse rows: the transposed half spectra of QR rows -> real rows, times
// 1/1024 (Z[k] = X_a[k] + i X_b[k] and its Hermitian extension, imaginary
// parts of k = 0 and 512 dropped as irfft does).  Loads the transform
// result of thread j of group g into a (pair of rows 2g, 2g + 1).
// Rows r0 .. r0 + QR - 1 (r0 = QR rb) of the half spectrum, inverse
// transformed: a holds the two real rows r0 + 2 g, r0 + 2 g + 1.
__device__ __forceinline__ void q1k_c2r_load(const float2* __restrict__ halfT, float2 (&lds)[2][Q_LDS],
                                             long long rb, const QTw& t, int g, int j,
                                             float2 (&a)[4]) {
  const long long r0 = rb * QR;
  const long long item = r0 / Q1K;
  const int rl = static_cast<int>(r0 - item * Q1K);
  const float4* in = reinterpret_cast<const float4*>(halfT + item * QH1 * Q1K + rl);
  constexpr int nh = Q1K / 2;
  for (int e = threadIdx.x; e < 2 * QH1; e += 2 * QT) {
    const int k = e >> 1, h = e & 1;
    const float4 v = in[(static_cast<long long>(k) * Q1K >> 1) + h];
    float2 aa = make_float2(v.x, v.y), bb = make_float2(v.z, v.w);
    if (k == 0 || k == nh) aa.y = bb.y = 0.0f;
    lds[h][q_pad(k)] = make_float2(__fsub_rn(aa.x, bb.y), __fadd_rn(aa.y, bb.x));
    if (k > 0 && k < nh)
      lds[h][q_pad(Q1K - k)] = make_float2(__fadd_rn(aa.x, bb.y), __fsub_rn(bb.x, aa.y));
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 4; ++r) a[r] = lds[g][q_pad(j + QT * r)];
  q_fft<true>(a, lds[g], t, j);
}

__global__ __launch_bounds__(2 * QT) void q1k_c2r_kernel(const float2* __restrict__ halfT,
                                                         float* __restrict__ y,
                                                         const float2* __restrict__ tw,
                                                         const int32_t* pred) {
  if (pred_off(pred)) return;
  __shared__ float2 lds[2][Q_LDS];
  const int g = threadIdx.x >> 8, j = threadIdx.x & (QT - 1);
  QTw t;
  q_twiddles(t, tw, j);
  float2 a[4];
  const long long rb = q_row_block();
  q1k_c2r_load(halfT, lds, rb, t, g, j, a);
  const long long ra = rb * QR + 2 * g;
  constexpr float sc = 1.0f / Q1K;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    y[ra * Q1K + j + QT * r] = __fmul_rn(a[r].x, sc);
    y[(ra + 1) * Q1K + j + QT * r] = __fmul_rn(a[r].y, sc);
  }
}

// ---------------------------------- fused AcceleratedPGM gradient step --
//
// For f(x) = 0.5 ||A x - y||^2 with A the 1024 x 1024 real circular
// convolution, one convolution per iteration serves both the gradient and
// the objective:
//   * G x (G = A^T A: the convolution by |Hs|^2, real on the half grid),
//     and G v = (1 + c) G x - c G x_prev for the extrapolated point
//     v = x + c (x - x_prev) (linearity), so u = v - (1/L) (G v - A^T y);
//   * ||A x - y||^2 = sum_k w_k |Hs[k] X[k] - Y[k]|^2 / (F0 F1) over the
//     half spectrum (Parseval; w_k = 1 on the first and the Nyquist column,
//     2 elsewhere), X and Y the spectra of x and y.
// All half-grid arrays are transposed (column k1 contiguous).

// Columns: forward, objective partial of the column (if obj), times gsT,
// inverse, times 1/1024.
__global__ __launch_bounds__(QT) void q1k_cols_gram_kernel(float2* __restrict__ halfT,
                                                           const float2* __restrict__ hsT,
                                                           const float* __restrict__ gsT,
                                                           const float2* __restrict__ YT,
                                                           double* __restrict__ obj,
                                                           const float2* __restrict__ tw,
                                                           const int32_t* pred) {
  if (pred_off(pred)) return;
  __shared__ float2 lds[Q_LDS];
  __shared__ double sh[4];
  const int j = threadIdx.x;
  const int c = blockIdx.x;
  const long long o = static_cast<long long>(c) * Q1K;
  float2* col = halfT + o;
  QTw t;
  q_twiddles(t, tw, j);
  float g[4];
  float2 h[4], yk[4], a[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    g[r] = gsT[o + j + QT * r];
    if (obj) {
      h[r] = hsT[o + j + QT * r];
      yk[r] = YT[o + j + QT * r];
    }
    a[r] = col[j + QT * r];
  }
  q_fft<false>(a, lds, t, j);
  if (obj) {
    double acc = 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float2 p = cmul(h[r], a[r]);
      const double dx = static_cast<double>(p.x) - static_cast<double>(yk[r].x);
      const double dy = static_cast<double>(p.y) - static_cast<double>(yk[r].y);
      acc += dx * dx + dy * dy;
    }
    acc = block_sum_f64(acc, sh);
    if (j == 0) obj[c] = (c == 0 || c == QH1 - 1) ? acc : 2.0 * acc;
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) a[r] = make_float2(__fmul_rn(g[r], a[r].x), __fmul_rn(g[r], a[r].y));
  q_fft<true>(a, lds, t, j);
  constexpr float sc = 1.0f / Q1K;
#pragma unroll
  for (int r = 0; r < 4; ++r)
    col[j + QT * r] = make_float2(__fmul_rn(a[r].x, sc), __fmul_rn(a[r].y, sc));
}

// Statistics of an AcceleratedPGM iteration (apgm_stats_block arguments).
struct ApgmStats {
  const double* obj;
  int nobj;
  double obj_scale;
  const double* tv;
  int ntv;
  double lam;
  const double* res;
  int nres;
  double* hist;
  int32_t* counter;
  int hist_cap;
};

// Inverse rows into G x, then the gradient step:
//   gv = (1 + c) gx_new - c gx_prev,  u = v + b (gv - aty)   (b = -1/L)
// with gx <- gx_new.  With STATS, workgroup 0 reduces the iteration's
// statistics row instead (from partials that the kernels before this one
// on the stream wrote) while workgroups 1 .. Q1K / QR take the row blocks,
// so the row needs no launch of its own.
template <bool STATS>
__global__ __launch_bounds__(2 * QT) void q1k_c2r_step_kernel(
    const float2* __restrict__ halfT, float* __restrict__ gx, const float* __restrict__ v,
    const float* __restrict__ aty, float c1, float c, float b, float* __restrict__ u,
    const float2* __restrict__ tw, ApgmStats st, const int32_t* pred) {
  if (pred_off(pred)) return;
  if (STATS && blockIdx.x == 0) {
    apgm_stats_block<2 * QT>(st.obj, st.nobj, st.obj_scale, st.tv, st.ntv, st.lam, st.res,
                             st.nres, st.hist, st.counter, st.hist_cap);
    return;
  }
  __shared__ float2 lds[2][Q_LDS];
  const int g = threadIdx.x >> 8, j = threadIdx.x & (QT - 1);
  QTw t;
  q_twiddles(t, tw, j);
  float2 a[4];
  const long long rb = q_row_block(STATS ? 1 : 0);
  q1k_c2r_load(halfT, lds, rb, t, g, j, a);
  const long long ra = rb * QR + 2 * g;
  constexpr float sc = 1.0f / Q1K;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
      const long long e = (ra + hh) * Q1K + j + QT * r;
      const float gn = __fmul_rn(hh ? a[r].y : a[r].x, sc);
      const float gv = __fsub_rn(__fmul_rn(c1, gn), __fmul_rn(c, gx[e]));
      u[e] = __fadd_rn(v[e], __fmul_rn(b, __fsub_rn(gv, aty[e])));
      gx[e] = gn;
    }
  }
}

// Hermitian part of the (F0, F1) spectrum on the half grid (F0, F1/2 + 1),
// stored transposed (hs[k1 F0 + k0]) if trans.
__global__ __launch_bounds__(256) void hermitian_half_kernel(const float2* __restrict__ h,
                                                             float2* __restrict__ hs, int F0,
                                                             int F1, int trans) {
  const int H1 = F1 / 2 + 1;
  const long long tot = static_cast<long long>(F0) * H1;
  for (long long e = blockIdx.x * 256ll + threadIdx.x; e < tot; e += 256ll * gridDim.x) {
    const int k0 = static_cast<int>(e / H1), k1 = static_cast<int>(e - static_cast<long long>(k0) * H1);
    const float2 a = h[static_cast<long long>(k0) * F1 + k1];
    const int m0 = (F0 - k0) % F0, m1 = (F1 - k1) % F1;
    const float2 b = h[static_cast<long long>(m0) * F1 + m1];
    hs[trans ? static_cast<long long>(k1) * F0 + k0 : e] =
        make_float2(__fmul_rn(0.5f, __fadd_rn(a.x, b.x)), __fmul_rn(0.5f, __fsub_rn(a.y, b.y)));
  }
}

// Elementwise complex product c = a * b (spectrum centring shift).
__global__ __launch_bounds__(256) void cmul_kernel(long long n, const float2* __restrict__ a,
                                                   const float2* __restrict__ b,
                                                   float2* __restrict__ c) {
  for (long long e = blockIdx.x * 256ll + threadIdx.x; e < n; e += 256ll * gridDim.x)
    c[e] = cmul(a[e], b[e]);
}

// ------------------------------------------------------------------ plan --

struct AxisTables {
  int n = 0, log2n = -1;
  float2* tw = nullptr;
};

struct Plan {
  int nd = 0;
  int F[3] = {1, 1, 1};
  long long batch = 1;
  long long fsize = 1;  // product of F
  AxisTables ax[3];
  float2* h = nullptr;     // spectrum (fsize)
  float2* work = nullptr;  // complex workspace (batch * fsize)
  // real 2D path (two power-of-two axes, F1 >= 4): Hermitian part of the
  // spectrum on the half grid (F0, F1/2 + 1) and lines per workgroup of the
  // row and column passes
  bool r2c = false;
  float2* hs = nullptr;
  int lpb_rows = 1, lpb_cols = 1;
};

// The 1024 x 1024 real path (transposed half spectrum, 256-thread lines).
bool q1k_path(const Plan& p) { return p.r2c && p.F[0] == Q1K && p.F[1] == Q1K; }

int log2_exact(int n) {
  if (n <= 0 || (n & (n - 1))) return -1;
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}

int make_tables(AxisTables& t, int n) {
  t.n = n;
  t.log2n = log2_exact(n);
  SCXR_REQUIRE(t.log2n >= 0 ? n <= MAX_POW2 : n <= MAX_DFT, SCXR_ERR_UNSUPPORTED,
               "FFT length %d unsupported (powers of two up to %d, other lengths up to %d)", n,
               MAX_POW2, MAX_DFT);
  const int cnt = n;  // exp(-2 pi i k / n), k < n
  std::vector<float2> tw(cnt);
  for (int k = 0; k < cnt; ++k) {
    const double a = -2.0 * M_PI * static_cast<double>(k) / static_cast<double>(n);
    tw[k] = make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
  }
  if (hipMalloc(&t.tw, sizeof(float2) * cnt) != hipSuccess)
    return fail(SCXR_ERR_NOMEM, "hipMalloc of FFT twiddles failed");
  SCXR_HIP(hipMemcpy(t.tw, tw.data(), sizeof(float2) * cnt, hipMemcpyHostToDevice));
  return SCXR_OK;
}

void destroy(Plan& p) {
  for (auto& a : p.ax)
    if (a.tw) (void)hipFree(a.tw), a.tw = nullptr;
  if (p.hs) (void)hipFree(p.hs), p.hs = nullptr;
  if (p.h) (void)hipFree(p.h), p.h = nullptr;
  if (p.work) (void)hipFree(p.work), p.work = nullptr;
}

// One pass along FFT axis a.
int run_pass(const Plan& p, int a, const void* in, void* out, bool in_real, bool out_real,
             bool inverse, bool fused, bool conj_h, const int32_t* pred, hipStream_t s) {
  Pass P{};
  const AxisTables& t = p.ax[a];
  P.n = t.n;
  P.log2n = t.log2n;
  long long stride = 1;
  for (int b = a + 1; b < p.nd; ++b) stride *= p.F[b];
  P.stride = stride;
  P.inner = stride;
  P.lines = p.batch * (p.fsize / t.n);
  if (t.log2n >= 2) {
    // Stockham: one LDS buffer of lpb lines (at most 8192 values); lines per
    // workgroup chosen so that a pass has at least ~1024 workgroups.
    const long long lines = p.batch * (p.fsize / t.n);
    int lpb = std::max(1, 8192 / t.n);
    { int q = 1; while (q * 2 <= lpb) q *= 2; lpb = q; }
    while (lpb > 1 && (lines + lpb - 1) / lpb < 1024) lpb /= 2;
    lpb = std::min(lpb, 64);
    P.lpb = lpb;
    P.in_real = in_real;
    P.out_real = out_real;
    P.inverse = inverse;
    P.fused = fused;
    P.conj_h = conj_h;
    P.tw = t.tw;
    P.h = p.h;
    const float scale = (inverse || fused) ? 1.0f / static_cast<float>(t.n) : 1.0f;
    const size_t lds = sizeof(float2) * lpb * t.n;
    const unsigned grid = static_cast<unsigned>((P.lines + lpb - 1) / lpb);
    hipLaunchKernelGGL(stockham_kernel, dim3(grid), dim3(SH_THREADS), lds, s, in, out, P, scale,
                       pred);
    SCXR_LAUNCHED("fft stockham pass");
    return SCXR_OK;
  }
  // Two LDS buffers for the fused and direct-DFT passes, one otherwise;
  // at most 64 KiB of LDS per workgroup in total.
  const bool two = fused || t.log2n < 0;
  const int cap = two ? LDS_ELEMS / 2 : LDS_ELEMS;
  SCXR_REQUIRE(t.n <= cap, SCXR_ERR_UNSUPPORTED,
               "FFT length %d too long for a %s pass (max %d)", t.n,
               fused ? "fused convolution" : "direct DFT", cap);
  int lpb = std::max(1, std::min(64, cap / t.n));
  if (t.log2n >= 0) {  // keep lpb a power of two
    int q = 1;
    while (q * 2 <= lpb) q *= 2;
    lpb = q;
  }
  P.lpb = lpb;
  P.in_real = in_real;
  P.out_real = out_real;
  P.inverse = inverse;
  P.fused = fused;
  P.conj_h = conj_h;
  P.tw = t.tw;
  P.h = p.h;
  const float scale = (inverse || fused) ? 1.0f / static_cast<float>(t.n) : 1.0f;
  const size_t lds = sizeof(float2) * (two ? 2 : 1) * lpb * t.n;
  const unsigned grid = static_cast<unsigned>((P.lines + lpb - 1) / lpb);
  hipLaunchKernelGGL(pass_kernel, dim3(grid), dim3(256), lds, s, in, out, P, scale, pred);
  SCXR_LAUNCHED("fft pass");
  return SCXR_OK;
}

// Real 2D path: rows forward (two real rows per complex FFT), fused column
// pass on the half spectrum, rows inverse.
int convolve_r2c(const Plan& p, const float* x, float* y, bool adj, const int32_t* pred,
                 hipStream_t s) {
  const AxisTables& t1 = p.ax[1];
  const int n = t1.n, H1 = n / 2 + 1;
  const long long R = p.batch * p.F[0];
  float2* half = p.work;  // R rows of H1 (fits in the batch * fsize workspace)
  const long long cl = (R + 1) / 2;  // complex row transforms
  if (q1k_path(p)) {
    const unsigned gr = static_cast<unsigned>(R / QR);
    hipLaunchKernelGGL(q1k_r2c_kernel, dim3(gr), dim3(2 * QT), 0, s, x, half, t1.tw, pred);
    SCXR_LAUNCHED("fft r2c rows (1024)");
    hipLaunchKernelGGL(q1k_cols_kernel, dim3(static_cast<unsigned>(p.batch * H1)), dim3(QT), 0, s,
                       half, p.hs, static_cast<int>(adj), p.ax[0].tw, pred);
    SCXR_LAUNCHED("fft half-spectrum columns (1024)");
    hipLaunchKernelGGL(q1k_c2r_kernel, dim3(gr), dim3(2 * QT), 0, s, half, y, t1.tw, pred);
    SCXR_LAUNCHED("fft c2r rows (1024)");
    return SCXR_OK;
  }
  int lr = p.lpb_rows;
  const unsigned gr = static_cast<unsigned>((cl + lr - 1) / lr);
  const size_t lds_r = sizeof(float2) * lr * n;
  hipLaunchKernelGGL(r2c_rows_kernel, dim3(gr), dim3(SH_THREADS), lds_r, s, x, half, n, t1.log2n,
                     R, lr, t1.tw, pred);
  SCXR_LAUNCHED("fft r2c rows");
  // columns of the half spectrum: n0-point transforms at stride H1
  const AxisTables& t0 = p.ax[0];
  Pass P{};
  P.n = t0.n;
  P.log2n = t0.log2n;
  P.stride = H1;
  P.inner = H1;
  P.lines = p.batch * H1;
  P.lpb = p.lpb_cols;
  P.fused = 1;
  P.conj_h = adj;
  P.tw = t0.tw;
  P.h = p.hs;
  const unsigned gc = static_cast<unsigned>((P.lines + P.lpb - 1) / P.lpb);
  hipLaunchKernelGGL(stockham_kernel, dim3(gc), dim3(SH_THREADS), sizeof(float2) * P.lpb * P.n, s,
                     half, half, P, 1.0f / static_cast<float>(t0.n), pred);
  SCXR_LAUNCHED("fft half-spectrum columns");
  hipLaunchKernelGGL(c2r_rows_kernel, dim3(gr), dim3(SH_THREADS), lds_r, s, half, y, n, t1.log2n,
                     R, lr, t1.tw, 1.0f / static_cast<float>(n), pred);
  SCXR_LAUNCHED("fft c2r rows");
  return SCXR_OK;
}

// y = real(ifftn(H' * fftn(x))) with H' = H or conj(H).
int convolve(const Plan& p, const float* x, float* y, bool adj, const int32_t* pred,
             hipStream_t s) {
  if (p.r2c) return convolve_r2c(p, x, y, adj, pred, s);
  if (p.nd == 1) return run_pass(p, 0, x, y, true, true, false, true, adj, pred, s);
  int rc;
  // forward along F{d-1} .. F1 (first pass reads the real input)
  for (int a = p.nd - 1, first = 1; a >= 1; --a, first = 0) {
    rc = run_pass(p, a, first ? static_cast<const void*>(x) : p.work, p.work, first, false, false,
                  false, false, pred, s);
    if (rc) return rc;
  }
  rc = run_pass(p, 0, p.work, p.work, false, false, false, true, adj, pred, s);
  if (rc) return rc;
  for (int a = 1; a < p.nd; ++a) {
    const bool last = (a == p.nd - 1);
    rc = run_pass(p, a, p.work, last ? static_cast<void*>(y) : p.work, false, last, true, false,
                  false, pred, s);
    if (rc) return rc;
  }
  return SCXR_OK;
}

// Complex fftn / ifftn of x (complex or real input) into out (complex).
int fftn(const Plan& p, const void* x, bool x_real, float2* out, bool inverse, const int32_t* pred,
         hipStream_t s) {
  for (int a = p.nd - 1, first = 1; a >= 0; --a, first = 0) {
    const int rc = run_pass(p, a, first ? x : static_cast<const void*>(out), out,
                            first && x_real, false, inverse, false, false, pred, s);
    if (rc) return rc;
  }
  return SCXR_OK;
}

}  // namespace fft
}  // namespace scxr

struct scxr_cconv_plan {
  scxr::fft::Plan p;
};

using namespace scxr::fft;

extern "C" {

int scxr_cconv_plan_create(int32_t ndim, const int32_t* fshape, int64_t batch,
                           scxr_cconv_plan** out) {
  SCXR_REQUIRE(out && fshape && ndim >= 1 && ndim <= 3 && batch >= 1, SCXR_ERR_ARG,
               "bad cconv plan arguments");
  auto* plan = new scxr_cconv_plan();
  Plan& p = plan->p;
  p.nd = ndim;
  p.batch = batch;
  p.fsize = 1;
  int rc = SCXR_OK;
  for (int a = 0; a < ndim && rc == SCXR_OK; ++a) {
    if (fshape[a] < 1) rc = fail(SCXR_ERR_SHAPE, "non-positive FFT length");
    if (rc == SCXR_OK) {
      p.F[a] = fshape[a];
      p.fsize *= fshape[a];
      rc = make_tables(p.ax[a], fshape[a]);
    }
  }
  if (rc == SCXR_OK && hipMalloc(&p.h, sizeof(float2) * p.fsize) != hipSuccess)
    rc = fail(SCXR_ERR_NOMEM, "hipMalloc of the spectrum failed");
  if (rc == SCXR_OK && hipMalloc(&p.work, sizeof(float2) * p.fsize * p.batch) != hipSuccess)
    rc = fail(SCXR_ERR_NOMEM, "hipMalloc of the FFT workspace failed");
  if (rc == SCXR_OK && ndim == 2 && p.ax[0].log2n >= 2 && p.ax[1].log2n >= 2) {
    p.r2c = true;
    const long long H1 = p.F[1] / 2 + 1;
    if (hipMalloc(&p.hs, sizeof(float2) * p.F[0] * H1) != hipSuccess)
      rc = fail(SCXR_ERR_NOMEM, "hipMalloc of the half spectrum failed");
    // lines per workgroup: as many as fit 8192 LDS values while the pass
    // keeps at least 512 workgroups (SCXR_FFT_LPB="rows,cols" overrides)
    auto pick = [](long long lines, int n) {
      int l = std::max(1, 8192 / n);
      while (l > 1 && (lines + l - 1) / l < 512) l /= 2;
      return l;
    };
    p.lpb_rows = pick((p.batch * p.F[0] + 1) / 2, p.F[1]);
    p.lpb_cols = pick(p.batch * H1, p.F[0]);
    if (const char* env = std::getenv("SCXR_FFT_LPB")) {
      int a = 0, b = 0;
      if (std::sscanf(env, "%d,%d", &a, &b) == 2 && a >= 1 && b >= 1 &&
          a * p.F[1] <= 8192 && b * p.F[0] <= 8192) {
        p.lpb_rows = a;
        p.lpb_cols = b;
      }
    }
  }
  if (rc != SCXR_OK) {
    destroy(p);
    delete plan;
    return rc;
  }
  *out = plan;
  return SCXR_OK;
}

int scxr_cconv_plan_destroy(scxr_cconv_plan* plan) {
  if (!plan) return SCXR_OK;
  destroy(plan->p);
  delete plan;
  return SCXR_OK;
}

int scxr_cconv_set_spectrum(scxr_cconv_plan* plan, const float* h, int32_t h_is_dft,
                            const float* shift, void* stream) {
  SCXR_REQUIRE(plan && h, SCXR_ERR_ARG, "null argument to cconv_set_spectrum");
  Plan& p = plan->p;
  const hipStream_t s = as_stream(stream);
  if (h_is_dft) {
    SCXR_HIP(hipMemcpyAsync(p.h, h, sizeof(float2) * p.fsize, hipMemcpyDeviceToDevice, s));
  } else {
    // h: real, zero-padded to the FFT shape (batch 1)
    Plan q = p;
    q.batch = 1;
    const int rc = fftn(q, h, true, p.h, false, nullptr, s);
    if (rc) return rc;
  }
  if (shift) {
    hipLaunchKernelGGL(cmul_kernel, dim3(256), dim3(256), 0, s, p.fsize, p.h,
                       reinterpret_cast<const float2*>(shift), p.h);
    SCXR_LAUNCHED("cconv shift");
  }
  if (p.r2c) {
    hipLaunchKernelGGL(hermitian_half_kernel, dim3(256), dim3(256), 0, s, p.h, p.hs, p.F[0],
                       p.F[1], static_cast<int>(q1k_path(p)));
    SCXR_LAUNCHED("cconv half spectrum");
  }
  return SCXR_OK;
}

int scxr_cconv_spectrum(const scxr_cconv_plan* plan, float* out, void* stream) {
  SCXR_REQUIRE(plan && out, SCXR_ERR_ARG, "null argument to cconv_spectrum");
  SCXR_HIP(hipMemcpyAsync(out, plan->p.h, sizeof(float2) * plan->p.fsize,
                          hipMemcpyDeviceToDevice, as_stream(stream)));
  return SCXR_OK;
}

int scxr_cconv_apply(const scxr_cconv_plan* plan, const float* x, float* y, int32_t adjoint,
                     const int32_t* pred, void* stream) {
  SCXR_REQUIRE(plan && x && y, SCXR_ERR_ARG, "null argument to cconv_apply");
  return convolve(plan->p, x, y, adjoint != 0, pred, as_stream(stream));
}

}  // extern "C"

// The three kernels of one fused gradient step; st (nullable) adds the
// statistics row to the last of them.
static int apgm_grad(const scxr_cconv_plan* plan, const float* x, const float* gs,
                     const float* Yh, double* obj, float* gx, const float* v, const float* aty,
                     float c, float b, float* u, const ApgmStats* st, const int32_t* pred,
                     hipStream_t s) {
  SCXR_REQUIRE(plan && x && gs && gx && v && aty && u && (!obj || Yh), SCXR_ERR_ARG,
               "null argument to cconv_apgm_grad");
  const Plan& p = plan->p;
  SCXR_REQUIRE(q1k_path(p) && p.batch == 1, SCXR_ERR_UNSUPPORTED,
               "fused APGM gradient needs a single 1024 x 1024 real convolution plan");
  constexpr int H1 = Q1K / 2 + 1;
  float2* half = p.work;
  hipLaunchKernelGGL(q1k_r2c_kernel, dim3(Q1K / QR), dim3(2 * QT), 0, s, x, half, p.ax[1].tw,
                     pred);
  SCXR_LAUNCHED("fft r2c rows (1024)");
  hipLaunchKernelGGL(q1k_cols_gram_kernel, dim3(H1), dim3(QT), 0, s, half, p.hs, gs,
                     reinterpret_cast<const float2*>(Yh), obj, p.ax[0].tw, pred);
  SCXR_LAUNCHED("fft gram columns (1024)");
  if (st) {
    hipLaunchKernelGGL(q1k_c2r_step_kernel<true>, dim3(Q1K / QR + 1), dim3(2 * QT), 0, s, half,
                       gx, v, aty, 1.0f + c, c, b, u, p.ax[1].tw, *st, pred);
    SCXR_LAUNCHED("fft c2r gradient step + statistics row (1024)");
  } else {
    hipLaunchKernelGGL(q1k_c2r_step_kernel<false>, dim3(Q1K / QR), dim3(2 * QT), 0, s, half, gx,
                       v, aty, 1.0f + c, c, b, u, p.ax[1].tw, ApgmStats{}, pred);
    SCXR_LAUNCHED("fft c2r gradient step (1024)");
  }
  return SCXR_OK;
}

extern "C" {

int scxr_cconv_apgm_grad(const scxr_cconv_plan* plan, const float* x, const float* gs,
                         const float* Yh, double* obj, float* gx, const float* v,
                         const float* aty, float c, float b, float* u, const int32_t* pred,
                         void* stream) {
  return apgm_grad(plan, x, gs, Yh, obj, gx, v, aty, c, b, u, nullptr, pred, as_stream(stream));
}

int scxr_cconv_apgm_grad_itstat(const scxr_cconv_plan* plan, const float* x, const float* gs,
                                const float* Yh, double* obj, float* gx, const float* v,
                                const float* aty, float c, float b, float* u, double obj_scale,
                                const double* tv, int32_t ntv, double lam, const double* res,
                                int32_t nres, double* hist, int32_t* counter, int32_t hist_cap,
                                const int32_t* pred, void* stream) {
  SCXR_REQUIRE(obj && Yh && tv && res && hist && counter && ntv > 0 && nres > 0 && hist_cap > 0,
               SCXR_ERR_ARG, "bad cconv_apgm_grad_itstat arguments");
  const ApgmStats st{obj, Q1K / 2 + 1, obj_scale, tv, ntv, lam, res, nres, hist, counter,
                     hist_cap};
  return apgm_grad(plan, x, gs, Yh, obj, gx, v, aty, c, b, u, &st, pred, as_stream(stream));
}

int scxr_fftn(const scxr_cconv_plan* plan, const float* x, int32_t x_real, float* out,
              int32_t inverse, const int32_t* pred, void* stream) {
  SCXR_REQUIRE(plan && x && out, SCXR_ERR_ARG, "null argument to fftn");
  return fftn(plan->p, x, x_real != 0, reinterpret_cast<float2*>(out), inverse != 0, pred,
              as_stream(stream));
}

}  // extern "C"
