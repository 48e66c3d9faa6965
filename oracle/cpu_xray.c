/*
 * C / OpenMP restatement of the 2D parallel-beam projector
 * (scico/linop/xray/_xray2d.py:223-351).
 *
 * TEST INFRASTRUCTURE ONLY: the CPU baseline of bench.py and a second,
 * independent implementation checked bit-for-bit against oracle/xray.py
 * (tests/test_oracle_cpu.py).  Compiled with -O3 -fno-fast-math
 * -ffp-contract=off so every float32 operation rounds as written.
 *
 * Forward: scatter-add per view (_xray2d.py:254-263), views split across
 * threads (each thread owns whole sinogram rows).  Adjoint: gather per
 * pixel (_xray2d.py:293-306), rows split across threads.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

typedef struct {
  float Pxmin, P0, P1, width;
} view_t;

/* _xray2d.py:326-347 (float32, reference operation order). */
static void views(int A, const float* angles, const double* x0, const double* dx, double y0,
                  view_t* v) {
  const float x00 = (float)x0[0], x01 = (float)x0[1];
  const float dx0 = (float)dx[0], dx1 = (float)dx[1];
  const float y0f = (float)y0;
  for (int a = 0; a < A; ++a) {
    const float u0 = (float)cos((double)angles[a]);
    const float u1 = (float)sin((double)angles[a]);
    const float t0 = x00 * u0, t1 = x01 * u1;
    const float Px0 = (t0 + t1) - y0f;
    const float P0 = dx0 * u0, P1 = dx1 * u1;
    const float c1 = Px0 + P0, c2 = Px0 + P1, c3 = c1 + P1;
    float m = Px0;
    if (c1 < m) m = c1;
    if (c2 < m) m = c2;
    if (c3 < m) m = c3;
    const float d1 = fabsf(P0 + P1), d2 = fabsf(P0 - P1);
    const float w = d1 > d2 ? d1 : d2, f = d1 > d2 ? d2 : d1;
    v[a].Pxmin = m;
    v[a].P0 = P0;
    v[a].P1 = P1;
    v[a].width = (w + f) / 2.0f;
  }
}

static inline void weight(const view_t* v, float fi, float fj, int* ind, float* w) {
  const float Px = (v->Pxmin + v->P0 * fi) + v->P1 * fj;
  const float fl = floorf(Px);
  float d = 1.0f - (Px - fl);
  if (v->width < d) d = v->width;
  *ind = (int)fl;
  *w = d / v->width;
}

int cpu_xr2d_weights(int M, int N, int A, const double* x0, const double* dx, double y0,
                     const float* angles, int32_t* inds, float* wts) {
  view_t* v = (view_t*)malloc(sizeof(view_t) * A);
  if (!v) return -1;
  views(A, angles, x0, dx, y0, v);
#pragma omp parallel for schedule(static)
  for (int a = 0; a < A; ++a)
    for (int i = 0; i < M; ++i)
      for (int j = 0; j < N; ++j) {
        const size_t e = ((size_t)a * M + i) * N + j;
        int ind;
        float w;
        weight(&v[a], (float)i, (float)j, &ind, &w);
        inds[e] = ind;
        wts[e] = w;
      }
  free(v);
  return 0;
}

int cpu_xr2d_forward(int M, int N, int A, int D, const double* x0, const double* dx, double y0,
                     const float* angles, const float* im, float* sino) {
  view_t* v = (view_t*)malloc(sizeof(view_t) * A);
  if (!v) return -1;
  views(A, angles, x0, dx, y0, v);
#pragma omp parallel for schedule(dynamic, 1)
  for (int a = 0; a < A; ++a) {
    float* y = sino + (size_t)a * D;
    for (int b = 0; b < D; ++b) y[b] = 0.0f;
    /* first scatter: bin ind with weight w (_xray2d.py:256-259) */
    for (int i = 0; i < M; ++i)
      for (int j = 0; j < N; ++j) {
        int ind;
        float w;
        weight(&v[a], (float)i, (float)j, &ind, &w);
        if (ind >= 0 && ind < D) y[ind] += im[(size_t)i * N + j] * w;
      }
    /* second scatter: bin ind + 1 with weight 1 - w (_xray2d.py:262-263) */
    for (int i = 0; i < M; ++i)
      for (int j = 0; j < N; ++j) {
        int ind;
        float w;
        weight(&v[a], (float)i, (float)j, &ind, &w);
        if (ind + 1 >= 0 && ind + 1 < D) y[ind + 1] += im[(size_t)i * N + j] * (1.0f - w);
      }
  }
  free(v);
  return 0;
}

int cpu_xr2d_adjoint(int M, int N, int A, int D, const double* x0, const double* dx, double y0,
                     const float* angles, const float* sino, float* im) {
  view_t* v = (view_t*)malloc(sizeof(view_t) * A);
  if (!v) return -1;
  views(A, angles, x0, dx, y0, v);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < N; ++j) {
      float s0 = 0.0f, s1 = 0.0f;
      for (int a = 0; a < A; ++a) {
        int ind;
        float w;
        weight(&v[a], (float)i, (float)j, &ind, &w);
        const float* y = sino + (size_t)a * D;
        if (ind >= 0 && ind < D) s0 += y[ind] * w;
        if (ind + 1 >= 0 && ind + 1 < D) s1 += y[ind + 1] * (1.0f - w);
      }
      im[(size_t)i * N + j] = s0 + s1;
    }
  free(v);
  return 0;
}

/* The same projections with float64 accumulation (each sum rounded to
 * float32 once): another valid summation order of the reference's sums,
 * used to size the rounding-noise floor of whole ADMM iterations. */
int cpu_xr2d_forward_f64acc(int M, int N, int A, int D, const double* x0, const double* dx,
                            double y0, const float* angles, const float* im, float* sino) {
  view_t* v = (view_t*)malloc(sizeof(view_t) * A);
  if (!v) return -1;
  views(A, angles, x0, dx, y0, v);
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 1)
  for (int a = 0; a < A; ++a) {
    double* y = (double*)calloc((size_t)D, sizeof(double));
    if (!y) {
      bad = 1;
      continue;
    }
    for (int i = 0; i < M; ++i)
      for (int j = 0; j < N; ++j) {
        int ind;
        float w;
        weight(&v[a], (float)i, (float)j, &ind, &w);
        const float p = im[(size_t)i * N + j];
        if (ind >= 0 && ind < D) y[ind] += (double)(p * w);
        if (ind + 1 >= 0 && ind + 1 < D) y[ind + 1] += (double)(p * (1.0f - w));
      }
    for (int b = 0; b < D; ++b) sino[(size_t)a * D + b] = (float)y[b];
    free(y);
  }
  free(v);
  return bad ? -1 : 0;
}

int cpu_xr2d_adjoint_f64acc(int M, int N, int A, int D, const double* x0, const double* dx,
                            double y0, const float* angles, const float* sino, float* im) {
  view_t* v = (view_t*)malloc(sizeof(view_t) * A);
  if (!v) return -1;
  views(A, angles, x0, dx, y0, v);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < N; ++j) {
      double s0 = 0.0, s1 = 0.0;
      for (int a = 0; a < A; ++a) {
        int ind;
        float w;
        weight(&v[a], (float)i, (float)j, &ind, &w);
        const float* y = sino + (size_t)a * D;
        if (ind >= 0 && ind < D) s0 += (double)(y[ind] * w);
        if (ind + 1 >= 0 && ind + 1 < D) s1 += (double)(y[ind + 1] * (1.0f - w));
      }
      im[(size_t)i * N + j] = (float)(s0 + s1);
    }
  free(v);
  return 0;
}

/*
 * 3D parallel-beam projector (scico/linop/xray/_xray3d.py:110-266) for a
 * separable geometry: in every view, detector row ri depends on the two
 * in-plane volume axes only (coefficient 0 on the slice axis s) and row
 * rs = 1 - ri on the slice axis only (rotation about axis s).  Every
 * voxel's four taps are the reference's (calc_weights_3d, :206-266),
 * computed with the general expression; the in-plane coordinate and
 * fraction of row ri are evaluated once per (pixel, view) and those of row
 * rs once per slice, which the zero coefficients make exact.  Sums follow
 * the reference order: forward, per view, tap (0,0), (1,0), (0,1), (1,1)
 * over all voxels in C order (np.add.at); adjoint, per voxel, views in
 * order and the four taps in that order (out = out + y * w).  f64acc
 * accumulates in float64 and rounds once.
 */
typedef struct {
  int ul;
  float t;
} tap1_t;

static inline tap1_t tap_row(const float* m, float x0, float x1, float x2) {
  const float P = ((m[0] * x0 + m[1] * x1) + m[2] * x2) + m[3];
  const float left = P - 0.25f;
  float t = ceilf(left) - left;
  if (t > 0.5f) t = 0.5f;
  tap1_t r;
  r.t = t;
  r.ul = (int)floorf(left);
  return r;
}

typedef struct {
  int X, Y, Z, s, ri, D0, D1;
} sep_t;

/* taps of row ri for every in-plane pixel (a, b) (a, b = the non-slice
 * axes in order) and of row rs for every slice, for view matrix m. */
static void sep_taps(const sep_t* g, const float* m, tap1_t* tin, tap1_t* ts) {
  const int n[3] = {g->X, g->Y, g->Z};
  const int pa = g->s == 0 ? 1 : 0, qa = g->s == 2 ? 1 : 2;
  const float* mi = m + 4 * g->ri;
  const float* ms = m + 4 * (1 - g->ri);
  for (int a = 0; a < n[pa]; ++a)
    for (int b = 0; b < n[qa]; ++b) {
      float x[3];
      x[g->s] = 0.5f;
      x[pa] = (float)a + 0.5f;
      x[qa] = (float)b + 0.5f;
      tin[(size_t)a * n[qa] + b] = tap_row(mi, x[0], x[1], x[2]);
    }
  for (int c = 0; c < n[g->s]; ++c) {
    float x[3] = {0.5f, 0.5f, 0.5f};
    x[g->s] = (float)c + 0.5f;
    ts[c] = tap_row(ms, x[0], x[1], x[2]);
  }
}

/* strides taking voxel (i, j, k) to its in-plane pixel and its slice */
static void sep_strides(const sep_t* g, size_t st[3], int sl[3]) {
  const int n[3] = {g->X, g->Y, g->Z};
  const int pa = g->s == 0 ? 1 : 0, qa = g->s == 2 ? 1 : 2;
  for (int a = 0; a < 3; ++a) {
    st[a] = a == pa ? (size_t)n[qa] : (a == qa ? 1u : 0u);
    sl[a] = a == g->s;
  }
}

static const int TAP_D0[4] = {0, 1, 0, 1}, TAP_D1[4] = {0, 0, 1, 1};

/* weight of tap tp (zero when its detector pixel is off the detector,
 * which the caller tests) */
static inline float sep_tap_weight(const sep_t* g, tap1_t a_in, tap1_t a_s, int tp, int* d0,
                                   int* d1) {
  const tap1_t r0 = g->ri == 0 ? a_in : a_s, r1 = g->ri == 0 ? a_s : a_in;
  const float h = 0.5f;
  const float f0 = TAP_D0[tp] ? h - r0.t : r0.t, f1 = TAP_D1[tp] ? h - r1.t : r1.t;
  *d0 = r0.ul + TAP_D0[tp];
  *d1 = r1.ul + TAP_D1[tp];
  return (f0 * f1) * 4.0f;
}

int cpu_xr3d_sep_forward(int X, int Y, int Z, int V, int D0, int D1, const float* mats, int s,
                         int ri, int f64acc, const float* vol, float* proj) {
  const sep_t g = {X, Y, Z, s, ri, D0, D1};
  const int n[3] = {X, Y, Z};
  const int pa = s == 0 ? 1 : 0, qa = s == 2 ? 1 : 2;
  const size_t npix = (size_t)n[pa] * n[qa];
  size_t st[3];
  int sl[3];
  sep_strides(&g, st, sl);
  int err = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
  for (int v = 0; v < V; ++v) {
    tap1_t* tin = (tap1_t*)malloc(sizeof(tap1_t) * npix);
    tap1_t* ts = (tap1_t*)malloc(sizeof(tap1_t) * n[s]);
    double* acc = f64acc ? (double*)calloc((size_t)D0 * D1, sizeof(double)) : NULL;
    float* out = proj + (size_t)v * D0 * D1;
    if (!tin || !ts || (f64acc && !acc)) {
      err = 1;
    } else {
      sep_taps(&g, mats + 8 * v, tin, ts);
      for (size_t e = 0; e < (size_t)D0 * D1; ++e) out[e] = 0.0f;
      /* np.add.at order: tap by tap, voxels in C order; a zero weight adds
       * nothing and is skipped */
      for (int tp = 0; tp < 4; ++tp)
        for (int i = 0; i < X; ++i)
          for (int j = 0; j < Y; ++j) {
            const size_t pij = i * st[0] + j * st[1];
            const int sij = i * sl[0] + j * sl[1];
            const float* xr = vol + ((size_t)i * Y + j) * Z;
            for (int k = 0; k < Z; ++k) {
              int d0, d1;
              const float w = sep_tap_weight(&g, tin[pij + k * st[2]], ts[sij + k * sl[2]], tp,
                                             &d0, &d1);
              if (w == 0.0f || d0 < 0 || d0 >= D0 || d1 < 0 || d1 >= D1) continue;
              const size_t o = (size_t)d0 * D1 + d1;
              if (f64acc)
                acc[o] += (double)w * (double)xr[k];
              else
                out[o] = out[o] + w * xr[k];
            }
          }
      if (f64acc)
        for (size_t e = 0; e < (size_t)D0 * D1; ++e) out[e] = (float)acc[e];
    }
    free(tin);
    free(ts);
    free(acc);
  }
  return err ? -1 : 0;
}

int cpu_xr3d_sep_adjoint(int X, int Y, int Z, int V, int D0, int D1, const float* mats, int s,
                         int ri, int f64acc, const float* proj, float* vol) {
  const sep_t g = {X, Y, Z, s, ri, D0, D1};
  const int n[3] = {X, Y, Z};
  const int pa = s == 0 ? 1 : 0, qa = s == 2 ? 1 : 2;
  const size_t npix = (size_t)n[pa] * n[qa];
  tap1_t* tin = (tap1_t*)malloc(sizeof(tap1_t) * npix * V);
  tap1_t* ts = (tap1_t*)malloc(sizeof(tap1_t) * n[s] * V);
  if (!tin || !ts) {
    free(tin);
    free(ts);
    return -1;
  }
#pragma omp parallel for schedule(static)
  for (int v = 0; v < V; ++v) sep_taps(&g, mats + 8 * v, tin + npix * v, ts + (size_t)n[s] * v);
  /* per in-plane pixel: every view in order, the slices in the inner
   * loop (each voxel still sums its views in order) */
  const size_t vst[3] = {(size_t)Y * Z, (size_t)Z, 1};
  int err = 0;
#pragma omp parallel for schedule(static) reduction(| : err)
  for (int a = 0; a < n[pa]; ++a) {
    float* accf = (float*)malloc(sizeof(float) * n[s]);
    double* accd = (double*)malloc(sizeof(double) * n[s]);
    if (!accf || !accd) {
      err = 1;
    } else {
      for (int b = 0; b < n[qa]; ++b) {
        const size_t pix = (size_t)a * n[qa] + b;
        for (int c = 0; c < n[s]; ++c) {
          accf[c] = 0.0f;
          accd[c] = 0.0;
        }
        for (int v = 0; v < V; ++v) {
          const float* y = proj + (size_t)v * D0 * D1;
          const tap1_t a_in = tin[npix * v + pix];
          const tap1_t* tsv = ts + (size_t)n[s] * v;
          for (int c = 0; c < n[s]; ++c)
            /* out = out + y * w per tap, in tap order; a tap whose pixel
             * is off the detector has weight 0 (the reference gathers a
             * clipped pixel times 0), and a zero weight adds nothing */
            for (int tp = 0; tp < 4; ++tp) {
              int d0, d1;
              const float w = sep_tap_weight(&g, a_in, tsv[c], tp, &d0, &d1);
              if (w == 0.0f || d0 < 0 || d0 >= D0 || d1 < 0 || d1 >= D1) continue;
              const float yv = y[(size_t)d0 * D1 + d1];
              if (f64acc)
                accd[c] += (double)yv * (double)w;
              else
                accf[c] = accf[c] + yv * w;
            }
        }
        for (int c = 0; c < n[s]; ++c)
          vol[a * vst[pa] + b * vst[qa] + c * vst[s]] = f64acc ? (float)accd[c] : accf[c];
      }
    }
    free(accf);
    free(accd);
  }
  free(tin);
  free(ts);
  return err ? -1 : 0;
}
