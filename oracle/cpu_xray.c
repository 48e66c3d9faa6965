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
