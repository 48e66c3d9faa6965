/*
 * scico_amd C ABI: hand-written gfx950 (MI355X) kernels for the
 * tomographic hot path of lanl/scico.
 *
 * Conventions
 *   - Every entry point is extern "C" with plain C types.  Arrays are
 *     caller-owned, contiguous, row-major device pointers (float32 unless
 *     stated); host pointers are marked "host".
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *     All device work is stream-ordered; no entry point synchronises the
 *     device, and hot-path calls allocate nothing (plans allocate their
 *     device tables and workspace once, at creation).
 *   - `pred` (nullable) is a device int32 predicate: while *pred == 0 the
 *     launched kernels return immediately.  NULL means "always run".  It
 *     lets callers replay a fixed-length iteration (e.g. CG with a
 *     device-side stopping test) inside a captured graph.
 *   - Return value: SCXR_OK (0) or a negative status; scxr_last_error()
 *     returns a thread-local message for the last failure.
 *   - A plan is immutable after creation; calls that use the same plan's
 *     workspace must be ordered on one stream.
 *
 * Reference citations are relative to the lanl/scico source tree.
 */
#ifndef SCICO_AMD_H
#define SCICO_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SCXR_VERSION 1

enum {
  SCXR_OK = 0,
  SCXR_ERR_ARG = -1,         /* invalid argument (null pointer, bad size)   */
  SCXR_ERR_SHAPE = -2,       /* shapes do not conform                        */
  SCXR_ERR_GEOM = -3,        /* unsupported projection geometry              */
  SCXR_ERR_HIP = -4,         /* a HIP runtime call or launch failed          */
  SCXR_ERR_NOMEM = -5,       /* device allocation failed (plan creation)     */
  SCXR_ERR_UNSUPPORTED = -6  /* feature not available for these parameters  */
};

/* ------------------------------------------------------------ library -- */

int scxr_version(void);
/* Thread-local message describing the last failure on this thread. */
const char* scxr_last_error(void);
/* Device properties of the current HIP device (any pointer may be NULL). */
int scxr_device_info(int32_t* cu_count, int32_t* warp_size, int64_t* lds_bytes,
                     char* arch_name, int32_t arch_len);

/* ------------------------------------------------ 2D parallel-beam X-ray --
 * Replaces scico.linop.xray.XRayTransform2D:
 *   constructor geometry   scico/linop/xray/_xray2d.py:86-136
 *   _calc_weights          scico/linop/xray/_xray2d.py:309-351
 *   project / _project     scico/linop/xray/_xray2d.py:138-146, 223-265
 *   back_project           scico/linop/xray/_xray2d.py:148-156, 267-306
 * Image (M, N) row-major, sinogram (A, D) row-major, D = det_count.
 * x0, dx, y0 are the constructor values (float64 on the host, rounded to
 * float32 as at the reference's jit boundary); angles are float32.
 */
typedef struct scxr_xr2d_plan scxr_xr2d_plan;

int scxr_xr2d_plan_create(int32_t M, int32_t N, int32_t A, int32_t D,
                          const double* x0 /* host, 2 */,
                          const double* dx /* host, 2 */, double y0,
                          const float* angles /* host, A */,
                          scxr_xr2d_plan** out);
int scxr_xr2d_plan_destroy(scxr_xr2d_plan* plan);

/* Per-view float32 geometry (host, A x 4): Pxmin, P0, P1, width
 * (_xray2d.py:326-347), for bit-exact checks. */
int scxr_xr2d_view_table(const scxr_xr2d_plan* plan, float* out);

/* sino = A im  (_xray2d.py:223-265).  Deterministic; no global atomics. */
int scxr_xr2d_forward(const scxr_xr2d_plan* plan, const float* im, float* sino,
                      const int32_t* pred, void* stream);

/* im = A^T sino  (_xray2d.py:267-306), gather form, exact adjoint of the
 * forward's weights. */
int scxr_xr2d_adjoint(const scxr_xr2d_plan* plan, const float* sino, float* im,
                      const int32_t* pred, void* stream);

/* Debug: detector bin indices (int32) and weights (float32), each (A, M, N),
 * exactly as the projector kernels compute them (_xray2d.py:331-349). */
int scxr_xr2d_weights(const scxr_xr2d_plan* plan, int32_t* inds, float* weights,
                      void* stream);


/* ------------------------------------------------------ vector kernels --
 * Elementwise algebra behind LinearOperator sums / scalings
 * (scico/linop/_linop.py:212-258) and deterministic float64 reductions
 * (sums of scico's jnp.sum / norm calls, e.g. scico/solver.py:380-397,
 * scico/loss.py:208-209, scico/functional/_norm.py:246-250).
 * Reductions write one float64 to the device pointer `out`; `ws` is a
 * device workspace of at least 1024 float64.
 */
enum {
  SCXR_RED_DOT = 0,    /* sum x*y                                  */
  SCXR_RED_SUMSQ = 1,  /* sum x^2                                  */
  SCXR_RED_SQDIST = 2, /* sum (x-y)^2                              */
  SCXR_RED_SUM = 3,    /* sum x                                    */
  SCXR_RED_L1 = 4,     /* sum |x|                                  */
  SCXR_RED_L21 = 5     /* x is (K, n): sum_j sqrt(sum_k x[k,j]^2)  */
};
enum { SCXR_POST_NONE = 0, SCXR_POST_SQRT = 1 };

int scxr_vec_axpby(int64_t n, float a, const float* x, float b, const float* y, float* z,
                   const int32_t* pred, void* stream);
/* z = x + y (sign > 0) or x - y (sign < 0). */
int scxr_vec_add(int64_t n, const float* x, float sign, const float* y, float* z,
                 const int32_t* pred, void* stream);
int scxr_vec_scale(int64_t n, float a, const float* x, float* z, const int32_t* pred,
                   void* stream);
int scxr_vec_fill(int64_t n, float v, float* z, const int32_t* pred, void* stream);
int scxr_vec_copy(int64_t n, const float* x, float* z, const int32_t* pred, void* stream);
/* *out = post(scale * reduction(x, y)). */
int scxr_vec_reduce(int32_t kind, int64_t n, int32_t K, const float* x, const float* y,
                    double scale, int32_t post, double* out, double* ws, const int32_t* pred,
                    void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SCICO_AMD_H */
