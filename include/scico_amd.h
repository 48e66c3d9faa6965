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

/* The tiled kernels hold a view group's detector rows in LDS and give each
 * lane a disjoint detector footprint; a geometry beyond that (a detector
 * too long for LDS, projected pixels much narrower or much wider than a
 * detector bin) gets general-geometry kernels with the same bins and
 * weights (see scxr_xr2d_info). */
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
/* scxr_xr2d_forward with the image's transpose imT (N x M) supplied by the
 * caller (e.g. by scxr_cg_iter_t) instead of computed by a transpose pass;
 * the same result bit for bit.  imT is read only when the plan has
 * column-fast views.  No counterpart in the reference (the CG loop's
 * fusion of the transpose into the direction update). */
int scxr_xr2d_forward_t(const scxr_xr2d_plan* plan, const float* im, const float* imT,
                        float* sino, const int32_t* pred, void* stream);

/* im = A^T sino  (_xray2d.py:267-306), gather form, exact adjoint of the
 * forward's weights. */
int scxr_xr2d_adjoint(const scxr_xr2d_plan* plan, const float* sino, float* im,
                      const int32_t* pred, void* stream);

/* Debug: detector bin indices (int32) and weights (float32), each (A, M, N),
 * exactly as the projector kernels compute them (_xray2d.py:331-349). */
int scxr_xr2d_weights(const scxr_xr2d_plan* plan, int32_t* inds, float* weights,
                      void* stream);

/* FBP row filter (scico/linop/xray/_xray2d.py:158-221): out (A, Dout) =
 * rows of y (A, D) linearly convolved with taps (device, D), keeping
 * outputs [shift, shift + Dout) of the full convolution. D <= 4096. */
int scxr_ramp_filter(int32_t A, int32_t D, int32_t Dout, int32_t shift, const float* taps,
                     const float* y, float* out, const int32_t* pred, void* stream);

/* Forward line split (a launch choice with no counterpart in the
 * reference): each forward group's image lines are divided over
 * `parts` workgroups (1..8) whose partial sinograms are added in part
 * order, which balances the CUs when there are few views.  parts = 0
 * restores the automatic choice made at plan creation (4 for images of at
 * least 256 x 256 pixels with fewer than two forward workgroups per CU,
 * else 1).  The split changes the
 * float32 summation order of a detector bin (results move by rounding,
 * relative L2 about 1e-7); with parts = 1 the plan's forward is bitwise
 * equal to the batched projector's for the same slice.  The effective
 * count (always 1 for the general-geometry forward) is written to
 * *effective when it is not null.  Calls on one plan on different streams
 * share its workspace and must not overlap.  Set the split before capturing
 * a graph that runs the plan's forward: a captured launch keeps the grid and
 * workspace of its capture, and a larger split replaces the workspace. */
int scxr_xr2d_set_forward_split(scxr_xr2d_plan* plan, int32_t parts, int32_t* effective);

/* Launch facts: out[0] forward workgroups (= partial count of
 * scxr_xr2d_forward_res), out[1] adjoint workgroups (= partial count of the
 * adjoint dot epilogue), out[2] views per forward workgroup, out[3] adjoint
 * detector window.  A geometry outside the tiled kernels' limits (see
 * scxr_xr2d_plan_create) runs general-geometry kernels instead; then out[0]
 * (forward) or out[1] (adjoint) is 0, and scxr_xr2d_forward_res and
 * scxr_xr2d_adjoint_epi with an epilogue or dot partials return
 * SCXR_ERR_UNSUPPORTED for that plan; scxr_xr2d_forward and
 * scxr_xr2d_adjoint work as for any other. */
int scxr_xr2d_info(const scxr_xr2d_plan* plan, int32_t* out /* host, 4 */);

/* Forward projection that also writes, per forward workgroup, the partial
 * sum of (y - A im)^2 over its sinogram rows (objective of
 * SquaredL2Loss, scico/loss.py:208-209). */
int scxr_xr2d_forward_res(const scxr_xr2d_plan* plan, const float* im, float* sino,
                          const float* y, double* res_ws, const int32_t* pred, void* stream);

/* Adjoint with a fused epilogue on the 2D finite-difference operator C
 * (scico/linop/_diff.py, append=0 or circular):
 *   kind 0: im = A^T sino
 *   kind 1: im = rho * C^T C p + A^T sino     (ADMM x-step operator,
 *           scico/optimize/_admmaux.py:222-229)
 *   kind 2: im = A^T sino + rho * C^T (z - u) (ADMM x-step right-hand side,
 *           scico/optimize/_admmaux.py:245-256; u may be NULL)
 * If dot_ws is non-NULL, each adjoint workgroup writes its partial of
 * sum(p * im) there (CG curvature p^T A p). */
int scxr_xr2d_adjoint_epi(const scxr_xr2d_plan* plan, const float* sino, float* im,
                          int32_t kind, float rho, int32_t circular, const float* p,
                          const float* z, const float* u, double* dot_ws, const int32_t* pred,
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
/* z = x * y elementwise. */
int scxr_vec_mul(int64_t n, const float* x, const float* y, float* z, const int32_t* pred,
                 void* stream);
/* z = (x >= thr) ? v : 0. */
int scxr_vec_ge(int64_t n, const float* x, float thr, float v, float* z, const int32_t* pred,
                void* stream);
/* z = x / *d with a device float64 divisor (0 divisor leaves x). */
int scxr_vec_div_dev(int64_t n, const float* x, const double* d, float* z, const int32_t* pred,
                     void* stream);
/* *out = post(scale * reduction(x, y)). */
int scxr_vec_reduce(int32_t kind, int64_t n, int32_t K, const float* x, const float* y,
                    double scale, int32_t post, double* out, double* ws, const int32_t* pred,
                    void* stream);
/* scxr_vec_reduce without the ws argument; allocates nothing.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The first-pass partials live in a static
 * device workspace of 1024 float64 (8 KiB per device, held until the
 * process ends; any n fits, since the first pass has at most 1024
 * workgroups).  Calls on one device share that workspace: calls that
 * overlap on one GPU (different streams, no ordering) race; calls ordered
 * on one stream do not.  Results are bitwise those of scxr_vec_reduce. */
int scxr_vec_reduce_noalloc(int32_t kind, int64_t n, int32_t K, const float* x,
                            const float* y, double scale, int32_t post, double* out,
                            const int32_t* pred, void* stream);

/* ---------------------------------------------- finite differences, prox --
 * FiniteDifference (scico/linop/_diff.py:25-272) over all axes of an
 * ndim-D array (ndim <= 3), append=0 (circular = 0) or circular; output
 * stacked (ndim, *shape) as VerticalStack collapses it
 * (scico/operator/_stack.py:126-128).
 */
int scxr_fd_forward(int32_t ndim, const int32_t* shape /* host */, const float* x, float* out,
                    int32_t circular, const int32_t* pred, void* stream);
/* out = scale * C^T g   (scico/linop/_stack.py:71-72) */
int scxr_fd_adjoint(int32_t ndim, const int32_t* shape, const float* g, float* out,
                    int32_t circular, float scale, const int32_t* pred, void* stream);
/* out = rho * C^T C x   (rho * FiniteDifference.gram_op) */
int scxr_fd_gram(int32_t ndim, const int32_t* shape, float rho, const float* x, float* out,
                 int32_t circular, const int32_t* pred, void* stream);
/* L21Norm.prox over axis 0 of v (K, n), K <= 8 (scico/functional/_norm.py:254-263) */
int scxr_l21_prox(int32_t K, int64_t n, const float* v, float lam, float* out,
                  const int32_t* pred, void* stream);
/* L1Norm.prox (scico/functional/_norm.py:110-116) */
int scxr_l1_prox(int64_t n, const float* v, float lam, float* out, const int32_t* pred,
                 void* stream);

/* ----------------------------------------------------------------- CG --
 * scico.solver.cg (scico/solver.py:326-405) with device-resident scalars.
 * state: float64[8 + 2 * (maxiter + 1)], flags: int32[maxiter + 1],
 * ws: float64[2 * scxr_elem_blocks(n)].  Usage per solve:
 *   Ax = A x0;  scxr_cg_start(b, Ax, r, p)
 *   for k in [0, maxiter): Ap = A p (predicated on &flags[k]), with the
 *     partials of p^T A p in pap_parts[pap_n];  scxr_cg_iter(k, ...)
 * state[2] holds the iteration count, state[3] sqrt(r^T r) / ||b||.
 */
int scxr_cg_start(int64_t n, const float* b, const float* Ax, float* r, float* p, double* state,
                  int32_t* flags, int32_t maxiter, double tol, double atol, double* ws,
                  void* stream);
int scxr_cg_iter(int64_t n, int32_t k, int32_t maxiter, float* x, float* r, float* p,
                 const float* Ap, double* state, int32_t* flags, const double* pap_parts,
                 int32_t pap_n, double* ws, void* stream);
/* scxr_cg_iter for an M x N image whose next direction p also feeds a 2D
 * projector: the update p = r + beta p additionally writes pT (N x M, the
 * transpose of p) for scxr_xr2d_forward_t, saving the projector's own
 * transpose pass.  Same results as scxr_cg_iter bit for bit (no
 * counterpart in the reference: a fusion of two device passes). */
int scxr_cg_iter_t(int32_t M, int32_t N, int32_t k, int32_t maxiter, float* x, float* r,
                   float* p, float* pT, const float* Ap, double* state, int32_t* flags,
                   const double* pap_parts, int32_t pap_n, double* ws, void* stream);
/* Workgroups (= partial count) of the elementwise passes over n elements. */
int32_t scxr_elem_blocks(int64_t n);

/* --------------------------------------------------------------- ADMM --
 * Fused z/u update of scico.optimize.ADMM.step (scico/optimize/_admm.py:
 * 364-378) for C = 2D FiniteDifference and g = lam * L21Norm (isotropic=1)
 * or lam * L1Norm (isotropic=0); thr = lam / rho.  z, u, zold are (2, M, N);
 * zold receives the previous z.  Partials (scxr_elem_blocks(M*N) each) of
 * ||C x - z||^2 and g(z) / lam go to ws_prml and ws_g.
 */
int scxr_admm_zu2d(int32_t M, int32_t N, const float* x, float* z, float* u, float* zold,
                   float thr, int32_t circular, int32_t isotropic, double* ws_prml,
                   double* ws_g, void* stream);
/* x-step right-hand side of ADMM-TV for an M x N image from a kept
 * aty = A^T y: out = aty + rho C^T (z - u) (_admmaux.py:245-256), the same
 * values as scxr_xr2d_adjoint_epi kind 2 on y without back projecting y. */
int scxr_admm_rhs2d(int32_t M, int32_t N, const float* aty, const float* z, const float* u,
                    float rho, int32_t circular, float* out, void* stream);
/* Partials of ||rho * C^T (z - zold)||^2 (dual residual, _admm.py:279-295). */
int scxr_admm_dual2d(int32_t M, int32_t N, const float* z, const float* zold, float rho,
                     int32_t circular, double* ws, void* stream);
/* nd variants (ndim <= 3, FiniteDifference over all axes) for operators
 * without projector epilogues (XRayTransform3D, ...):
 *   out = base + rho C^T (z - u)              (_admmaux.py:245-256)
 *   out = rho C^T C x + ata, partials of x.out (_admmaux.py:258-268)
 * out may alias base / ata. */
int scxr_admm_rhs_nd(int32_t ndim, const int32_t* shape /* host */, const float* base,
                     const float* z, const float* u, float rho, int32_t circular, float* out,
                     void* stream);
int scxr_admm_lhs_nd(int32_t ndim, const int32_t* shape /* host */, const float* x,
                     const float* ata, float rho, int32_t circular, float* out, double* dot_ws,
                     const int32_t* pred, void* stream);
int scxr_admm_zu_nd(int32_t ndim, const int32_t* shape /* host */, const float* x, float* z,
                    float* u, float* zold, float thr, int32_t circular, int32_t isotropic,
                    double* ws_prml, double* ws_g, void* stream);
int scxr_admm_dual_nd(int32_t ndim, const int32_t* shape /* host */, const float* z,
                      const float* zold, float rho, int32_t circular, double* ws, void* stream);
/* Appends (Objective, Prml Rsdl, Dual Rsdl, CG It, CG Res) to hist[5 * *counter]
 * (scico/optimize/_admm.py:171-198, 241-295) and increments *counter. */
int scxr_admm_itstat(const double* res, int32_t nres, const double* g, int32_t ng,
                     const double* prml, int32_t nprml, const double* dual, int32_t ndual,
                     const double* cg_state, double lam, double rho, double* hist,
                     int32_t* counter, int32_t hist_cap, void* stream);

/* ----------------------------------------------------------------- TV --
 * Approximate scaled prox of IsotropicTVNorm / AnisotropicTVNorm over all
 * axes (scico/functional/_tvnorm.py:117-236: Haar transform of the
 * edge-padded input, shrink of the high-pass, adjoint, crop, times 1/K),
 * fused into one stencil pass; tau = sqrt(2) * K * lam, K = 2 * ndim.
 * out must not alias v.
 */
int scxr_tv_prox(int32_t ndim, const int32_t* shape, const float* v, float* out, float tau,
                 int32_t isotropic, int32_t circular, const int32_t* pred, void* stream);
/* TV value norm(FD(x)) (scico/functional/_tvnorm.py:89-115) into the device
 * double *out; ws: float64[scxr_elem_blocks(n) + 1]. */
/* tv_prox with a per-pixel weight map: threshold tau * w[p] at coefficient
 * location p (w nullable = all ones).  Used by the differentiable unrolled
 * reconstruction; no reference counterpart beyond TVNorm.prox with w = 1. */
int scxr_tv_prox_w(int32_t ndim, const int32_t* shape /* host */, const float* v,
                   const float* w, float* out, float tau, int32_t isotropic,
                   int32_t circular, const int32_t* pred, void* stream);
/* Float count of the workspace of scxr_tv_prox_vjp (2 ndim prod(padded shape)). */
int64_t scxr_tv_prox_vjp_ws(int32_t ndim, const int32_t* shape, int32_t circular);
/* Vector-Jacobian product of scxr_tv_prox_w: given the cotangent gout of
 * out, gv = (d out / d v)^T gout and gw = (d out / d w)^T gout (gw nullable). */
int scxr_tv_prox_vjp(int32_t ndim, const int32_t* shape /* host */, const float* v,
                     const float* w, const float* gout, float tau, int32_t isotropic,
                     int32_t circular, float* gv, float* gw, float* ws, void* stream);
int scxr_tv_norm(int32_t ndim, const int32_t* shape, const float* x, int32_t isotropic,
                 int32_t circular, double* out, double* ws, const int32_t* pred, void* stream);

/* ------------------------------- fused AcceleratedPGM deconvolution --
 * One AcceleratedPGM-TV iteration (scico/optimize/_pgm.py:183-252) on a
 * 1024 x 1024 image with f = 0.5 ||A x - y||^2, A a real CircularConvolve,
 * g = lam TV (2D), constant step 1/L, in four launches that also record
 * its statistics row:
 *   scxr_tv_prox_apgm:  x <- prox_{tau TV}(u) (tau = lam / L), the
 *     partials of ||x - v||^2 and of TV(x), v <- x + c (x - x_prev)
 *     (c = (t_k - 1) / t_{k+1});
 *   scxr_cconv_apgm_grad_itstat: G x and the next gradient step u, with
 *     the objective's partials, and the Objective and Residual row (the
 *     three launches of scxr_cconv_apgm_grad, the last of which also
 *     reduces the row as scxr_apgm_itstat does).
 */
/* Workgroup (partial) count of scxr_tv_prox_apgm for an M x N image. */
int32_t scxr_tv_prox_apgm_parts(int32_t M, int32_t N);
/* x <- prox of u as scxr_tv_prox (2D); res_ws[b] = sum over workgroup b of
 * (x - v)^2 and tv_ws[b] its share of the TV norm of the new x (as
 * scxr_tv_norm; both float64); v <- x + c (x - x_prev) with x_prev the x
 * on entry.  u, x and v are distinct M x N buffers. */
int scxr_tv_prox_apgm(int32_t M, int32_t N, const float* u, float* x, float* v, float tau,
                      int32_t isotropic, int32_t circular, float c, double* res_ws,
                      double* tv_ws, const int32_t* pred, void* stream);
/* hist[2 row] = obj_scale * sum(obj) + lam * sum(tv), hist[2 row + 1] =
 * sqrt(sum(res)), row = *counter (skipped at or beyond hist_cap), then
 * *counter += 1. */
int scxr_apgm_itstat(const double* obj, int32_t nobj, double obj_scale, const double* tv,
                     int32_t ntv, double lam, const double* res, int32_t nres, double* hist,
                     int32_t* counter, int32_t hist_cap, void* stream);


/* ------------------------------------------------ CircularConvolve / FFT --
 * Replaces scico.linop.CircularConvolve (scico/linop/_circconv.py:79-209)
 * for real float32 input of shape (batch..., F0[, F1[, F2]]).  FFT lengths:
 * powers of two up to 4096 (radix-2 in LDS) or any length up to 512
 * (direct DFT).  The plan owns the spectrum (complex64, F-shaped) and a
 * complex workspace of batch * prod(F).  Two power-of-two axes convolve on
 * half spectra (F0 x (F1/2 + 1)) with the Hermitian part of the spectrum,
 * which equals real(ifftn(H * fftn(x))) for any H.
 */
typedef struct scxr_cconv_plan scxr_cconv_plan;
int scxr_cconv_plan_create(int32_t ndim, const int32_t* fshape /* host */, int64_t batch,
                           scxr_cconv_plan** out);
int scxr_cconv_plan_destroy(scxr_cconv_plan* plan);
/* Spectrum = fftn(h) of the real, zero-padded F-shaped h (h_is_dft = 0) or
 * a copy of the complex64 h (h_is_dft = 1); then times shift (complex64,
 * nullable) -- the h_center shift of _circconv.py:155-186. */
int scxr_cconv_set_spectrum(scxr_cconv_plan* plan, const float* h, int32_t h_is_dft,
                            const float* shift, void* stream);
/* Copy the complex64 spectrum to out (device). */
int scxr_cconv_spectrum(const scxr_cconv_plan* plan, float* out, void* stream);
/* y = real(ifftn(H * fftn(x))) or, adjoint != 0, with conj(H)
 * (_circconv.py:188-209). */
int scxr_cconv_apply(const scxr_cconv_plan* plan, const float* x, float* y, int32_t adjoint,
                     const int32_t* pred, void* stream);
/* Gradient step of the fused AcceleratedPGM iteration, for a plan of one
 * 1024 x 1024 real convolution (SCXR_ERR_UNSUPPORTED otherwise); G = A^T A
 * is the convolution by gs = |Hs|^2 (Hs the Hermitian part of the spectrum
 * on the half grid k0 < 1024, k1 < 513; gs float32 and Yh complex64 are
 * stored transposed, gs[k1 1024 + k0]):
 *   gx <- G x;  u <- v + b ((1 + c) gx - c gx_prev - aty)
 * with gx_prev the gx on entry (G v for v = x + c (x - x_prev), by
 * linearity; c = 0 at the first step, with gx zero), aty = A^T y and
 * b = -1/L.  If obj is non-NULL, obj[k] (513 doubles) = w_k sum over the
 * column k of |Hs X - Yh|^2 (X = fft2(x), Yh = fft2(y) on the half grid;
 * w_k = 1 for k = 0 and 512, else 2), so that sum(obj) / 1024^2
 * = ||A x - y||^2 (Parseval).  No reference counterpart as one call: it
 * is the gradient of SquaredL2Loss (scico/loss.py:208-209) and the
 * objective evaluation of _pgm.py:151-156, one convolution instead of
 * three. */
int scxr_cconv_apgm_grad(const scxr_cconv_plan* plan, const float* x, const float* gs,
                         const float* Yh, double* obj, float* gx, const float* v,
                         const float* aty, float c, float b, float* u, const int32_t* pred,
                         void* stream);
/* scxr_cconv_apgm_grad with obj and Yh required, followed in its last
 * launch by the statistics row of scxr_apgm_itstat(obj, 513, obj_scale,
 * tv, ntv, lam, res, nres, hist, counter, hist_cap): the row sums the
 * objective partials of this call and the prox partials (tv, res) that
 * earlier work on the stream wrote.  The sums' double roundings may differ
 * from scxr_apgm_itstat's (512 threads instead of 1024).  A call whose
 * predicate is 0 records no row.  No reference counterpart as one call
 * (see scxr_cconv_apgm_grad; the row is _pgm.py:242-248's itstat). */
int scxr_cconv_apgm_grad_itstat(const scxr_cconv_plan* plan, const float* x, const float* gs,
                                const float* Yh, double* obj, float* gx, const float* v,
                                const float* aty, float c, float b, float* u, double obj_scale,
                                const double* tv, int32_t ntv, double lam, const double* res,
                                int32_t nres, double* hist, int32_t* counter, int32_t hist_cap,
                                const int32_t* pred, void* stream);
/* Complex fftn / ifftn over the plan's F axes (x real if x_real != 0). */
int scxr_fftn(const scxr_cconv_plan* plan, const float* x, int32_t x_real, float* out,
              int32_t inverse, const int32_t* pred, void* stream);

/* ------------------------------------------- ragged batches of 2D slices --
 * Many independent XRayTransform2D problems (different M, N, A, D) in
 * single launches.  The reference batches same-shape images with
 * jax.vmap / pmap (scico/flax/examples/data_generation.py:153-190); ragged
 * batches have no reference counterpart.  Slice s's image occupies
 * elements [ioff[s], ioff[s] + M N) of packed image vectors, its sinogram
 * [soff[s], soff[s] + A D) of packed sinograms, and its FD stacks (2, M, N)
 * start at 2 ioff[s].  Per-slice results are bitwise those of the
 * single-slice entry points on the same slice.
 *   dims: host, ns x 4 (M, N, A, D); angles: host, concatenated (sum A);
 *   geo: host, ns x 5 (x0[0], x0[1], dx[0], dx[1], y0) as in
 *   scxr_xr2d_plan_create; vpg: views per forward group (1, 2, 4).
 * pred (nullable) holds one device predicate per slice at pred[s * pstride]
 * (a slice whose predicate is 0 is skipped). */
typedef struct scxr_xr2d_batch scxr_xr2d_batch;

int scxr_xr2d_batch_create(int32_t ns, const int32_t* dims, const float* angles,
                           const double* geo, int32_t vpg, scxr_xr2d_batch** out);
int scxr_xr2d_batch_destroy(scxr_xr2d_batch* plan);
/* out (host, 11): ns, image elements, sinogram elements, forward items,
 * adjoint items, transpose items, CG/ADMM blocks, max D, vpg, adjoint
 * tiles (dot partials), forward groups (residual partials). */
int scxr_xr2d_batch_info(const scxr_xr2d_batch* plan, int64_t* out);
int scxr_xr2d_batch_offsets(const scxr_xr2d_batch* plan, int64_t* ioff /* host, ns */,
                            int64_t* soff /* host, ns */);
/* Persistent grid sizes (0 = one workgroup per item) and item order
 * (0 slice-major, 1 round robin, 2 decreasing cost); set by the autotuner. */
int scxr_xr2d_batch_configure(scxr_xr2d_batch* plan, int32_t fwd_grid, int32_t adj_grid,
                              int32_t order);
/* View groups per adjoint tile: 3 (default; each slice's back projection
 * bitwise that of scxr_xr2d_adjoint) or 2 (fewer threads per tile, more
 * tiles resident on batches of many tiles; each pixel's sum is split over
 * the views differently, so results agree with the single-slice path to
 * float32 rounding only).  No reference counterpart: a tuning choice of
 * the batched launch, set by the autotuner. */
int scxr_xr2d_batch_set_adjoint_groups(scxr_xr2d_batch* plan, int32_t groups);
/* sino = A im for every slice; with y and res_ws also per-group partials of
 * ||y - A im||^2 (res_ws has "forward groups" doubles). */
int scxr_xr2d_batch_forward(const scxr_xr2d_batch* plan, const float* im, float* sino,
                            const float* y, double* res_ws, const int32_t* pred,
                            int32_t pstride, void* stream);
/* im = A^T sino with the epilogues of scxr_xr2d_adjoint_epi per slice
 * (dot_ws has "adjoint tiles" doubles). */
int scxr_xr2d_batch_adjoint(const scxr_xr2d_batch* plan, const float* sino, float* im,
                            int32_t kind, float rho, int32_t circular, const float* p,
                            const float* z, const float* u, double* dot_ws, const int32_t* pred,
                            int32_t pstride, void* stream);
/* Per-slice CG and ADMM passes (scxr_cg_start / scxr_cg_iter / scxr_admm_zu2d
 * / scxr_admm_dual2d / scxr_admm_itstat on every slice):
 *   state: ns x (8 + 2 (maxiter + 1)) doubles, flags: ns x (maxiter + 1),
 *   ws, ws_prml, ws_g, dual: 2 x (CG/ADMM blocks) doubles,
 *   res: forward-group partials, pap_parts: adjoint-tile partials,
 *   hist: ns x hist_cap x 5, counters: ns int32. */
int scxr_batch_cg_start(const scxr_xr2d_batch* plan, const float* b, const float* Ax, float* r,
                        float* p, double* state, int32_t* flags, int32_t maxiter, double tol,
                        double atol, double* ws, void* stream);
int scxr_batch_cg_iter(const scxr_xr2d_batch* plan, int32_t k, int32_t maxiter, float* x,
                       float* r, float* p, const float* Ap, double* state, int32_t* flags,
                       const double* pap_parts, double* ws, void* stream);
int scxr_batch_admm_zu(const scxr_xr2d_batch* plan, const float* x, float* z, float* u,
                       float* zold, float thr, int32_t circular, int32_t isotropic,
                       double* ws_prml, double* ws_g, void* stream);
/* out = aty + rho C^T (z - u) per slice from a kept aty = A^T y: the same
 * values as scxr_xr2d_batch_adjoint's right-hand-side epilogue (kind 2,
 * _admmaux.py:245-256) without back projecting y again. */
int scxr_batch_admm_rhs(const scxr_xr2d_batch* plan, const float* aty, const float* z,
                        const float* u, float rho, int32_t circular, float* out, void* stream);
int scxr_batch_admm_dual(const scxr_xr2d_batch* plan, const float* z, const float* zold,
                         float rho, int32_t circular, double* ws, void* stream);
int scxr_batch_admm_itstat(const scxr_xr2d_batch* plan, const double* res, const double* g,
                           const double* prml, const double* dual, const double* cg_state,
                           int32_t maxiter, double lam, double rho, double* hist,
                           int32_t* counters, int32_t hist_cap, void* stream);

/* ------------------------------------------------ 3D parallel-beam X-ray --
 * Replaces scico.linop.xray.XRayTransform3D:
 *   constructor            scico/linop/xray/_xray3d.py:54-93
 *   project / _project     scico/linop/xray/_xray3d.py:95-159
 *   back_project           scico/linop/xray/_xray3d.py:102-108, 161-204
 *   _calc_weights          scico/linop/xray/_xray3d.py:206-266
 * Volume (X, Y, Z) row-major, projections (V, D0, D1) row-major; matrices
 * (host, V x 2 x 4 float32) as the operator holds them (_xray3d.py:80).
 * Geometries where one detector axis follows one volume axis exactly
 * (single rotation axis, e.g. matrices_from_euler_angles with seq "Y")
 * run in slice mode: slice-parallel kernels, deterministic, no atomics.
 * Others run in general mode (gathers in a fixed order, also
 * deterministic).  allow_slice is a set of flags: bit 0 clear forces
 * general mode; bit 1 set keeps the slice adjoint that reads the detector
 * rows from global memory instead of staging them in LDS (the two give
 * bitwise-equal results; the staged one is faster).  Pass 1 for the
 * default.
 */
typedef struct scxr_xr3d_plan scxr_xr3d_plan;

int scxr_xr3d_plan_create(int32_t X, int32_t Y, int32_t Z, int32_t V,
                          const float* matrices /* host, V x 2 x 4 */, int32_t D0,
                          int32_t D1, int32_t allow_slice, scxr_xr3d_plan** out);
int scxr_xr3d_plan_destroy(scxr_xr3d_plan* plan);
/* out (host, 9): mode (0 general, 1 slice), slice volume axis, slice
 * detector axis, column offset, forward candidates per row, volume used in
 * place, projections used in place, in-plane detector size, adjoint rows
 * staged in LDS (1) or read from global memory (0). */
int scxr_xr3d_info(const scxr_xr3d_plan* plan, int32_t* out);
/* proj = H vol  (_xray3d.py:110-159). */
int scxr_xr3d_forward(const scxr_xr3d_plan* plan, const float* vol, float* proj,
                      const int32_t* pred, void* stream);
/* vol = H^T proj  (_xray3d.py:161-204). */
int scxr_xr3d_adjoint(const scxr_xr3d_plan* plan, const float* proj, float* vol,
                      const int32_t* pred, void* stream);
/* Debug: the reference's weight arrays of one view (_xray3d.py:206-266):
 * ul (2, X, Y, Z) int32 and w (4, X, Y, Z) float32, from the same device
 * functions the projection kernels use. */
int scxr_xr3d_weights(const scxr_xr3d_plan* plan, int32_t view, int32_t* ul, float* w,
                      void* stream);


#ifdef __cplusplus
}
#endif

#endif /* SCICO_AMD_H */
