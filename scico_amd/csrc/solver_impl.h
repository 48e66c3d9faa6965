// Internal constants of the solver kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace scxr {
namespace solver {

// CG state block layout (float64 slots).
enum { ST_BN = 0, ST_TOL2 = 1, ST_NITER = 2, ST_RELRES = 3, ST_NUM = 8 };

constexpr int MAX_K = 8;          // largest stack size of the L21 prox
constexpr int ELEM_BLOCKS = 1024; // grid cap (and partial count) of elementwise passes

unsigned elem_grid(int64_t n);

// Batched solves over ragged slices packed in one vector: slice s owns
// elements [off[s], off[s] + M N) (and 2 M N of the stacked FD arrays at
// 2 off[s]) and runs on G[s] = elem_grid(M N) virtual blocks, the
// single-slice grid, so each slice's reductions are bitwise those of a
// single-slice solve.  Per-block partials of slice s live at woff[s].
struct SegTabs {
  int ns = 0;
  int nblk = 0;                      // total virtual blocks
  const int* blk_seg = nullptr;      // virtual block -> slice
  const int* blk_first = nullptr;    // first virtual block of slice s
  const int* G = nullptr;            // virtual blocks of slice s
  const int64_t* off = nullptr;      // element offset of slice s
  const int2* MN = nullptr;          // image shape of slice s
  const int* woff = nullptr;         // partials offset of slice s (prefix of G)
  const int* pap_off = nullptr;      // projector adjoint tile partials of slice s
  const int* pap_n = nullptr;
  const int* res_off = nullptr;      // projector forward group partials of slice s
  const int* res_n = nullptr;
};

int batch_cg_start(const SegTabs& t, const float* b, const float* Ax, float* r, float* p,
                   double* state, int32_t* flags, int maxiter, double tol, double atol,
                   double* ws, hipStream_t s);
int batch_cg_iter(const SegTabs& t, int k, int maxiter, float* x, float* r, float* p,
                  const float* Ap, double* state, int32_t* flags, const double* pap_parts,
                  double* ws, hipStream_t s);
int batch_admm_zu(const SegTabs& t, const float* x, float* z, float* u, float* zold, float thr,
                  int circ, int iso, double* ws_prml, double* ws_g, hipStream_t s);
int batch_admm_rhs(const SegTabs& t, const float* aty, const float* z, const float* u, float rho,
                   int circ, float* out, hipStream_t s);
int batch_admm_dual(const SegTabs& t, const float* z, const float* zold, float rho, int circ,
                    double* ws, hipStream_t s);
int batch_admm_itstat(const SegTabs& t, const double* res, const double* g, const double* prml,
                      const double* dual, const double* cg_state, int maxiter, double lam,
                      double rho, double* hist, int32_t* counters, int hist_cap, hipStream_t s);

}  // namespace solver
}  // namespace scxr
