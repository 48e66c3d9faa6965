// Internal constants of the solver kernels.
#pragma once

namespace scxr {
namespace solver {

// CG state block layout (float64 slots).
enum { ST_BN = 0, ST_TOL2 = 1, ST_NITER = 2, ST_RELRES = 3, ST_NUM = 8 };

constexpr int MAX_K = 8;          // largest stack size of the L21 prox
constexpr int ELEM_BLOCKS = 1024; // grid cap (and partial count) of elementwise passes

unsigned elem_grid(int64_t n);

}  // namespace solver
}  // namespace scxr
