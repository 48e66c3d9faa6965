// Internal interface of the 2D projector (shared by the 2D, per-slice 3D,
// batched and solver translation units).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace scxr {
namespace xr2d {

// Adjoint tiling: ADJ_TI x (ADJ_TJ * PPT) pixels per workgroup (PPT per
// thread, PPT = 1 or 2),
// ADJ_VC views staged per LDS chunk, at most ADJ_WFAST bins per window.
constexpr int ADJ_TI = 16;
constexpr int ADJ_TJ = 16;
constexpr int ADJ_VC = 32;
constexpr int ADJ_VS = 3;  // groups of 256 threads splitting the views of a tile
constexpr int ADJ_WFAST = 64;  // window slots (one per lane)

// Forward: FWD_WAVES waves x ch chains (2, or 1 for wide detectors)
// private LDS copies of (D + 4) bins per owned view.
constexpr int FWD_WAVES = 8;
constexpr int FWD_SP_MAX = 8;  // most workgroups one forward group's lines are split over
constexpr size_t FWD_LDS_MAX = 64 * 1024;
constexpr int FWD_SMAX = 64;  // largest lane stride of the forward projector
inline size_t fwd_lds_bytes(int vpg, int D, int ch) {
  return sizeof(float) * static_cast<size_t>(vpg) * FWD_WAVES * ch *
         (static_cast<size_t>(D) + 4);
}
// Chains per wave for vpg views of D bins: 2 when the rows fit, else 1
// (0: too wide).
inline int fwd_chains(int vpg, int D) {
  if (fwd_lds_bytes(vpg, D, 2) <= FWD_LDS_MAX) return 2;
  if (fwd_lds_bytes(vpg, D, 1) <= FWD_LDS_MAX) return 1;
  return 0;
}
// Single-slice plans of one view per workgroup may use the device's whole
// LDS (160 KiB on gfx950) for detectors too wide for FWD_LDS_MAX.
inline int fwd_chains_ext(int vpg, int D, size_t lds_cap) {
  const int c = fwd_chains(vpg, D);
  if (c > 0) return c;
  return (vpg == 1 && fwd_lds_bytes(1, D, 1) <= lds_cap) ? 1 : 0;
}

struct Plan {
  int M = 0, N = 0, A = 0, D = 0;
  int vpg = 1;         // views per forward workgroup
  int fwd_ch = 2;      // forward chains per wave
  int n_groups = 0;    // forward workgroups
  int fwd_sp = 1;      // workgroups per forward group, splitting its lines
  int adj_w = 0;       // adjoint window width (bins)
  int adj_ppt = 1;     // adjoint pixels per thread
  bool adj_fract = false;  // Px >= 1 everywhere: fraction-instruction weights
  bool any_colfast = false;
  // general-geometry gathers (fwd_generic_kernel / adj_generic_kernel) in
  // place of the tiled kernels, which cannot hold this geometry
  bool fwd_generic = false, adj_generic = false;
  void* dev_block = nullptr;  // device tables (one allocation)
  float4* geom = nullptr;     // (Pxmin, P0, P1, width) per view
  double* rw = nullptr;       // RN64(1 / width) per view
  int4* groups = nullptr;     // forward groups: (first, count, class, 0)
  int* order = nullptr;       // views grouped by class
  int* gblock = nullptr;      // forward group of each forward block index
  float* clo = nullptr;       // adjoint window origin offset per view
  float* work = nullptr;      // transposed image workspace (M * N)
  float* parts = nullptr;     // fwd_sp_cap partial sinograms (A x D each)
  int fwd_sp_cap = 1;         // parts the workspace holds (1: none)
  std::vector<float4> host_geom;
};

// Adjoint epilogues (applied per pixel after the back projection sum).
enum { EPI_NONE = 0, EPI_GRAM = 1, EPI_RHS = 2 };
struct AdjEpi {
  int kind = EPI_NONE;
  int circular = 0;
  float rho = 0.0f;
  const float* p = nullptr;  // EPI_GRAM: image whose FD gram is added; dot partner
  const float* z = nullptr;  // EPI_RHS: (2, M, N)
  const float* u = nullptr;  // EPI_RHS: (2, M, N), may be null
  double* dot_ws = nullptr;  // per-workgroup partials of sum p * out (nullable)
};

void view_geometry(int A, const float* angles, const double* x0, const double* dx, double y0,
                   std::vector<float4>& geom, std::vector<double>& rw);

// Host tables of one slice (geometry, forward groups, adjoint windows).
struct HostTables {
  std::vector<float4> geom;
  std::vector<double> rw;
  std::vector<int4> groups;
  std::vector<int> order;
  std::vector<float> gcost;  // estimated relative time of each forward group
  std::vector<float> clo;
  int adj_w = 0;
  int adj_ppt = 1;
  bool any_colfast = false;
  bool px_ge1 = false;  // Px >= 1 for every pixel and view
  bool fwd_generic = false, adj_generic = false;  // see Plan
};
// vpg: views per forward group; ppt: adjoint pixels per thread (0 = 2 when
// the adjoint window fits, else 1).  fwd_ch: forward chains per wave (0:
// the rows do not fit).  With allow_generic, geometries the tiled kernels
// cannot hold set fwd_generic / adj_generic instead of failing.
int build_tables(HostTables& h, int M, int N, int A, int D, const double* x0, const double* dx,
                 double y0, const float* angles, int vpg, int ppt, int fwd_ch = -1,
                 bool allow_generic = false);

// Batched (ragged) slices: device tables of a batch plan.  Slice s has
// shape (M, N, A, D) = shape[s]; its image starts at ioff[s] in packed
// image vectors, its sinogram at soff[s], its views at voff[s] in the
// packed view tables, its forward groups at goff[s] and its adjoint tiles
// at toff[s].  Work items name (slice, group / tile).
struct BatchTabs {
  int ns = 0;
  const int4* shape = nullptr;    // (M, N, A, D)
  const int4* meta = nullptr;     // (adjoint window W, forward groups, tiles x, tiles y)
  const int64_t* ioff = nullptr;
  const int64_t* soff = nullptr;
  const int* voff = nullptr;
  const int* goff = nullptr;
  const int* toff = nullptr;
  const float4* geom = nullptr;
  const double* rw = nullptr;
  const int4* groups = nullptr;
  const int* order = nullptr;
  const float* clo = nullptr;
  const int2* fitems = nullptr;   // forward items (slice, group)
  const int2* aitems = nullptr;   // adjoint items (slice, tile)
  const int4* titems = nullptr;   // transpose items (slice, tile row, tile col, 0)
  int nf = 0, na = 0, nt = 0;
  int ppt = 1;                    // adjoint pixels per thread (all slices)
  int fract = 0;                  // 1: Px >= 1 for every pixel of every slice
  int ch = 2;                     // forward chains per wave (all slices)
  int avs = ADJ_VS;               // adjoint view groups per tile (2 or ADJ_VS)
};

// Epilogue arguments of the batched adjoint (packed vectors; z, u are
// (2, M, N) per slice at offset 2 ioff[s]; dot partials per tile at toff).
struct BatchEpi {
  int kind = EPI_NONE;
  int circular = 0;
  float rho = 0.0f;
  const float* p = nullptr;
  const float* z = nullptr;
  const float* u = nullptr;
  double* dot_ws = nullptr;
};

// Launchers (grid = persistent workgroup count, 0 = one per item).  pred
// (nullable) holds one predicate per slice at pred[s * pstride].
int batch_forward(const BatchTabs& t, int V, int Dmax, int grid, const float* im, float* imT,
                  float* sino, const float* yres, double* res_ws, const int32_t* pred,
                  int pstride, hipStream_t s);
int batch_adjoint(const BatchTabs& t, int grid, const float* sino, float* im,
                  const BatchEpi& epi, const int32_t* pred, int pstride, hipStream_t s);
int create(Plan& p, int M, int N, int A, int D, const double* x0, const double* dx, double y0,
           const float* angles);
void destroy(Plan& p);
int auto_split(const Plan& p);
int set_split(Plan& p, int parts);
// imT: the image transposed (N x M) by the caller, or null to transpose it
// into the plan's workspace first.
int forward(const Plan& p, const float* im, float* sino, const int32_t* pred, hipStream_t s,
            const float* yres, double* res_ws, const float* imT = nullptr);
int adjoint(const Plan& p, const float* sino, float* im, const int32_t* pred, hipStream_t s,
            const AdjEpi& epi);

}  // namespace xr2d
}  // namespace scxr
