// Ragged batches of 2D slices (many independent XRayTransform2D problems of
// different image size, view count and detector count) in single
// launches: the projector runs as a persistent grid over (slice, group)
// and (slice, tile) work items, and the CG / ADMM passes run every slice
// on its single-slice grid in one launch.  Every slice's arithmetic is
// that of the single-slice path, so the per-slice results are bitwise
// the single-slice results.
//
// The reference batches slices with jax.vmap / jax.pmap over images of one
// shape (scico/flax/examples/data_generation.py:153-190); ragged batches
// have no reference counterpart.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "common.h"
#include "solver_impl.h"
#include "xray2d_impl.h"

using namespace scxr;

namespace scxr {
namespace batch {

struct Batch {
  int ns = 0, vpg = 1, Dmax = 0;
  int fwd_grid = 0, adj_grid = 0, order = 0;
  int64_t n_img = 0, n_sino = 0;
  int n_tiles = 0, n_groups = 0;
  bool any_colfast = false;
  xr2d::BatchTabs bt;
  solver::SegTabs st;
  void* block = nullptr;   // device tables
  float* imT = nullptr;    // transposed images (packed like the images)
  // host copies
  std::vector<int4> shape, meta;
  std::vector<int64_t> ioff, soff;
  std::vector<int> voff, goff, toff, ngroups, ntiles;
  std::vector<int4> groups;
  std::vector<float> gcost;  // estimated relative time of each forward group
  std::vector<int2> fitems_nat, aitems_nat;
  std::vector<int4> titems;
  int2* fitems = nullptr;  // device item arrays (reordered by configure)
  int2* aitems = nullptr;
};

// Item orders: 0 slice-major, 1 round robin over slices, 2 decreasing cost
// (forward group: the plan's estimate, phases x lines x modelled LDS bank
// cost of its views; adjoint tile ~ views).
static std::vector<int2> order_items(const Batch& b, const std::vector<int2>& nat, bool fwd,
                                     int mode) {
  if (mode == 0) return nat;
  std::vector<int2> out;
  out.reserve(nat.size());
  if (mode == 1) {
    std::vector<std::vector<int2>> per(b.ns);
    for (const int2& w : nat) per[w.x].push_back(w);
    std::vector<size_t> pos(b.ns, 0);
    bool more = true;
    while (more) {
      more = false;
      for (int s = 0; s < b.ns; ++s)
        if (pos[s] < per[s].size()) {
          out.push_back(per[s][pos[s]++]);
          more = true;
        }
    }
    return out;
  }
  std::vector<std::pair<double, int2>> c;
  c.reserve(nat.size());
  for (const int2& w : nat) {
    const int4 sh = b.shape[w.x];
    double cost;
    if (fwd) {
      cost = b.gcost[b.goff[w.x] + w.y];
    } else {
      cost = static_cast<double>(sh.z);
    }
    c.push_back({cost, w});
  }
  std::stable_sort(c.begin(), c.end(), [](const auto& a, const auto& b2) { return a.first > b2.first; });
  for (const auto& e : c) out.push_back(e.second);
  return out;
}

void destroy(Batch& b) {
  if (b.block) (void)hipFree(b.block);
  if (b.imT) (void)hipFree(b.imT);
  b.block = nullptr;
  b.imT = nullptr;
}

int create(Batch& b, int ns, const int32_t* dims, const float* angles, const double* geo,
           int vpg) {
  b.ns = ns;
  b.vpg = vpg;
  std::vector<float4> geom;
  std::vector<double> rw;
  std::vector<int> order;
  std::vector<float> clo;
  std::vector<int2> MN;
  std::vector<int> G, woff, blk_seg, blk_first;
  int64_t io = 0, so = 0;
  int vo = 0, go = 0, to = 0, wo = 0;
  // Tables of every slice; one adjoint tile shape for the whole batch (two
  // pixels per thread unless a slice's window needs one).
  std::vector<xr2d::HostTables> tabs(ns);
  std::vector<const float*> angp(ns);
  {
    const float* ang = angles;
    int ppt = 2;
    for (int s = 0; s < ns; ++s) {
      const int M = dims[4 * s], N = dims[4 * s + 1], A = dims[4 * s + 2], D = dims[4 * s + 3];
      SCXR_REQUIRE(M > 0 && N > 0 && A > 0 && D > 0, SCXR_ERR_SHAPE,
                   "slice %d: non-positive size (M=%d N=%d A=%d D=%d)", s, M, N, A, D);
      const double* g5 = geo + 5 * s;
      angp[s] = ang;
      const int rc = xr2d::build_tables(tabs[s], M, N, A, D, g5, g5 + 2, g5[4], ang, vpg, 0);
      if (rc) {
        const std::string msg = scxr_last_error();
        return fail(rc, "slice %d: %s", s, msg.c_str());
      }
      ppt = std::min(ppt, tabs[s].adj_ppt);
      ang += A;
    }
    for (int s = 0; s < ns; ++s) {
      if (tabs[s].adj_ppt == ppt) continue;
      const double* g5 = geo + 5 * s;
      const int rc = xr2d::build_tables(tabs[s], dims[4 * s], dims[4 * s + 1], dims[4 * s + 2],
                                        dims[4 * s + 3], g5, g5 + 2, g5[4], angp[s], vpg, ppt);
      if (rc) return rc;
    }
    b.bt.ppt = ppt;
    b.bt.fract = 1;
    for (int s = 0; s < ns; ++s) b.bt.fract = b.bt.fract && tabs[s].px_ge1;
  }
  {
    int Dm = 0;
    for (int s = 0; s < ns; ++s) Dm = std::max(Dm, dims[4 * s + 3]);
    b.bt.ch = xr2d::fwd_chains(vpg, Dm);
    SCXR_REQUIRE(b.bt.ch > 0, SCXR_ERR_UNSUPPORTED,
                 "det_count %d too large for the forward projector with vpg %d", Dm, vpg);
  }
  for (int s = 0; s < ns; ++s) {
    const int M = dims[4 * s], N = dims[4 * s + 1], A = dims[4 * s + 2], D = dims[4 * s + 3];
    const xr2d::HostTables& h = tabs[s];
    b.any_colfast = b.any_colfast || h.any_colfast;
    b.Dmax = std::max(b.Dmax, D);
    const int tx = static_cast<int>(ceil_div(N, xr2d::ADJ_TJ * b.bt.ppt));
    const int ty = static_cast<int>(ceil_div(M, xr2d::ADJ_TI));
    b.shape.push_back(make_int4(M, N, A, D));
    b.meta.push_back(make_int4(h.adj_w, static_cast<int>(h.groups.size()), tx, ty));
    b.ioff.push_back(io);
    b.soff.push_back(so);
    b.voff.push_back(vo);
    b.goff.push_back(go);
    b.toff.push_back(to);
    b.ngroups.push_back(static_cast<int>(h.groups.size()));
    b.ntiles.push_back(tx * ty);
    for (size_t g = 0; g < h.groups.size(); ++g) b.fitems_nat.push_back(make_int2(s, static_cast<int>(g)));
    for (int t = 0; t < tx * ty; ++t) b.aitems_nat.push_back(make_int2(s, t));
    if (h.any_colfast)
      for (int i = 0; i < static_cast<int>(ceil_div(M, 32)); ++i)
        for (int j = 0; j < static_cast<int>(ceil_div(N, 32)); ++j)
          b.titems.push_back(make_int4(s, i, j, 0));
    geom.insert(geom.end(), h.geom.begin(), h.geom.end());
    rw.insert(rw.end(), h.rw.begin(), h.rw.end());
    order.insert(order.end(), h.order.begin(), h.order.end());
    clo.insert(clo.end(), h.clo.begin(), h.clo.end());
    b.groups.insert(b.groups.end(), h.groups.begin(), h.groups.end());
    b.gcost.insert(b.gcost.end(), h.gcost.begin(), h.gcost.end());
    const int Gs = static_cast<int>(solver::elem_grid(static_cast<int64_t>(M) * N));
    MN.push_back(make_int2(M, N));
    G.push_back(Gs);
    woff.push_back(wo);
    blk_first.push_back(static_cast<int>(blk_seg.size()));
    for (int k = 0; k < Gs; ++k) blk_seg.push_back(s);
    io += static_cast<int64_t>(M) * N;
    so += static_cast<int64_t>(A) * D;
    vo += A;
    go += static_cast<int>(h.groups.size());
    to += tx * ty;
    wo += Gs;
  }
  b.n_img = io;
  b.n_sino = so;
  b.n_tiles = to;
  b.n_groups = go;

  // one device block for every table
  size_t total = 0;
  auto need = [&](size_t bytes) { total += (bytes + 255) & ~static_cast<size_t>(255); };
  const size_t nf = b.fitems_nat.size(), na = b.aitems_nat.size(), nt = b.titems.size();
  need(sizeof(int4) * ns);       // shape
  need(sizeof(int4) * ns);       // meta
  need(sizeof(int64_t) * ns);    // ioff
  need(sizeof(int64_t) * ns);    // soff
  need(sizeof(int) * ns * 3);    // voff goff toff
  need(sizeof(float4) * geom.size());
  need(sizeof(double) * rw.size());
  need(sizeof(int4) * b.groups.size());
  need(sizeof(int) * order.size());
  need(sizeof(float) * clo.size());
  need(sizeof(int2) * nf);
  need(sizeof(int2) * na);
  need(sizeof(int4) * std::max<size_t>(nt, 1));
  need(sizeof(int) * blk_seg.size());
  need(sizeof(int) * ns * 5);    // blk_first G woff pap_n res_n
  need(sizeof(int2) * ns);       // MN
  if (hipMalloc(&b.block, total) != hipSuccess) {
    (void)hipGetLastError();
    return fail(SCXR_ERR_NOMEM, "hipMalloc of %zu bytes for the batch tables failed", total);
  }
  char* base = static_cast<char*>(b.block);
  bool copied = true;
  auto carve = [&](const void* src, size_t bytes) -> void* {
    char* r = base;
    base += (bytes + 255) & ~static_cast<size_t>(255);
    if (bytes && src && hipMemcpy(r, src, bytes, hipMemcpyHostToDevice) != hipSuccess)
      copied = false;
    return r;
  };
#define CARVE(T, dst, vec) \
  dst = static_cast<T>(carve((vec).data(), sizeof(*(vec).data()) * (vec).size()))
  CARVE(const int4*, b.bt.shape, b.shape);
  CARVE(const int4*, b.bt.meta, b.meta);
  CARVE(const int64_t*, b.bt.ioff, b.ioff);
  CARVE(const int64_t*, b.bt.soff, b.soff);
  {
    std::vector<int> vgt(b.voff);
    vgt.insert(vgt.end(), b.goff.begin(), b.goff.end());
    vgt.insert(vgt.end(), b.toff.begin(), b.toff.end());
    const int* p = static_cast<const int*>(carve(vgt.data(), sizeof(int) * vgt.size()));
    b.bt.voff = p;
    b.bt.goff = p + ns;
    b.bt.toff = p + 2 * ns;
  }
  CARVE(const float4*, b.bt.geom, geom);
  CARVE(const double*, b.bt.rw, rw);
  CARVE(const int4*, b.bt.groups, b.groups);
  CARVE(const int*, b.bt.order, order);
  CARVE(const float*, b.bt.clo, clo);
  b.fitems = static_cast<int2*>(carve(b.fitems_nat.data(), sizeof(int2) * nf));
  b.aitems = static_cast<int2*>(carve(b.aitems_nat.data(), sizeof(int2) * na));
  b.bt.titems = static_cast<const int4*>(
      carve(nt ? b.titems.data() : nullptr, sizeof(int4) * std::max<size_t>(nt, 1)));
  b.bt.fitems = b.fitems;
  b.bt.aitems = b.aitems;
  b.bt.ns = ns;
  b.bt.nf = static_cast<int>(nf);
  b.bt.na = static_cast<int>(na);
  b.bt.nt = static_cast<int>(nt);
  CARVE(const int*, b.st.blk_seg, blk_seg);
  {
    std::vector<int> five(blk_first);
    five.insert(five.end(), G.begin(), G.end());
    five.insert(five.end(), woff.begin(), woff.end());
    five.insert(five.end(), b.ntiles.begin(), b.ntiles.end());
    five.insert(five.end(), b.ngroups.begin(), b.ngroups.end());
    const int* p = static_cast<const int*>(carve(five.data(), sizeof(int) * five.size()));
    b.st.blk_first = p;
    b.st.G = p + ns;
    b.st.woff = p + 2 * ns;
    b.st.pap_n = p + 3 * ns;
    b.st.res_n = p + 4 * ns;
  }
  CARVE(const int2*, b.st.MN, MN);
#undef CARVE
  b.st.ns = ns;
  b.st.nblk = static_cast<int>(blk_seg.size());
  b.st.off = b.bt.ioff;
  b.st.pap_off = b.bt.toff;
  b.st.res_off = b.bt.goff;
  SCXR_REQUIRE(copied, SCXR_ERR_HIP, "copying the batch tables to the device failed");
  if (b.any_colfast) {
    if (hipMalloc(&b.imT, sizeof(float) * static_cast<size_t>(b.n_img)) != hipSuccess) {
      (void)hipGetLastError();
      return fail(SCXR_ERR_NOMEM, "hipMalloc of the batch transpose workspace failed");
    }
  }
  return SCXR_OK;
}

int configure(Batch& b, int fwd_grid, int adj_grid, int order) {
  SCXR_REQUIRE(order >= 0 && order <= 2, SCXR_ERR_ARG, "item order must be 0, 1 or 2");
  b.fwd_grid = std::max(0, fwd_grid);
  b.adj_grid = std::max(0, adj_grid);
  if (order != b.order) {
    const std::vector<int2> f = order_items(b, b.fitems_nat, true, order);
    const std::vector<int2> a = order_items(b, b.aitems_nat, false, order);
    SCXR_HIP(hipMemcpy(b.fitems, f.data(), sizeof(int2) * f.size(), hipMemcpyHostToDevice));
    SCXR_HIP(hipMemcpy(b.aitems, a.data(), sizeof(int2) * a.size(), hipMemcpyHostToDevice));
    b.order = order;
  }
  return SCXR_OK;
}

}  // namespace batch
}  // namespace scxr

struct scxr_xr2d_batch {
  scxr::batch::Batch b;
};

extern "C" {

int scxr_xr2d_batch_create(int32_t ns, const int32_t* dims, const float* angles,
                           const double* geo, int32_t vpg, scxr_xr2d_batch** out) {
  SCXR_REQUIRE(out && dims && angles && geo, SCXR_ERR_ARG, "null argument to batch_create");
  SCXR_REQUIRE(ns > 0, SCXR_ERR_ARG, "empty batch");
  SCXR_REQUIRE(vpg == 1 || vpg == 2 || vpg == 4, SCXR_ERR_ARG, "vpg must be 1, 2 or 4");
  auto* p = new scxr_xr2d_batch();
  const int rc = batch::create(p->b, ns, dims, angles, geo, vpg);
  if (rc != SCXR_OK) {
    batch::destroy(p->b);
    delete p;
    return rc;
  }
  *out = p;
  return SCXR_OK;
}

int scxr_xr2d_batch_destroy(scxr_xr2d_batch* plan) {
  if (!plan) return SCXR_OK;
  batch::destroy(plan->b);
  delete plan;
  return SCXR_OK;
}

int scxr_xr2d_batch_info(const scxr_xr2d_batch* plan, int64_t* out) {
  SCXR_REQUIRE(plan && out, SCXR_ERR_ARG, "null argument to batch_info");
  const auto& b = plan->b;
  out[0] = b.ns;
  out[1] = b.n_img;
  out[2] = b.n_sino;
  out[3] = b.bt.nf;
  out[4] = b.bt.na;
  out[5] = b.bt.nt;
  out[6] = b.st.nblk;
  out[7] = b.Dmax;
  out[8] = b.vpg;
  out[9] = b.n_tiles;
  out[10] = b.n_groups;
  return SCXR_OK;
}

int scxr_xr2d_batch_offsets(const scxr_xr2d_batch* plan, int64_t* ioff, int64_t* soff) {
  SCXR_REQUIRE(plan && ioff && soff, SCXR_ERR_ARG, "null argument to batch_offsets");
  const auto& b = plan->b;
  std::copy(b.ioff.begin(), b.ioff.end(), ioff);
  std::copy(b.soff.begin(), b.soff.end(), soff);
  return SCXR_OK;
}

int scxr_xr2d_batch_configure(scxr_xr2d_batch* plan, int32_t fwd_grid, int32_t adj_grid,
                              int32_t order) {
  SCXR_REQUIRE(plan, SCXR_ERR_ARG, "null argument to batch_configure");
  return batch::configure(plan->b, fwd_grid, adj_grid, order);
}

int scxr_xr2d_batch_set_adjoint_groups(scxr_xr2d_batch* plan, int32_t groups) {
  SCXR_REQUIRE(plan, SCXR_ERR_ARG, "null argument to batch_set_adjoint_groups");
  SCXR_REQUIRE(groups == 2 || groups == xr2d::ADJ_VS, SCXR_ERR_ARG,
               "adjoint view groups must be 2 or %d", xr2d::ADJ_VS);
  plan->b.bt.avs = groups;
  return SCXR_OK;
}

int scxr_xr2d_batch_forward(const scxr_xr2d_batch* plan, const float* im, float* sino,
                            const float* y, double* res_ws, const int32_t* pred,
                            int32_t pstride, void* stream) {
  SCXR_REQUIRE(plan && im && sino, SCXR_ERR_ARG, "null argument to batch_forward");
  SCXR_REQUIRE(!y == !res_ws, SCXR_ERR_ARG, "batch_forward: y and res_ws go together");
  const auto& b = plan->b;
  return xr2d::batch_forward(b.bt, b.vpg, b.Dmax, b.fwd_grid, im, b.imT, sino, y, res_ws, pred,
                             pstride, as_stream(stream));
}

int scxr_xr2d_batch_adjoint(const scxr_xr2d_batch* plan, const float* sino, float* im,
                            int32_t kind, float rho, int32_t circular, const float* p,
                            const float* z, const float* u, double* dot_ws, const int32_t* pred,
                            int32_t pstride, void* stream) {
  SCXR_REQUIRE(plan && im && sino, SCXR_ERR_ARG, "null argument to batch_adjoint");
  SCXR_REQUIRE(kind >= xr2d::EPI_NONE && kind <= xr2d::EPI_RHS, SCXR_ERR_ARG,
               "bad epilogue kind %d", kind);
  SCXR_REQUIRE(kind != xr2d::EPI_GRAM || p, SCXR_ERR_ARG, "gram epilogue needs p");
  SCXR_REQUIRE(kind != xr2d::EPI_RHS || z, SCXR_ERR_ARG, "rhs epilogue needs z");
  SCXR_REQUIRE(!dot_ws || p, SCXR_ERR_ARG, "dot partials need p");
  xr2d::BatchEpi e;
  e.kind = kind;
  e.circular = circular;
  e.rho = rho;
  e.p = p;
  e.z = z;
  e.u = u;
  e.dot_ws = dot_ws;
  const auto& b = plan->b;
  return xr2d::batch_adjoint(b.bt, b.adj_grid, sino, im, e, pred, pstride, as_stream(stream));
}

int scxr_batch_cg_start(const scxr_xr2d_batch* plan, const float* b, const float* Ax, float* r,
                        float* p, double* state, int32_t* flags, int32_t maxiter, double tol,
                        double atol, double* ws, void* stream) {
  SCXR_REQUIRE(plan && b && Ax && r && p && state && flags && ws, SCXR_ERR_ARG,
               "bad batch_cg_start arguments");
  SCXR_REQUIRE(maxiter >= 0, SCXR_ERR_ARG, "negative maxiter");
  return solver::batch_cg_start(plan->b.st, b, Ax, r, p, state, flags, maxiter, tol, atol, ws,
                                as_stream(stream));
}

int scxr_batch_cg_iter(const scxr_xr2d_batch* plan, int32_t k, int32_t maxiter, float* x,
                       float* r, float* p, const float* Ap, double* state, int32_t* flags,
                       const double* pap_parts, double* ws, void* stream) {
  SCXR_REQUIRE(plan && x && r && p && Ap && state && flags && pap_parts && ws, SCXR_ERR_ARG,
               "bad batch_cg_iter arguments");
  SCXR_REQUIRE(k >= 0 && k < maxiter, SCXR_ERR_ARG, "bad batch_cg_iter index");
  return solver::batch_cg_iter(plan->b.st, k, maxiter, x, r, p, Ap, state, flags, pap_parts, ws,
                               as_stream(stream));
}

int scxr_batch_admm_zu(const scxr_xr2d_batch* plan, const float* x, float* z, float* u,
                       float* zold, float thr, int32_t circular, int32_t isotropic,
                       double* ws_prml, double* ws_g, void* stream) {
  SCXR_REQUIRE(plan && x && z && u && zold && ws_prml && ws_g, SCXR_ERR_ARG,
               "bad batch_admm_zu arguments");
  return solver::batch_admm_zu(plan->b.st, x, z, u, zold, thr, circular, isotropic, ws_prml,
                               ws_g, as_stream(stream));
}

int scxr_batch_admm_rhs(const scxr_xr2d_batch* plan, const float* aty, const float* z,
                        const float* u, float rho, int32_t circular, float* out, void* stream) {
  SCXR_REQUIRE(plan && aty && z && u && out, SCXR_ERR_ARG, "bad batch_admm_rhs arguments");
  return solver::batch_admm_rhs(plan->b.st, aty, z, u, rho, circular, out, as_stream(stream));
}

int scxr_batch_admm_dual(const scxr_xr2d_batch* plan, const float* z, const float* zold,
                         float rho, int32_t circular, double* ws, void* stream) {
  SCXR_REQUIRE(plan && z && zold && ws, SCXR_ERR_ARG, "bad batch_admm_dual arguments");
  return solver::batch_admm_dual(plan->b.st, z, zold, rho, circular, ws, as_stream(stream));
}

int scxr_batch_admm_itstat(const scxr_xr2d_batch* plan, const double* res, const double* g,
                           const double* prml, const double* dual, const double* cg_state,
                           int32_t maxiter, double lam, double rho, double* hist,
                           int32_t* counters, int32_t hist_cap, void* stream) {
  SCXR_REQUIRE(plan && res && g && prml && dual && hist && counters, SCXR_ERR_ARG,
               "bad batch_admm_itstat arguments");
  return solver::batch_admm_itstat(plan->b.st, res, g, prml, dual, cg_state, maxiter, lam, rho,
                                   hist, counters, hist_cap, as_stream(stream));
}

}  // extern "C"
