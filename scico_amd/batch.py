"""Ragged batches of 2D X-ray problems on single launches.

:class:`XRayTransform2DBatch` holds many :class:`XRayTransform2D`
geometries of different image size, view count and detector count.  Images
(and sinograms) of all slices are packed into one device vector; the
projector runs as one persistent launch over (slice, view group) and
(slice, pixel tile) work items.  :meth:`XRayTransform2DBatch.autotune`
times the launch parameters (views per forward group, persistent grid
size, item order) on the batch itself and keeps the fastest.

:class:`ADMMTVBatch` runs ADMM-TV (the fused iteration of
``scico_amd.optimize.ADMM`` for ``SquaredL2Loss + lam * L21(FD)``) on every
slice at once; one iteration is a fixed sequence of launches and can be
captured in a HIP graph.  Each slice's iterates and statistics are bitwise
those of the single-slice fused solver (the batch uses the single-slice
arithmetic and reduction order per slice).

The reference batches same-shape images with ``jax.vmap`` / ``jax.pmap``
(``scico/flax/examples/data_generation.py:153-190``).
"""

from __future__ import annotations

import ctypes
import time
import weakref
from typing import Optional, Sequence

import numpy as np
import torch

from . import _device as dev
from . import _vec as vec
from ._lib import call, load
from .solver import ST_NUM


def _defaults(shape, angles, x0=None, dx=None, y0=None, det_count=None):
    """XRayTransform2D constructor defaults (``_xray2d.py:86-121``)."""
    nx = tuple(int(s) for s in shape)
    if dx is None:
        dx = 2 * (np.sqrt(2) / 2,)
    if np.isscalar(dx):
        dx = 2 * (float(dx),)
    dx = tuple(float(d) for d in dx)
    if x0 is None:
        x0 = -(np.array(nx) * np.array(dx)) / 2
    if det_count is None:
        det_count = int(np.ceil(np.linalg.norm(nx)))
    if y0 is None:
        y0 = -det_count / 2
    return nx, np.asarray(angles, dtype=np.float32).reshape(-1), np.asarray(x0, np.float64), dx, \
        float(y0), int(det_count)


class XRayTransform2DBatch:
    """Batch of independent 2D parallel-beam transforms (one per slice)."""

    def __init__(self, shapes: Sequence, angles: Sequence, det_counts=None, x0s=None, dxs=None,
                 y0s=None, vpg: int = 1):
        ns = len(shapes)
        if ns == 0 or len(angles) != ns:
            raise ValueError("shapes and angles must be non-empty and of equal length")
        self.slices = []
        for s in range(ns):
            self.slices.append(_defaults(
                shapes[s], angles[s], None if x0s is None else x0s[s],
                None if dxs is None else dxs[s], None if y0s is None else y0s[s],
                None if det_counts is None else det_counts[s]))
        self.ns = ns
        self.input_shapes = [sl[0] for sl in self.slices]
        self.output_shapes = [(len(sl[1]), sl[5]) for sl in self.slices]
        self.vpg = int(vpg)
        self._plan = None
        self._retired = []
        self.generation = 0
        self.config = (0, 0, 0, 3)
        self._create()

    def _create(self):
        dev.require_gpu()
        lib = load()
        dims = np.array([[sl[0][0], sl[0][1], len(sl[1]), sl[5]] for sl in self.slices],
                        dtype=np.int32)
        ang = np.ascontiguousarray(np.concatenate([sl[1] for sl in self.slices]).astype(np.float32))
        geo = np.array([[sl[2][0], sl[2][1], sl[3][0], sl[3][1], sl[4]] for sl in self.slices],
                       dtype=np.float64)
        h = ctypes.c_void_p()
        call("scxr_xr2d_batch_create", self.ns,
             dims.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
             ang.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
             geo.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), self.vpg, ctypes.byref(h))
        if self._plan is not None:
            # A graph captured from the old plan still reads its tables:
            # keep them until this object goes away instead of freeing them.
            self._retired.append(self._plan)
        self._plan = h
        self.generation += 1
        weakref.finalize(self, lib.scxr_xr2d_batch_destroy, h)
        info = (ctypes.c_int64 * 11)()
        call("scxr_xr2d_batch_info", h, info)
        (_, self.n_img, self.n_sino, self.n_fwd_items, self.n_adj_items, self.n_t_items,
         self.n_blocks, self.Dmax, _, self.n_tiles, self.n_groups) = [int(v) for v in info]
        ioff = (ctypes.c_int64 * self.ns)()
        soff = (ctypes.c_int64 * self.ns)()
        call("scxr_xr2d_batch_offsets", h, ioff, soff)
        self.ioff = np.array(ioff[:], dtype=np.int64)
        self.soff = np.array(soff[:], dtype=np.int64)
        self.configure(*self.config)

    @property
    def plan(self):
        return self._plan

    def configure(self, fwd_grid: int = 0, adj_grid: int = 0, order: int = 0,
                  adj_groups: int = 3):
        """Persistent grid sizes, item order, and view groups per adjoint
        tile (3: every slice's back projection bitwise the single-slice
        one; 2: the views of each pixel summed in two shares, equal to it
        to float32 rounding)."""
        call("scxr_xr2d_batch_configure", self._plan, int(fwd_grid), int(adj_grid), int(order))
        call("scxr_xr2d_batch_set_adjoint_groups", self._plan, int(adj_groups))
        self.config = (int(fwd_grid), int(adj_grid), int(order), int(adj_groups))

    # --- packing
    def pack_images(self, images) -> torch.Tensor:
        out = torch.empty(self.n_img, dtype=torch.float32, device=dev.DEVICE)
        for s, im in enumerate(images):
            M, N = self.input_shapes[s]
            out[self.ioff[s]: self.ioff[s] + M * N].copy_(dev.as_device(im).reshape(-1))
        return out

    def unpack_images(self, packed):
        return [packed[self.ioff[s]: self.ioff[s] + M * N].view(M, N)
                for s, (M, N) in enumerate(self.input_shapes)]

    def pack_sinograms(self, sinos) -> torch.Tensor:
        out = torch.empty(self.n_sino, dtype=torch.float32, device=dev.DEVICE)
        for s, y in enumerate(sinos):
            A, D = self.output_shapes[s]
            out[self.soff[s]: self.soff[s] + A * D].copy_(dev.as_device(y).reshape(-1))
        return out

    def unpack_sinograms(self, packed):
        return [packed[self.soff[s]: self.soff[s] + A * D].view(A, D)
                for s, (A, D) in enumerate(self.output_shapes)]

    # --- projections of the packed batch
    def _check_pred(self, pred, pstride):
        if pred is not None:
            need = (self.ns - 1) * int(pstride) + 1
            dev.check_buffer(pred, need, "pred", np.int32, at_least=True)

    def project(self, im, out=None, y=None, res_ws=None, pred=None, pstride: int = 1):
        """Forward projection of the packed images ``im`` (``n_img``) into
        packed sinograms (``n_sino``); with ``y`` and ``res_ws`` also the
        forward-group partials of ``||y - A im||^2``."""
        dev.check_buffer(im, self.n_img, "im")
        out = dev.output_buffer(out, (self.n_sino,))
        dev.check_buffer(y, self.n_sino, "y", allow_none=True)
        if (y is None) != (res_ws is None):
            raise ValueError("y and res_ws go together")
        dev.check_buffer(res_ws, self.n_groups, "res_ws", np.float64, allow_none=True,
                         at_least=True)
        self._check_pred(pred, pstride)
        call("scxr_xr2d_batch_forward", self._plan, im.data_ptr(), out.data_ptr(), dev.ptr(y),
             dev.ptr(res_ws), dev.ptr(pred), int(pstride), dev.stream())
        return out

    def back_project(self, sino, out=None, kind=0, rho=0.0, circular=False, p=None, z=None,
                     u=None, dot_ws=None, pred=None, pstride: int = 1):
        """Back projection of packed sinograms with the epilogues of
        ``scxr_xr2d_adjoint_epi`` per slice (``kind`` 0, 1 or 2)."""
        dev.check_buffer(sino, self.n_sino, "sino")
        out = dev.output_buffer(out, (self.n_img,))
        if kind not in (0, 1, 2):
            raise ValueError(f"kind must be 0, 1 or 2, got {kind}")
        dev.check_buffer(p, self.n_img, "p", allow_none=kind != 1)
        dev.check_buffer(z, 2 * self.n_img, "z", allow_none=kind != 2)
        dev.check_buffer(u, 2 * self.n_img, "u", allow_none=True)
        dev.check_buffer(dot_ws, self.n_tiles, "dot_ws", np.float64, allow_none=True,
                         at_least=True)
        self._check_pred(pred, pstride)
        call("scxr_xr2d_batch_adjoint", self._plan, sino.data_ptr(), out.data_ptr(), int(kind),
             float(rho), int(circular), dev.ptr(p), dev.ptr(z), dev.ptr(u), dev.ptr(dot_ws),
             dev.ptr(pred), int(pstride), dev.stream())
        return out

    # --- autotuner
    def _time(self, im, sino, out_s, out_i, reps):
        torch.cuda.synchronize()
        self.project(im, out=out_s)
        self.back_project(sino, out=out_i)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            self.project(im, out=out_s)
            self.back_project(sino, out=out_i)
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    def autotune(self, vpgs=(1, 2), grids=(0, 2, 4, 8), orders=(0, 1, 2), reps: int = 3,
                 cu_count: Optional[int] = None, verbose: bool = False, adj_groups=(3,)):
        """Time (views per forward group, persistent grid, item order,
        adjoint view groups) on this batch and keep the fastest; returns the
        table of timings (ms per forward + adjoint pass).  Grid g > 0 means
        g workgroups per CU.  Only vpg 1 with 3 adjoint groups keeps every
        slice bitwise equal to the single-slice path."""
        if cu_count is None:
            cu = ctypes.c_int32()
            call("scxr_device_info", ctypes.byref(cu), None, None, None, 0)
            cu_count = int(cu.value)
        gen = torch.Generator(device=dev.DEVICE).manual_seed(0)
        im = torch.rand(self.n_img, device=dev.DEVICE, generator=gen)
        sino = torch.rand(self.n_sino, device=dev.DEVICE, generator=gen)
        out_s, out_i = torch.empty_like(sino), torch.empty_like(im)
        table = []
        best = None
        for v in vpgs:
            if self.vpg != v:
                self.vpg = v
                try:
                    self._create()
                except Exception:
                    continue  # vpg not supported for this detector size
            for o in orders:
                for gmul in grids:
                    for ag in adj_groups:
                        g = gmul * cu_count
                        self.configure(g, g, o, ag)
                        t = self._time(im, sino, out_s, out_i, reps)
                        table.append({"vpg": v, "order": o, "grid": g, "adj_groups": ag,
                                      "ms": t})
                        if verbose:
                            print(f"  autotune vpg={v} order={o} grid={g} adj_groups={ag}: "
                                  f"{t:.3f} ms", flush=True)
                        if best is None or t < best[0]:
                            best = (t, v, g, o, ag)
        _, v, g, o, ag = best
        if self.vpg != v:
            self.vpg = v
            self._create()
        self.configure(g, g, o, ag)
        self.tuned = {"vpg": v, "grid": g, "order": o, "adj_groups": ag, "ms": best[0],
                      "table": table}
        return self.tuned


class ADMMTVBatch:
    """Fused ADMM-TV on every slice of a batch (x0 = 0, z = C x0, u = 0),
    with ``SquaredL2Loss``, ``lam * L21Norm`` (or ``L1Norm``) of
    ``FiniteDifference(append=0)`` and a CG x-step, as the single-slice
    fused ADMM."""

    def __init__(self, H: XRayTransform2DBatch, y, lam: float, rho: float = 1.0,
                 cg_maxiter: int = 10, cg_tol: float = 1e-4, cg_atol: float = 0.0,
                 isotropic: bool = True, circular: bool = False, hist_cap: int = 64):
        self.H = H
        self.generation = H.generation
        self.y = y if isinstance(y, torch.Tensor) else H.pack_sinograms(y)
        self.lam, self.rho = float(lam), float(rho)
        self.iso, self.circ = bool(isotropic), int(bool(circular))
        self.maxiter = int(cg_maxiter)
        self.tol, self.atol = float(cg_tol), float(cg_atol)
        self.thr = float(np.float32((1.0 / self.rho) * self.lam))
        f32 = dict(dtype=torch.float32, device=dev.DEVICE)
        f64 = dict(dtype=torch.float64, device=dev.DEVICE)
        n, ns = H.n_img, H.ns
        self.x = torch.zeros(n, **f32)
        self.z = torch.zeros(2 * n, **f32)
        self.u = torch.zeros(2 * n, **f32)
        self.zold = torch.zeros(2 * n, **f32)
        self.r, self.p = torch.empty(n, **f32), torch.empty(n, **f32)
        self.Ap, self.rhs = torch.empty(n, **f32), torch.empty(n, **f32)
        self.sino = torch.empty(H.n_sino, **f32)
        # A x of the current iterate (x = 0: exactly 0), left by the residual
        # projection that ends an iteration for the x-step of the next
        self.ax = torch.zeros(H.n_sino, **f32)
        # A^T y for the x-step right-hand side (y is fixed), back projected
        # again only when the plan's configuration changes (the adjoint's view
        # groups decide its rounding)
        self.aty = torch.empty(n, **f32)
        self._aty_key = None
        self.SS = ST_NUM + 2 * (self.maxiter + 1)
        self.state = torch.zeros(ns * self.SS, **f64)
        self.flags = torch.zeros(ns * (self.maxiter + 1), dtype=torch.int32, device=dev.DEVICE)
        self.ws = torch.empty(2 * H.n_blocks, **f64)
        self.pap = torch.empty(H.n_tiles, **f64)
        self.ws_prml = torch.empty(H.n_blocks, **f64)
        self.ws_g = torch.empty(H.n_blocks, **f64)
        self.ws_dual = torch.empty(H.n_blocks, **f64)
        self.ws_res = torch.empty(H.n_groups, **f64)
        self.hist_cap = int(hist_cap)
        self.hist = torch.zeros((ns, self.hist_cap, 5), **f64)
        self.counters = torch.zeros(ns, dtype=torch.int32, device=dev.DEVICE)

    def reset(self):
        """Restart from x = 0 (z = C 0 = 0, u = 0) with empty statistics;
        device fills only, so it may run between graph replays."""
        for t in (self.x, self.z, self.u, self.zold, self.ax):
            vec.fill(t, 0.0)
        vec.fill(self.counters.view(torch.float32), 0.0)

    def enqueue(self):
        """One ADMM iteration on every slice (no host synchronisation).

        The workspaces are sized for the plan ``H`` had when this solver was
        built; if ``H`` was re-created since (``autotune``), this raises."""
        H, P, s = self.H, (lambda t: t.data_ptr()), dev.stream()
        if H.generation != self.generation:
            raise RuntimeError("the batch plan was re-created (autotune) after this solver was "
                               "built; build a new ADMMTVBatch")
        F = self.maxiter + 1
        key = (H.config, self.y._version)
        if self._aty_key != key:  # new configuration, or y modified in place
            H.back_project(self.y, out=self.aty)
            self._aty_key = key
        call("scxr_batch_admm_rhs", H.plan, P(self.aty), P(self.z), P(self.u), self.rho,
             self.circ, P(self.rhs), s)
        H.back_project(self.ax, out=self.Ap, kind=1, rho=self.rho, circular=self.circ, p=self.x)
        call("scxr_batch_cg_start", H.plan, P(self.rhs), P(self.Ap), P(self.r), P(self.p),
             P(self.state), P(self.flags), self.maxiter, self.tol, self.atol, P(self.ws), s)
        for k in range(self.maxiter):
            flag = self.flags[k:]
            H.project(self.p, out=self.sino, pred=flag, pstride=F)
            H.back_project(self.sino, out=self.Ap, kind=1, rho=self.rho, circular=self.circ,
                           p=self.p, dot_ws=self.pap, pred=flag, pstride=F)
            call("scxr_batch_cg_iter", H.plan, k, self.maxiter, P(self.x), P(self.r), P(self.p),
                 P(self.Ap), P(self.state), P(self.flags), P(self.pap), P(self.ws), s)
        call("scxr_batch_admm_zu", H.plan, P(self.x), P(self.z), P(self.u), P(self.zold),
             self.thr, self.circ, int(self.iso), P(self.ws_prml), P(self.ws_g), s)
        call("scxr_batch_admm_dual", H.plan, P(self.z), P(self.zold), self.rho, self.circ,
             P(self.ws_dual), s)
        H.project(self.x, out=self.ax, y=self.y, res_ws=self.ws_res)
        call("scxr_batch_admm_itstat", H.plan, P(self.ws_res), P(self.ws_g), P(self.ws_prml),
             P(self.ws_dual), P(self.state), self.maxiter, self.lam, self.rho, P(self.hist),
             P(self.counters), self.hist_cap, s)

    def images(self):
        return self.H.unpack_images(self.x)
