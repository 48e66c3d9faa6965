"""Differentiable projector and unrolled reconstruction for PyTorch.

* :class:`XRayProject` / :class:`XRayBackProject`: ``torch.autograd.Function``
  wrappers of an X-ray transform (2D or 3D) whose backward is the other
  hand-written kernel (the adjoint is exact, so the VJP of the projection
  is the back projection and vice versa).  This is the role of the
  ``jax.custom_vjp`` that scico's native projector wrappers carry
  (``scico/linop/xray/astra/_astra_2d.py:140-143``): ``SquaredL2Loss.grad``
  differentiates through ``A``.
* ``torch.ops.scico_amd.xray_project`` / ``xray_back_project``: the same
  pair registered as PyTorch custom ops (``torch.library``), keyed by an
  operator handle from :func:`register_operator`.
* :func:`unrolled_apgm_tv`: ``K`` AcceleratedPGM-TV steps
  (``scico/optimize/_pgm.py:183-252``) with a per-pixel TV weight map,
  differentiable with respect to the sinogram and the weight map.  The
  backward pass recomputes the forward from checkpoints every
  ``checkpoint_every`` steps and runs the adjoint recursion with the
  projector kernels and the TV-prox VJP kernel.
"""

from __future__ import annotations

import ctypes
import itertools
import weakref
from typing import List

import numpy as np
import torch

from . import _device as dev
from . import _vec as vec
from ._lib import call, load

f32 = np.float32


# ----------------------------------------------------------- projector --


class XRayProject(torch.autograd.Function):
    """``y = H x``; backward ``dx = H^T dy``."""

    @staticmethod
    def forward(ctx, x, H):
        ctx.H = H
        return H.project(x.detach().contiguous())

    @staticmethod
    def backward(ctx, gy):
        return ctx.H.back_project(gy.contiguous()), None


class XRayBackProject(torch.autograd.Function):
    """``x = H^T y``; backward ``dy = H dx``."""

    @staticmethod
    def forward(ctx, y, H):
        ctx.H = H
        return H.back_project(y.detach().contiguous())

    @staticmethod
    def backward(ctx, gx):
        return ctx.H.project(gx.contiguous()), None


_OPS: dict = {}
_next_handle = itertools.count(1)


def register_operator(H) -> int:
    """Handle for ``torch.ops.scico_amd.xray_project(x, handle)``."""
    h = next(_next_handle)
    _OPS[h] = weakref.ref(H)
    return h


def _op(handle: int):
    H = _OPS.get(int(handle))
    H = H() if H is not None else None
    if H is None:
        raise ValueError(f"no live X-ray operator registered under handle {handle}")
    return H


@torch.library.custom_op("scico_amd::xray_project", mutates_args=())
def _xray_project(x: torch.Tensor, handle: int) -> torch.Tensor:
    return _op(handle).project(x.contiguous())


@torch.library.custom_op("scico_amd::xray_back_project", mutates_args=())
def _xray_back_project(y: torch.Tensor, handle: int) -> torch.Tensor:
    return _op(handle).back_project(y.contiguous())


@_xray_project.register_fake
def _(x, handle):
    return x.new_empty(_op(handle).output_shape)


@_xray_back_project.register_fake
def _(y, handle):
    return y.new_empty(_op(handle).input_shape)


def _proj_bwd(ctx, g):
    return torch.ops.scico_amd.xray_back_project(g, ctx.handle), None


def _bproj_bwd(ctx, g):
    return torch.ops.scico_amd.xray_project(g, ctx.handle), None


def _save_handle(ctx, inputs, output):
    ctx.handle = inputs[1]


torch.library.register_autograd("scico_amd::xray_project", _proj_bwd, setup_context=_save_handle)
torch.library.register_autograd("scico_amd::xray_back_project", _bproj_bwd,
                                setup_context=_save_handle)


# ------------------------------------------------- weighted TV prox --


def _tau(lam: float, nd: int) -> float:
    """Threshold ``sqrt(2) K lam`` in float32, as ``TVNorm.prox``."""
    return float(f32(f32(np.sqrt(f32(2.0))) * f32(2 * nd)) * f32(lam))


def tv_prox_weighted(v, lam, w=None, out=None, isotropic=True, circular=False):
    """Haar TV prox of ``lam * sum_p w[p] |grad x|_p`` (``w=None``: TVNorm.prox)."""
    shape = tuple(v.shape)
    nd = len(shape)
    sh = (ctypes.c_int32 * nd)(*shape)
    out = torch.empty_like(v) if out is None else out
    call("scxr_tv_prox_w", nd, sh, v.data_ptr(), dev.ptr(w), out.data_ptr(), _tau(lam, nd),
         int(isotropic), int(circular), None, dev.stream())
    return out


def tv_prox_weighted_vjp(v, lam, w, gout, gv=None, gw=None, isotropic=True, circular=False):
    """``(gv, gw)``: cotangents of ``v`` and ``w`` for cotangent ``gout``."""
    shape = tuple(v.shape)
    nd = len(shape)
    sh = (ctypes.c_int32 * nd)(*shape)
    nws = int(load().scxr_tv_prox_vjp_ws(nd, sh, int(circular)))
    ws = torch.empty(nws, dtype=torch.float32, device=v.device)
    gv = torch.empty_like(v) if gv is None else gv
    if w is not None and gw is None:
        gw = torch.empty_like(w)
    call("scxr_tv_prox_vjp", nd, sh, v.data_ptr(), dev.ptr(w), gout.data_ptr(), _tau(lam, nd),
         int(isotropic), int(circular), gv.data_ptr(), dev.ptr(gw if w is not None else None),
         ws.data_ptr(), dev.stream())
    return gv, gw


# ---------------------------------------------- unrolled APGM-TV --


def _momentum(K: int) -> List[float]:
    """AcceleratedPGM's float32 ``(t_old - 1) / t`` for steps 0..K-1."""
    cs, t = [], 1.0
    for _ in range(K):
        t_old = t
        t = float(f32(0.5) * (f32(1) + np.sqrt(f32(1) + f32(4) * f32(t_old) ** 2)))
        cs.append(float(f32(f32(t_old) - f32(1)) / f32(t)))
    return cs


def _step(H, y, v, x_old, w, L, lam, c, iso):
    """One AcceleratedPGM step; returns ``(x, v_next, u)``."""
    r = vec.add(H.project(v), -1.0, y)
    u = vec.axpby(1.0, v, -(1.0 / L), H.back_project(r))
    x = tv_prox_weighted(u, (1.0 / L) * lam, w, isotropic=iso)
    v_next = vec.axpby(1.0, x, c, vec.add(x, -1.0, x_old))
    return x, v_next, u


class _UnrolledAPGMTV(torch.autograd.Function):
    @staticmethod
    def forward(ctx, y, w, x0, H, L, lam, K, every, iso):
        y, w = y.detach().contiguous(), w.detach().contiguous()
        cs = _momentum(K)
        x = x0.detach().contiguous().clone()
        v = x.clone()
        ckpts = {}
        for k in range(K):
            if k % every == 0:
                ckpts[k] = (x.clone(), v.clone())
            x, v, _ = _step(H, y, v, x, w, L, lam, cs[k], iso)
        ctx.H, ctx.L, ctx.lam, ctx.K, ctx.every, ctx.iso, ctx.cs = H, L, lam, K, every, iso, cs
        ctx.ckpts = ckpts
        ctx.save_for_backward(y, w)
        return x

    @staticmethod
    def backward(ctx, gx):
        y, w = ctx.saved_tensors
        H, L, lam, K, every, iso, cs = ctx.H, ctx.L, ctx.lam, ctx.K, ctx.every, ctx.iso, ctx.cs
        gam = 1.0 / L
        xbar = gx.detach().contiguous().clone()  # cotangent of x_{k+1}
        vbar = torch.zeros_like(xbar)  # cotangent of v_{k+1}
        ybar = torch.zeros_like(y)
        wbar = torch.zeros_like(w)
        gw = torch.empty_like(w)
        for s in reversed(sorted(ctx.ckpts)):
            # recompute the segment's prox inputs u_k from the checkpoint
            x, v = (t.clone() for t in ctx.ckpts[s])
            us = []
            for k in range(s, min(s + every, K)):
                x, v, u = _step(H, y, v, x, w, L, lam, cs[k], iso)
                us.append(u)
            for k in reversed(range(s, min(s + every, K))):
                u = us[k - s]
                # v_{k+1} = x_{k+1} + c_k (x_{k+1} - x_k)
                xk1 = vec.axpby(1.0, xbar, 1.0 + cs[k], vbar)
                xk_part = vec.scale(-cs[k], vbar)
                # x_{k+1} = prox(u_k; w)
                ubar, _ = tv_prox_weighted_vjp(u, gam * lam, w, xk1, gw=gw, isotropic=iso)
                vec.add(wbar, 1.0, gw, out=wbar)
                # u_k = v_k - gam A^T (A v_k - y)
                Au = H.project(ubar)
                vec.axpby(1.0, ybar, gam, Au, out=ybar)
                vbar = vec.axpby(1.0, ubar, -gam, H.back_project(Au))
                xbar = xk_part
        return ybar, wbar, None, None, None, None, None, None, None


def unrolled_apgm_tv(H, y, w, L: float, lam: float, K: int = 20, checkpoint_every: int = 5,
                     x0=None, isotropic: bool = True):
    """``x_K`` of ``K`` AcceleratedPGM steps on
    ``0.5 ||H x - y||^2 + lam * sum_p w[p] |grad x|_p`` (Haar TV prox),
    step ``1/L``, from ``x0`` (zeros by default).  Differentiable with
    respect to ``y`` and ``w`` (torch autograd)."""
    if x0 is None:
        x0 = torch.zeros(H.input_shape, dtype=torch.float32, device=dev.DEVICE)
    return _UnrolledAPGMTV.apply(y, w, x0, H, float(L), float(lam), int(K),
                                 int(checkpoint_every), bool(isotropic))
