"""L1 and L21 norms (``scico/functional/_norm.py:65-116, 208-296``) on the
HIP prox and reduction kernels."""

from __future__ import annotations

import torch

from .. import _device as dev
from .. import _pred
from .. import _vec as vec
from .._lib import call
from ._functional import Functional


class SquaredL2Norm(Functional):
    """``sum |x|^2``; prox ``v / (1 + 2 lam)`` (``_norm.py:119-158``)."""

    has_eval = True
    has_prox = True

    def __call__(self, x):
        return vec.reduce(vec.RED_SUMSQ, dev.as_device(x))

    def prox(self, v, lam=1.0, **kwargs):
        return vec.scale(1.0 / (1.0 + 2.0 * float(lam)), dev.as_device(v))


class L1Norm(Functional):
    """``sum |x|``; prox is soft thresholding."""

    has_eval = True
    has_prox = True

    def __call__(self, x):
        return vec.reduce(vec.RED_L1, dev.as_device(x))

    def prox(self, v, lam=1.0, out=None, **kwargs):
        v = dev.as_device(v)
        out = torch.empty_like(v) if out is None else out
        call("scxr_l1_prox", v.numel(), v.data_ptr(), float(lam), out.data_ptr(),
             dev.ptr(_pred.current()), dev.stream())
        return out


class L21Norm(Functional):
    """``sum_j ||x[:, j]||_2`` with the l2 norm over ``l2_axis`` (0 only)."""

    has_eval = True
    has_prox = True

    def __init__(self, l2_axis=0):
        if l2_axis != 0:
            raise NotImplementedError("L21Norm supports l2_axis=0 (the TV stack axis).")
        self.l2_axis = l2_axis

    def __call__(self, x):
        x = dev.as_device(x)
        return vec.reduce(vec.RED_L21, x, K=x.shape[0])

    def prox(self, v, lam=1.0, out=None, **kwargs):
        v = dev.as_device(v)
        out = torch.empty_like(v) if out is None else out
        K = v.shape[0]
        call("scxr_l21_prox", K, v.numel() // K, v.data_ptr(), float(lam), out.data_ptr(),
             dev.ptr(_pred.current()), dev.stream())
        return out
