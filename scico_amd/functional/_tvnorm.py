"""Total variation norms (``scico/functional/_tvnorm.py:33-353``).

``__call__`` is norm(FiniteDifference(x)) (append=0 unless circular) and
``prox`` is the Haar-based approximate prox, both as fused HIP kernels.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import _device as dev
from .. import _pred
from .._lib import call, load
from ._functional import Functional


class TVNorm(Functional):
    has_eval = True
    has_prox = True

    def __init__(self, isotropic: bool, circular: bool = False, axes=None, input_shape=None,
                 input_dtype=np.float32):
        if axes is not None:
            raise NotImplementedError("TV over a subset of axes is not supported.")
        self.isotropic = bool(isotropic)
        self.circular = bool(circular)
        self.axes = axes
        self.input_shape = None if input_shape is None else tuple(input_shape)

    def _shape(self, x):
        shape = tuple(x.shape)
        if not 1 <= len(shape) <= 3:
            raise NotImplementedError("TVNorm supports 1 to 3 dimensions.")
        return (ctypes.c_int32 * len(shape))(*shape), len(shape)

    def __call__(self, x):
        x = dev.as_device(x)
        shp, nd = self._shape(x)
        out = torch.empty((), dtype=torch.float64, device=dev.DEVICE)
        ws = torch.empty(int(load().scxr_elem_blocks(x.numel())) + 1, dtype=torch.float64,
                         device=dev.DEVICE)
        call("scxr_tv_norm", nd, shp, x.data_ptr(), int(self.isotropic), int(self.circular),
             out.data_ptr(), ws.data_ptr(), dev.ptr(_pred.current()), dev.stream())
        return out

    def prox(self, v, lam=1.0, out=None, **kwargs):
        """``_prox_core`` with threshold ``sqrt(2) * K * lam`` (float32 order)."""
        v = dev.as_device(v)
        shp, nd = self._shape(v)
        K = 2 * nd
        tau = float(np.float32(np.float32(np.sqrt(np.float32(2.0))) * np.float32(K)) *
                    np.float32(lam))
        out = torch.empty_like(v) if out is None else out
        call("scxr_tv_prox", nd, shp, v.data_ptr(), out.data_ptr(), tau, int(self.isotropic),
             int(self.circular), dev.ptr(_pred.current()), dev.stream())
        return out


class IsotropicTVNorm(TVNorm):
    """Isotropic TV (L21 of the gradient), ``_tvnorm.py:297-353``."""

    def __init__(self, circular: bool = False, axes=None, input_shape=None,
                 input_dtype=np.float32):
        super().__init__(True, circular, axes, input_shape, input_dtype)


class AnisotropicTVNorm(TVNorm):
    """Anisotropic TV (L1 of the gradient), ``_tvnorm.py:239-295``."""

    def __init__(self, circular: bool = False, axes=None, input_shape=None,
                 input_dtype=np.float32):
        super().__init__(False, circular, axes, input_shape, input_dtype)
