"""Finite-difference operator on the HIP stencil kernels.

Mirrors ``scico.linop.FiniteDifference`` (``scico/linop/_diff.py:25-272``)
for the boundary handling used on the hot path: ``append=0`` (a trailing
zero difference, used by ADMM-TV) and ``circular=True``.  The output is
the stacked ``(ndim, *input_shape)`` array that ``VerticalStack``
collapses to (``scico/operator/_stack.py:126-128``).
"""

from __future__ import annotations

import ctypes

import numpy as np

from .. import _device as dev
from .. import _pred
from .._lib import call
from ._linop import LinearOperator


class FiniteDifference(LinearOperator):
    """Stack of per-axis forward differences over all axes (ndim <= 3)."""

    def __init__(self, input_shape, input_dtype=np.float32, axes=None, prepend=None,
                 append=None, circular=False, jit=True, **kwargs):
        input_shape = tuple(int(s) for s in input_shape)
        nd = len(input_shape)
        if axes is not None:
            ax = (axes,) if isinstance(axes, int) else tuple(axes)
            ax = tuple(a + nd if a < 0 else a for a in ax)
            if any(a >= nd for a in ax):
                raise ValueError(f"Invalid axes {axes} for input_shape {input_shape}.")
            if tuple(sorted(ax)) != tuple(range(nd)):
                raise NotImplementedError("Only differences over all axes are supported.")
        if prepend not in (None, 0, 1):
            raise ValueError(f"Invalid value {prepend} for argument 'prepend'.")
        if append not in (None, 0, 1):
            raise ValueError(f"Invalid value {append} for argument 'append'.")
        if circular and (prepend is not None or append is not None):
            raise ValueError("Arguments 'prepend' and 'append' must be None when circular=True.")
        if not circular and not (prepend is None and append == 0):
            raise NotImplementedError(
                "Only append=0 (with prepend=None) or circular=True boundary handling is supported."
            )
        if not 1 <= nd <= 3:
            raise NotImplementedError("FiniteDifference supports 1 to 3 dimensions.")
        if np.dtype(input_dtype) != np.float32:
            raise NotImplementedError("FiniteDifference supports float32 input.")
        self.axes = tuple(range(nd))
        self.circular = bool(circular)
        self.append = append
        self.prepend = prepend
        self._shape_c = (ctypes.c_int32 * nd)(*input_shape)
        super().__init__(input_shape=input_shape, output_shape=(nd, *input_shape),
                         input_dtype=np.float32, output_dtype=np.float32)

    def _eval(self, x, out=None, pred=None):
        x = dev.input_array(x, self.input_shape, "x")
        out = dev.output_buffer(out, self.output_shape)
        call("scxr_fd_forward", len(self.input_shape), self._shape_c, x.data_ptr(),
             out.data_ptr(), int(self.circular), dev.ptr(_pred.resolve(pred)), dev.stream())
        return out

    def _adj(self, g, out=None, scale=1.0, pred=None):
        g = dev.input_array(g, self.output_shape, "g")
        out = dev.output_buffer(out, self.input_shape)
        call("scxr_fd_adjoint", len(self.input_shape), self._shape_c, g.data_ptr(),
             out.data_ptr(), int(self.circular), float(scale), dev.ptr(_pred.resolve(pred)),
             dev.stream())
        return out

    def gram(self, x, rho=1.0, out=None, pred=None):
        """``rho * C^T C x`` in one stencil pass."""
        x = dev.input_array(x, self.input_shape, "x")
        out = dev.output_buffer(out, self.input_shape)
        call("scxr_fd_gram", len(self.input_shape), self._shape_c, float(rho), x.data_ptr(),
             out.data_ptr(), int(self.circular), dev.ptr(_pred.resolve(pred)), dev.stream())
        return out
