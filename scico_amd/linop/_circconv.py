"""Circular convolution on the HIP FFT kernels.

Drop-in for ``scico.linop.CircularConvolve`` (``scico/linop/_circconv.py``)
for real float32 input: ``A x = real(ifftn(h_dft * fftn(x)))`` over the
trailing ``ndims`` axes (leading axes are batch axes), adjoint with
``conj(h_dft)``.  ``h_dft = fftn(h, s=shape)`` is formed on the device;
the ``h_center`` shift follows ``_circconv.py:155-186``.
"""

from __future__ import annotations

import ctypes
import math
import weakref

import numpy as np
import torch

from .. import _device as dev
from .. import _pred
from .._lib import call, load
from ._linop import LinearOperator


def dft_center_shift(fft_shape, h_center) -> np.ndarray:
    """Complex64 centring shift (``_circconv.py:155-186``)."""
    if isinstance(h_center, (float, int)):
        offset = -np.array([h_center])
    else:
        offset = -np.array(dev.to_numpy(h_center) if isinstance(h_center, torch.Tensor)
                           else h_center).reshape(-1)
    factors = np.ix_(*tuple(
        np.select(
            [np.arange(s) < s / 2, np.arange(s) == s / 2, np.arange(s) > s / 2],
            [np.exp(-1j * k * 2 * np.pi * np.arange(s) / s), np.cos(k * np.pi),
             np.exp(1j * k * 2 * np.pi * (s - np.arange(s)) / s)],
        )
        for k, s in zip(offset, fft_shape)
    ))
    factors = tuple(f.astype(np.complex64) for f in factors)
    return np.asarray(math.prod(factors), dtype=np.complex64)


class CircularConvolve(LinearOperator):
    """``A x = h * x`` (circular) over the trailing ``ndims`` axes."""

    def __init__(self, h, input_shape, ndims=None, input_dtype=np.float32, h_is_dft=False,
                 h_center=None, jit=True, **kwargs):
        input_shape = tuple(int(s) for s in input_shape)
        self.ndims = len(input_shape) if ndims is None else int(ndims)
        if h_is_dft and h_center is not None:
            raise ValueError("Argument 'h_center' must be None when h_is_dft=True.")
        if np.dtype(input_dtype) != np.float32:
            raise NotImplementedError("CircularConvolve supports real float32 input.")
        if not 1 <= self.ndims <= 3:
            raise NotImplementedError("CircularConvolve supports 1 to 3 convolution axes.")
        self.h_center = h_center
        self.fft_shape = input_shape[-self.ndims:]
        self.batch = int(np.prod(input_shape[: len(input_shape) - self.ndims], dtype=np.int64))
        hh = h if isinstance(h, torch.Tensor) else np.asarray(h)
        if hh.ndim > self.ndims and all(int(n) == 1 for n in hh.shape[: hh.ndim - self.ndims]):
            # leading singleton axes broadcast over the batch axes
            hh = hh.reshape(tuple(hh.shape[hh.ndim - self.ndims:]))
        if hh.ndim != self.ndims:
            raise NotImplementedError("Filter batch axes (multiple filters) are not supported.")
        self._h_is_dft = bool(h_is_dft)
        self._h_src = hh
        self._plan = None
        self.x_fft_axes = list(range(len(input_shape) - self.ndims, len(input_shape)))
        self.ifft_axes = list(self.x_fft_axes)
        self.real = True
        super().__init__(input_shape=input_shape, output_shape=input_shape,
                         input_dtype=np.float32, output_dtype=np.float32)

    @property
    def plan(self):
        if self._plan is None:
            dev.require_gpu()
            lib = load()
            fs = (ctypes.c_int32 * self.ndims)(*self.fft_shape)
            h = ctypes.c_void_p()
            call("scxr_cconv_plan_create", self.ndims, fs, self.batch, ctypes.byref(h))
            self._plan = h
            self._fin = weakref.finalize(self, lib.scxr_cconv_plan_destroy, h)
            if self._h_is_dft:
                hd = self._h_src
                hd = hd if isinstance(hd, torch.Tensor) else torch.as_tensor(
                    np.ascontiguousarray(hd, dtype=np.complex64))
                hd = hd.to(dev.DEVICE, torch.complex64).contiguous()
                if tuple(hd.shape) != self.fft_shape:
                    raise ValueError(f"h_dft shape {tuple(hd.shape)} != {self.fft_shape}")
                call("scxr_cconv_set_spectrum", h, hd.data_ptr(), 1, None, dev.stream())
            else:
                # zero-pad (or crop) h to the FFT shape, as fftn(h, s=shape) does
                hs = dev.to_numpy(self._h_src).astype(np.float32)
                hp = np.zeros(self.fft_shape, np.float32)
                sl = tuple(slice(0, min(a, b)) for a, b in zip(hs.shape, self.fft_shape))
                hp[sl] = hs[sl]
                hpd = dev.as_device(hp)
                sh = None
                if self.h_center is not None:
                    sh = torch.as_tensor(dft_center_shift(self.fft_shape, self.h_center),
                                         device=dev.DEVICE)
                call("scxr_cconv_set_spectrum", h, hpd.data_ptr(), 0, dev.ptr(sh), dev.stream())
                torch.cuda.current_stream().synchronize()  # hpd / sh are temporaries
        return self._plan

    @property
    def h_dft(self) -> torch.Tensor:
        out = torch.empty(self.fft_shape, dtype=torch.complex64, device=dev.DEVICE)
        call("scxr_cconv_spectrum", self.plan, out.data_ptr(), dev.stream())
        return out

    def _apply(self, x, adjoint, out=None, pred=None):
        x = dev.input_array(x, self.input_shape, "x")
        out = dev.output_buffer(out, self.input_shape)
        call("scxr_cconv_apply", self.plan, x.data_ptr(), out.data_ptr(), int(adjoint),
             dev.ptr(_pred.resolve(pred)), dev.stream())
        return out

    def _eval(self, x, out=None, pred=None):
        return self._apply(x, False, out, pred)

    def _adj(self, y, out=None, pred=None):
        return self._apply(y, True, out, pred)

    def fftn(self, x, inverse=False):
        """Complex fftn (ifftn) of ``x`` over the convolution axes."""
        x = x if isinstance(x, torch.Tensor) else torch.as_tensor(np.asarray(x))
        x = x.to(dev.DEVICE).contiguous()
        if tuple(x.shape) != tuple(self.input_shape):
            raise ValueError(f"Shapes do not conform: x array shape {tuple(x.shape)} != "
                             f"expected shape {tuple(self.input_shape)}")
        real = not x.is_complex()
        if real:
            x = x.to(torch.float32)
        else:
            x = x.to(torch.complex64)
        out = torch.empty(self.input_shape, dtype=torch.complex64, device=dev.DEVICE)
        call("scxr_fftn", self.plan, x.data_ptr(), int(real), out.data_ptr(), int(inverse),
             None, dev.stream())
        return out

    @classmethod
    def from_operator(cls, H, ndims=None, center=None, device=None, jit=True):
        """``CircularConvolve`` from the impulse response of an LSI operator
        (``_circconv.py:258-307``)."""
        ndims = len(H.input_shape) if ndims is None else ndims
        if center is None:
            center = tuple(d // 2 for d in H.input_shape[-ndims:])
        d = np.zeros(H.input_shape, np.float32)
        d[(Ellipsis,) + tuple(center)] = 1.0
        Hd = dev.to_numpy(H(d))
        return cls(Hd, H.input_shape, ndims=ndims, input_dtype=H.input_dtype,
                   h_center=np.array(center))
