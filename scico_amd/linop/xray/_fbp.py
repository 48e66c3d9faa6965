"""Filtered back projection (``scico/linop/xray/_xray2d.py:158-221``)."""

from __future__ import annotations

import numpy as np

from ... import _device as dev
from ... import _vec as vec
from ..._lib import call


def ramp_filter(x, tau: float) -> np.ndarray:
    """Kak & Slaney eq. 61 taps (``_xray2d.py:199-221``), float32."""
    x = np.asarray(x, dtype=np.int32)
    out = np.zeros(x.shape, np.float32)
    out[x == 0] = np.float32(1.0 / (4.0 * tau**2))
    odd = (x % 2) != 0
    out[odd] = (-1.0 / (x[odd].astype(np.float64) ** 2 * np.pi**2 * tau**2)).astype(np.float32)
    return out


def filter_window(N: int) -> tuple[int, int]:
    """First index and length of the filtered row within the full linear
    convolution of length ``2 N - 1``.  The reference slices
    ``[(N - 1) // 2 : -(N - 1) // 2]``, where ``-(N - 1) // 2`` is
    ``floor(-(N - 1) / 2)``, so the row keeps exactly ``N`` bins for every
    ``N`` (``_xray2d.py:192-194``)."""
    r = range(2 * N - 1)[(N - 1) // 2 : -(N - 1) // 2]
    if len(r) != N:  # N = 1: the slice [0:0] is empty
        raise ValueError(f"fbp needs at least 2 detector bins, got {N}")
    return r.start, len(r)


def fbp(H, y):
    """``x = (pi dx0 dx1 / A) * mask * H^T(ramp * y)``.

    The rows are filtered by direct linear convolution (same result as the
    reference's zero-padded FFT, up to float32 rounding), keeping the
    window of the full convolution that the reference's slice
    ``[(N - 1) // 2 : -(N - 1) // 2]`` keeps (``_xray2d.py:192-194``).
    """
    y = dev.as_device(y)
    if tuple(y.shape) != tuple(H.output_shape):
        raise ValueError(f"Shapes do not conform: input array shape {tuple(y.shape)} != "
                         f"output_shape {tuple(H.output_shape)}")
    A, N = y.shape
    s, Dout = filter_window(N)
    if H.fbp_filter is None:
        H.fbp_filter = dev.as_device(ramp_filter(np.arange(N) - s, 1.0))
    if H.fbp_mask is None:
        ones = vec.fill(dev.empty(H.output_shape), 1.0)
        H.fbp_mask = vec.ge(H.back_project(ones), np.float32(A * (1.0 - 1e-5)))
    hy = dev.empty((A, Dout))
    call("scxr_ramp_filter", A, N, Dout, s, H.fbp_filter.data_ptr(), y.data_ptr(),
         hy.data_ptr(), None, dev.stream())
    bp = H.back_project(hy)
    c = float(np.float32(np.pi * float(np.asarray(H.dx[0])) * float(np.asarray(H.dx[1])) / A))
    return vec.mul(vec.scale(c, H.fbp_mask), bp)
