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


def fbp(H, y):
    """``x = (pi dx0 dx1 / A) * mask * H^T(ramp * y)``.

    The rows are filtered by direct linear convolution (same result as the
    reference's zero-padded FFT, up to float32 rounding).  For even
    ``det_count`` the reference's slice keeps ``det_count + 1`` filtered
    bins; that row length is reproduced, and back projected with the same
    geometry over ``det_count + 1`` bins.
    """
    y = dev.as_device(y)
    A, N = y.shape
    s = (N - 1) // 2
    Dout = (2 * N - 1) - 2 * s
    if H.fbp_filter is None:
        H.fbp_filter = dev.as_device(ramp_filter(np.arange(N) - s, 1.0))
    if H.fbp_mask is None:
        ones = vec.fill(dev.empty(H.output_shape), 1.0)
        H.fbp_mask = vec.ge(H.back_project(ones), np.float32(A * (1.0 - 1e-5)))
    hy = dev.empty((A, Dout))
    call("scxr_ramp_filter", A, N, Dout, s, H.fbp_filter.data_ptr(), y.data_ptr(),
         hy.data_ptr(), None, dev.stream())
    if Dout == N:
        bp = H.back_project(hy)
    else:
        Hx = getattr(H, "_fbp_wide", None)
        if Hx is None:
            from ._xray2d import XRayTransform2D

            Hx = XRayTransform2D(H.input_shape, H.angles, x0=H.x0, dx=H.dx, y0=H.y0,
                                 det_count=Dout)
            H._fbp_wide = Hx
        bp = Hx.back_project(hy)
    c = float(np.float32(np.pi * float(np.asarray(H.dx[0])) * float(np.asarray(H.dx[1])) / A))
    return vec.mul(vec.scale(c, H.fbp_mask), bp)
