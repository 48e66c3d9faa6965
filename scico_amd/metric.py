"""Image quality metrics (``scico/metric.py:20-163``).

Sums run in the library's float64 reduction kernel; each function returns
a Python float.
"""

from __future__ import annotations

import math
from typing import Optional

from . import _device as dev
from . import _vec as vec


def _f(t) -> float:
    return float(t.item())


def _n(x) -> int:
    return int(x.numel())


def mae(reference, comparison) -> float:
    """Mean absolute error (``metric.py:20-31``)."""
    r, c = dev.as_device(reference), dev.as_device(comparison)
    return _f(vec.reduce(vec.RED_L1, vec.add(r, -1.0, c))) / _n(r)


def mse(reference, comparison) -> float:
    """Mean squared error (``metric.py:34-45``)."""
    r, c = dev.as_device(reference), dev.as_device(comparison)
    return _f(vec.reduce(vec.RED_SQDIST, r, c)) / _n(r)


def _var(x) -> float:
    n = _n(x)
    m = _f(vec.reduce(vec.RED_SUM, x)) / n
    return max(_f(vec.reduce(vec.RED_SUMSQ, x)) / n - m * m, 0.0)


def _db(rt: float) -> float:
    if rt == 0.0:
        return -math.inf
    return 10.0 * math.log10(rt) if math.isfinite(rt) else math.inf


def snr(reference, comparison) -> float:
    """``10 log10(var(reference) / mse)`` (``metric.py:48-62``)."""
    r = dev.as_device(reference)
    e = mse(r, comparison)
    return _db(_var(r) / e if e > 0 else math.inf)


def psnr(reference, comparison, signal_range: Optional[float] = None) -> float:
    """``10 log10(range^2 / mse)``, range = max - min of the reference by
    default (``metric.py:65-92``)."""
    r = dev.as_device(reference)
    if signal_range is None:
        signal_range = abs(float(r.max().item()) - float(r.min().item()))
    e = mse(r, comparison)
    return _db(signal_range**2 / e if e > 0 else math.inf)


def isnr(reference, degraded, restored) -> float:
    """``10 log10(mse(ref, degraded) / mse(ref, restored))`` (``metric.py:95-118``)."""
    e = mse(reference, restored)
    return _db(mse(reference, degraded) / e if e > 0 else math.inf)


def bsnr(blurry, noisy) -> float:
    """``10 log10(var(blurry) / var(noisy - blurry))`` (``metric.py:121-139``)."""
    b, n = dev.as_device(blurry), dev.as_device(noisy)
    v = _var(vec.add(n, -1.0, b))
    return _db(_var(b) / v if v > 0 else math.inf)


def rel_res(ax, b) -> float:
    """``||b - ax|| / max(||ax||, ||b||)``, 0 when both are 0 (``metric.py:142-163``)."""
    a, bb = dev.as_device(ax), dev.as_device(b)
    nrm = max(_f(vec.norm(a)), _f(vec.norm(bb)))
    if nrm == 0.0:
        return 0.0
    return _f(vec.reduce(vec.RED_SQDIST, bb, a)) ** 0.5 / nrm
