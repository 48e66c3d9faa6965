"""Python wrappers of the vector kernels (``scxr_vec_*``).

Reductions return 0-d float64 device tensors ("device scalars"): nothing
is copied to the host until the caller asks for it.
"""

from __future__ import annotations

import torch

from . import _device as dev
from . import _pred
from ._lib import call

RED_DOT, RED_SUMSQ, RED_SQDIST, RED_SUM, RED_L1, RED_L21 = range(6)
POST_NONE, POST_SQRT = 0, 1
RED_BLOCKS = 1024


def _n(x: torch.Tensor) -> int:
    return x.numel()


def add(x, sign, y, out=None, pred=None):
    """``x + y`` (sign > 0) or ``x - y``."""
    out = torch.empty_like(x) if out is None else out
    call("scxr_vec_add", _n(x), x.data_ptr(), float(sign), y.data_ptr(), out.data_ptr(),
         dev.ptr(_pred.resolve(pred)), dev.stream())
    return out


def axpby(a, x, b, y, out=None, pred=None):
    out = torch.empty_like(x) if out is None else out
    call("scxr_vec_axpby", _n(x), float(a), x.data_ptr(), float(b), y.data_ptr(), out.data_ptr(),
         dev.ptr(_pred.resolve(pred)), dev.stream())
    return out


def scale(a, x, out=None, pred=None):
    out = torch.empty_like(x) if out is None else out
    call("scxr_vec_scale", _n(x), float(a), x.data_ptr(), out.data_ptr(), dev.ptr(_pred.resolve(pred)),
         dev.stream())
    return out


def fill(out, v, pred=None):
    call("scxr_vec_fill", _n(out), float(v), out.data_ptr(), dev.ptr(_pred.resolve(pred)), dev.stream())
    return out


def copy(x, out=None, pred=None):
    out = torch.empty_like(x) if out is None else out
    call("scxr_vec_copy", _n(x), x.data_ptr(), out.data_ptr(), dev.ptr(_pred.resolve(pred)), dev.stream())
    return out


def workspace() -> torch.Tensor:
    return torch.empty(RED_BLOCKS, dtype=torch.float64, device=dev.DEVICE)


def reduce(kind, x, y=None, K=1, scale_=1.0, post=POST_NONE, out=None, pred=None):
    """Device float64 scalar ``post(scale * reduction)``."""
    out = torch.empty((), dtype=torch.float64, device=dev.DEVICE) if out is None else out
    n = _n(x) // K if kind == RED_L21 else _n(x)
    call("scxr_vec_reduce", kind, n, K, x.data_ptr(), dev.ptr(y), float(scale_), post,
         out.data_ptr(), workspace().data_ptr(), dev.ptr(_pred.resolve(pred)), dev.stream())
    return out


def dot(x, y, **kw):
    return reduce(RED_DOT, x, y, **kw)


def norm(x, **kw):
    """Euclidean norm (float64 accumulation)."""
    return reduce(RED_SUMSQ, x, post=POST_SQRT, **kw)


def mul(x, y, out=None, pred=None):
    out = torch.empty_like(x) if out is None else out
    call("scxr_vec_mul", _n(x), x.data_ptr(), y.data_ptr(), out.data_ptr(),
         dev.ptr(_pred.resolve(pred)), dev.stream())
    return out


def ge(x, thr, v=1.0, out=None, pred=None):
    """``(x >= thr) * v``."""
    out = torch.empty_like(x) if out is None else out
    call("scxr_vec_ge", _n(x), x.data_ptr(), float(thr), float(v), out.data_ptr(),
         dev.ptr(_pred.resolve(pred)), dev.stream())
    return out


def div_dev(x, d, out=None, pred=None):
    """``x / d`` with ``d`` a 0-d float64 device tensor."""
    out = torch.empty_like(x) if out is None else out
    call("scxr_vec_div_dev", _n(x), x.data_ptr(), d.data_ptr(), out.data_ptr(),
         dev.ptr(_pred.resolve(pred)), dev.stream())
    return out
