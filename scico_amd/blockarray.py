"""Block arrays (``scico/numpy/_blockarray.py``): an ordered list of device
arrays of possibly different shapes, used where an operator or a split
variable has several components (``VerticalStack`` outputs, the ``z`` and
``u`` variables of proximal ADMM).

Arithmetic runs block by block in the library's vector kernels.
"""

from __future__ import annotations

import torch

from . import _vec as vec


class BlockArray(list):
    """List of device tensors with blockwise arithmetic."""

    @property
    def shape(self) -> tuple:
        return tuple(tuple(b.shape) for b in self)

    @property
    def dtype(self):
        return self[0].dtype

    @property
    def device(self):
        return self[0].device

    def numel(self) -> int:
        return sum(int(b.numel()) for b in self)

    def clone(self) -> "BlockArray":
        return BlockArray(b.clone() for b in self)

    def __add__(self, other):
        if isinstance(other, BlockArray):
            return BlockArray(vec.add(a, 1.0, b) for a, b in zip(self, other))
        return NotImplemented

    def __sub__(self, other):
        if isinstance(other, BlockArray):
            return BlockArray(vec.add(a, -1.0, b) for a, b in zip(self, other))
        return NotImplemented

    def __neg__(self):
        return BlockArray(vec.scale(-1.0, a) for a in self)

    def __mul__(self, s):
        if isinstance(s, (int, float)):
            return BlockArray(vec.scale(float(s), a) for a in self)
        return NotImplemented

    __rmul__ = __mul__

    # a list compares by identity in `in` checks of callers; keep list equality
    __hash__ = None


def blockarray(arrays) -> BlockArray:
    from . import _device as dev

    return BlockArray(dev.as_device(a) for a in arrays)


# ---- helpers over a tensor or a BlockArray


def axpby(a: float, x, b: float, y):
    """``a x + b y``."""
    if isinstance(x, BlockArray):
        return BlockArray(vec.axpby(a, xi, b, yi) for xi, yi in zip(x, y))
    return vec.axpby(a, x, b, y)


def add(x, sign: float, y):
    """``x + sign y``."""
    if isinstance(x, BlockArray):
        return BlockArray(vec.add(xi, sign, yi) for xi, yi in zip(x, y))
    return vec.add(x, sign, y)


def scale(a: float, x):
    if isinstance(x, BlockArray):
        return BlockArray(vec.scale(a, xi) for xi in x)
    return vec.scale(a, x)


def sumsq(x) -> torch.Tensor:
    """Device float64 sum of squares over all blocks."""
    if isinstance(x, BlockArray):
        out = None
        for xi in x:
            s = vec.reduce(vec.RED_SUMSQ, xi)
            out = s if out is None else out + s
        return out
    return vec.reduce(vec.RED_SUMSQ, x)


def norm(x) -> torch.Tensor:
    """Euclidean norm over all blocks (device float64)."""
    return torch.sqrt(sumsq(x))


def zeros_like(x):
    if isinstance(x, BlockArray):
        return BlockArray(torch.zeros_like(xi) for xi in x)
    return torch.zeros_like(x)


def zeros(shape, device):
    """Zeros of a shape or a tuple of block shapes."""
    if len(shape) and isinstance(shape[0], tuple):
        return BlockArray(torch.zeros(s, dtype=torch.float32, device=device) for s in shape)
    return torch.zeros(shape, dtype=torch.float32, device=device)


def all_finite(x) -> bool:
    blocks = x if isinstance(x, BlockArray) else [x]
    return all(bool(torch.isfinite(b).all().item()) for b in blocks)
