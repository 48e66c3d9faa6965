"""Functional base classes (``scico/functional/_functional.py:22-378``).

Values are returned as 0-d float64 device tensors (no host sync).
"""

from __future__ import annotations

import numpy as np


class Functional:
    """Base functional: ``__call__``, ``prox(v, lam)``, ``grad``."""

    has_eval = False
    has_prox = False

    def __call__(self, x):
        raise NotImplementedError(f"{type(self).__name__} cannot be evaluated")

    def prox(self, v, lam=1.0, **kwargs):
        raise NotImplementedError(f"{type(self).__name__} has no prox")

    def grad(self, x):
        raise NotImplementedError(f"{type(self).__name__} has no gradient")

    def __mul__(self, other):
        if np.isscalar(other) or (hasattr(other, "shape") and np.asarray(other).size == 1):
            return ScaledFunctional(self, float(other))
        return NotImplemented

    __rmul__ = __mul__


class ScaledFunctional(Functional):
    """``scale * functional`` (``_functional.py:141-198``)."""

    def __init__(self, functional: Functional, scale: float):
        self.functional = functional
        self.scale = float(scale)
        self.has_eval = functional.has_eval
        self.has_prox = functional.has_prox

    def __call__(self, x):
        return self.scale * self.functional(x)

    def __mul__(self, other):
        if np.isscalar(other):
            return ScaledFunctional(self.functional, float(other) * self.scale)
        return NotImplemented

    __rmul__ = __mul__

    def prox(self, v, lam=1.0, **kwargs):
        # prox_{lam * (scale f)} = prox_{(lam * scale) f}   (_functional.py:198)
        return self.functional.prox(v, lam * self.scale, **kwargs)


class ZeroFunctional(Functional):
    """The zero functional (prox = identity)."""

    has_eval = True
    has_prox = True

    def __call__(self, x):
        import torch

        return torch.zeros((), dtype=torch.float64, device=x.device)

    def prox(self, v, lam=1.0, **kwargs):
        return v


class SeparableFunctional(Functional):
    """``f(x) = sum_i f_i(x_i)`` over the blocks of a :class:`BlockArray`
    (``scico/functional/_functional.py:201-270``); the prox is blockwise."""

    def __init__(self, functional_list):
        self.functional_list = list(functional_list)
        self.has_eval = all(fi.has_eval for fi in self.functional_list)
        self.has_prox = all(fi.has_prox for fi in self.functional_list)

    def __call__(self, x):
        if len(x) != len(self.functional_list):
            raise ValueError(
                f"Number of blocks in x, {len(x)}, and length of functional_list, "
                f"{len(self.functional_list)}, do not match."
            )
        out = None
        for fi, xi in zip(self.functional_list, x):
            t = fi(xi)
            out = t if out is None else out + t
        return out

    def prox(self, v, lam=1.0, **kwargs):
        from ..blockarray import BlockArray

        if len(v) != len(self.functional_list):
            raise ValueError(
                f"Number of blocks in v, {len(v)}, and length of functional_list, "
                f"{len(self.functional_list)}, do not match."
            )
        return BlockArray(fi.prox(vi, lam, **kwargs) for fi, vi in zip(self.functional_list, v))
