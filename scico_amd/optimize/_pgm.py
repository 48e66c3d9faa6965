"""Proximal gradient methods (``scico/optimize/_pgm.py:30-252``,
``scico/optimize/_pgmaux.py:23-62``)."""

from __future__ import annotations

from typing import List

import numpy as np
import torch

from .. import _device as dev
from .. import _vec as vec
from ._common import Optimizer


class PGMStepSize:
    """Constant reciprocal step size ``L`` (``_pgmaux.py:23-62``)."""

    def internal_init(self, pgm):
        self.pgm = pgm

    def update(self, v):
        return self.pgm.L


class PGM(Optimizer):
    def __init__(self, f, g, L0: float, x0, step_size=None, **kwargs):
        self.f = f
        if g.has_prox is not True:
            raise ValueError(f"Functional 'g' ({type(g)}) must have a prox method.")
        self.g = g
        if step_size is None:
            step_size = PGMStepSize()
        if type(step_size) is not PGMStepSize:
            raise NotImplementedError("Only the constant PGMStepSize is supported.")
        self.step_size = step_size
        self.step_size.internal_init(self)
        self.L = float(L0)
        self.fixed_point_residual = float("inf")
        self.x = dev.as_device(x0).clone()
        super().__init__(**kwargs)

    def x_step(self, v, L):
        """``g.prox(v - (1/L) grad f(v), 1/L)`` (``_pgm.py:106-117``)."""
        u = vec.axpby(1.0, v, -(1.0 / L), self.f.grad(v))
        return self.g.prox(u, 1.0 / L)

    def _working_vars_finite(self) -> bool:
        return bool(torch.isfinite(self.x).all().item())

    def _objective_evaluatable(self):
        return self.f.has_eval and self.g.has_eval

    def _itstat_extra_fields(self):
        return {"L": "%9.3e", "Residual": "%9.3e"}, ["L", "norm_residual()"]

    def _state_variable_names(self) -> List[str]:
        return ["x", "L"]

    def minimizer(self):
        return self.x

    def objective(self, x=None):
        """``f(x) + g(x)`` (``_pgm.py:151-156``), device float64."""
        x = self.x if x is None else x
        return self.f(x) + self.g(x)

    def norm_residual(self):
        return self.fixed_point_residual

    def step(self):
        self.L = self.step_size.update(self.x)
        x = self.x_step(self.x, self.L)
        self.fixed_point_residual = vec.reduce(vec.RED_SQDIST, self.x, x, post=vec.POST_SQRT)
        self.x = x


class AcceleratedPGM(PGM):
    """FISTA (``_pgm.py:183-252``)."""

    def __init__(self, f, g, L0: float, x0, step_size=None, **kwargs):
        super().__init__(f=f, g=g, L0=L0, x0=x0, step_size=step_size, **kwargs)
        self.v = self.x.clone()
        self.t = 1.0

    def step(self):
        x_old = self.x
        self.L = self.step_size.update(self.v)
        self.x = self.x_step(self.v, self.L)
        self.fixed_point_residual = vec.reduce(vec.RED_SQDIST, self.x, self.v,
                                               post=vec.POST_SQRT)
        # t_{k+1} = (1 + sqrt(1 + 4 t_k^2)) / 2 in float32, as the reference
        # evaluates it after the first step (the sequence does not depend on
        # the data, so it is computed on the host).
        t_old = self.t
        f32 = np.float32
        self.t = float(f32(0.5) * (f32(1) + np.sqrt(f32(1) + f32(4) * f32(t_old) ** 2)))
        c = float(f32(f32(t_old) - f32(1)) / f32(self.t))
        self.v = vec.axpby(1.0, self.x, c, vec.add(self.x, -1.0, x_old))

    def _working_vars_finite(self) -> bool:
        return bool(torch.isfinite(self.x).all().item()) and bool(torch.isfinite(self.v).all().item())

    def _state_variable_names(self) -> List[str]:
        return ["x", "v", "t", "L"]
