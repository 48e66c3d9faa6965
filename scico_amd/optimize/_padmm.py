"""Proximal ADMM (``scico/optimize/_padmm.py:30-412``).

Solves ``argmin_x f(x) + g(z)  s.t.  A x + B z = c`` with the iterations
(``_padmm.py:330-350``)

    x <- prox_{f / (rho mu)}(x - (1/mu) A^T (2 u - u_old))
    z <- prox_{g / (rho nu)}(z - (1/nu) B^T (A x + B z - c + u))
    u <- u + A x + B z - c

``B = None`` means ``B = -I``.  Variables may be :class:`BlockArray` (for
example ``A = VerticalStack((C, alpha D))`` with a separable ``g``), so no
x-step solve is needed: each iteration is one ``A`` and one ``A^T``
application plus blockwise vector passes.
"""

from __future__ import annotations

from typing import List, Optional, Tuple

import torch

from .. import _device as dev
from .. import blockarray as blk
from ..linop import Identity, LinearOperator, operator_norm
from ._common import Optimizer


class ProximalADMMBase(Optimizer):
    """Shared state and statistics of the proximal ADMM solvers
    (``_padmm.py:30-168``)."""

    def __init__(self, f, g, rho: float, mu: float, nu: float, xshape, zshape, ushape,
                 x0=None, z0=None, u0=None, fast_dual_residual: bool = True, **kwargs):
        self.f = f
        self.g = g
        self.rho = float(rho)
        self.mu = float(mu)
        self.nu = float(nu)
        self.fast_dual_residual = bool(fast_dual_residual)
        self.x = blk.zeros(xshape, dev.DEVICE) if x0 is None else dev.as_device(x0)
        self.z = blk.zeros(zshape, dev.DEVICE) if z0 is None else dev.as_device(z0)
        self.z_old = self.z
        self.u = blk.zeros(ushape, dev.DEVICE) if u0 is None else dev.as_device(u0)
        self.u_old = self.u
        super().__init__(**kwargs)

    def _working_vars_finite(self) -> bool:
        return blk.all_finite(self.x) and blk.all_finite(self.z) and blk.all_finite(self.u)

    def _objective_evaluatable(self):
        return self.f.has_eval and self.g.has_eval

    def _itstat_extra_fields(self):
        return ({"Prml Rsdl": "%9.3e", "Dual Rsdl": "%9.3e"},
                ["norm_primal_residual()", "norm_dual_residual()"])

    def _state_variable_names(self) -> List[str]:
        return ["x", "z", "z_old", "u", "u_old"]

    def minimizer(self):
        return self.x

    def objective(self, x=None, z=None):
        """``f(x) + g(z)`` (``_padmm.py:136-168``)."""
        if (x is None) != (z is None):
            raise ValueError("Both or neither of arguments 'x' and 'z' must be supplied")
        if x is None:
            x, z = self.x, self.z
        return self.f(x) + self.g(z)


class ProximalADMM(ProximalADMMBase):
    """Proximal ADMM with linear constraint ``A x + B z = c``
    (``_padmm.py:171-412``)."""

    def __init__(self, f, g, A: LinearOperator, rho: float, mu: float, nu: float,
                 B: Optional[LinearOperator] = None, c=None, x0=None, z0=None, u0=None,
                 fast_dual_residual: bool = True, **kwargs):
        self.A = A
        self.B = B  # None: B = -I, applied as a negation
        self.c = c
        zshape = A.output_shape if B is None else B.input_shape
        super().__init__(f, g, rho, mu, nu, A.input_shape, zshape, A.output_shape, x0=x0,
                         z0=z0, u0=u0, fast_dual_residual=fast_dual_residual, **kwargs)

    def _Bz(self, z):
        return blk.scale(-1.0, z) if self.B is None else self.B(z)

    def _BH(self, w):
        return blk.scale(-1.0, w) if self.B is None else self.B.adj(w)

    def _constraint(self, Ax, z):
        """``A x + B z - c``."""
        if self.B is None:
            r = blk.add(Ax, -1.0, z)
        else:
            r = blk.add(Ax, 1.0, self.B(z))
        if self.c is not None:
            r = blk.add(r, -1.0, dev.as_device(self.c))
        return r

    def norm_primal_residual(self, x=None, z=None):
        """``||A x + B z - c||`` (``_padmm.py:288-317``)."""
        if (x is None) != (z is None):
            raise ValueError("Both or neither of arguments 'x' and 'z' must be supplied")
        if x is None:
            x, z = self.x, self.z
        return blk.norm(self._constraint(self.A(x), z))

    def norm_dual_residual(self):
        """``||z - z_old||`` (fast) or ``||A^T B (z - z_old)||`` (``_padmm.py:319-346``)."""
        d = blk.add(self.z, -1.0, self.z_old)
        if not self.fast_dual_residual:
            d = self.A.adj(self._Bz(d))
        return blk.norm(d)

    def step(self):
        """One iteration (``_padmm.py:348-361``)."""
        t = blk.axpby(2.0, self.u, -1.0, self.u_old)
        proxarg = blk.axpby(1.0, self.x, -1.0 / self.mu, self.A.adj(t))
        self.x = self.f.prox(proxarg, 1.0 / (self.rho * self.mu), v0=self.x)
        Ax = self.A(self.x)
        w = blk.add(self._constraint(Ax, self.z), 1.0, self.u)
        proxarg = blk.axpby(1.0, self.z, -1.0 / self.nu, self._BH(w))
        self.z_old = self.z
        self.z = self.g.prox(proxarg, 1.0 / (self.rho * self.nu), v0=self.z)
        self.u_old = self.u
        self.u = blk.add(self.u, 1.0, self._constraint(Ax, self.z))

    @staticmethod
    def estimate_parameters(A: LinearOperator, B: Optional[LinearOperator] = None,
                            factor: Optional[float] = 1.01, maxiter: int = 100, key=None,
                            xdevice=None, zdevice=None) -> Tuple[float, float]:
        """``mu = ||A||^2``, ``nu = ||B||^2`` (1 for ``B = -I``), times
        ``factor`` (``_padmm.py:363-412``)."""
        mu = float(operator_norm(A, maxiter=maxiter, key=key)) ** 2
        nu = 1.0 if B is None else float(operator_norm(B, maxiter=maxiter, key=key)) ** 2
        if factor is None:
            return (mu, nu)
        return (factor * mu, factor * nu)
