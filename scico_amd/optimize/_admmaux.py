"""ADMM x-step solvers (``scico/optimize/_admmaux.py:37-269``)."""

from __future__ import annotations

from functools import reduce
from typing import Any, Optional

from .. import _vec as vec
from ..linop import LinearOperator
from ..loss import SquaredL2Loss
from ..solver import cg as scico_cg


class SubproblemSolver:
    """Base class of ADMM x-step solvers (``_admmaux.py:37-66``)."""

    def internal_init(self, admm):
        self.admm = admm


class GenericSubproblemSolver(SubproblemSolver):
    """Placeholder for the scipy-minimize x-step (not on the hot path)."""

    def internal_init(self, admm):
        raise NotImplementedError(
            "GenericSubproblemSolver is not supported; use LinearSubproblemSolver."
        )


class LinearSubproblemSolver(SubproblemSolver):
    r"""x-step for a quadratic ``f``: CG on
    ``(A^T W A + sum_i rho_i C_i^T C_i) x = A^T W y + sum_i rho_i C_i^T (z_i - u_i)``.
    """

    def __init__(self, cg_kwargs: Optional[dict[str, Any]] = None, cg_function: str = "scico",
                 device=None):
        default = {"tol": 1e-4, "maxiter": 100}
        if cg_kwargs:
            default.update(cg_kwargs)
        self.cg_kwargs = default
        if cg_function not in ("scico", "jax"):
            raise ValueError(
                f"Argument 'cg_function' must be one of 'jax', 'scico'; got {cg_function}."
            )
        self.cg_function = cg_function
        self.cg = scico_cg
        self.info = None
        self.device = device

    def internal_init(self, admm):
        if admm.f is not None:
            if not isinstance(admm.f, SquaredL2Loss):
                raise TypeError(
                    "LinearSubproblemSolver requires f to be a scico.loss.SquaredL2Loss; "
                    f"got {type(admm.f)}."
                )
            if not isinstance(admm.f.A, LinearOperator):
                raise TypeError(
                    "LinearSubproblemSolver requires f.A to be a scico.linop.LinearOperator; "
                    f"got {type(admm.f.A)}."
                )
        super().internal_init(admm)
        lhs = reduce(lambda a, b: a + b,
                     [rho * C.gram_op for rho, C in zip(admm.rho_list, admm.C_list)])
        if admm.f is not None:
            lhs = lhs + admm.f.hessian
        self.lhs_op = lhs

    def compute_rhs(self):
        """``A^T W y * 2 scale + sum_i rho_i C_i^T (z_i - u_i)`` (``_admmaux.py:231-256``)."""
        admm = self.admm
        rhs = None
        if admm.f is not None:
            ATy = admm.f.A.adj(admm.f.y)
            rhs = vec.scale(2.0 * admm.f.scale, ATy)
        for rho, C, z, u in zip(admm.rho_list, admm.C_list, admm.z_list, admm.u_list):
            t = vec.scale(rho, C.adj(vec.add(z, -1.0, u)))
            rhs = t if rhs is None else vec.add(rhs, 1.0, t)
        return rhs

    def solve(self, x0):
        rhs = self.compute_rhs()
        x, self.info = self.cg(self.lhs_op, rhs, x0, **self.cg_kwargs)
        return x
