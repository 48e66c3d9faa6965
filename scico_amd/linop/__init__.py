"""Linear operators (``scico.linop`` surface for the hot path)."""

from ._diff import FiniteDifference
from ._linop import ComposedLinearOperator, Identity, LinearOperator, valid_adjoint

__all__ = ["FiniteDifference", "LinearOperator", "ComposedLinearOperator", "Identity", "valid_adjoint"]
