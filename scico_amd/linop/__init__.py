"""Linear operators (``scico.linop`` surface for the hot path)."""

from ._linop import ComposedLinearOperator, Identity, LinearOperator, valid_adjoint

__all__ = ["LinearOperator", "ComposedLinearOperator", "Identity", "valid_adjoint"]
