"""Linear operators (``scico.linop`` surface for the hot path)."""

from ._circconv import CircularConvolve
from ._diff import FiniteDifference
from ._linop import ComposedLinearOperator, Identity, LinearOperator, valid_adjoint
from ._stack import VerticalStack
from ._util import operator_norm, power_iteration

__all__ = ["CircularConvolve", "FiniteDifference", "LinearOperator", "ComposedLinearOperator", "Identity", "VerticalStack", "valid_adjoint",
           "power_iteration", "operator_norm"]
