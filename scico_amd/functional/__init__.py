"""Functionals (``scico.functional`` surface for the hot path)."""

from ._functional import Functional, ScaledFunctional, ZeroFunctional
from ._norm import L1Norm, L21Norm

__all__ = ["Functional", "ScaledFunctional", "ZeroFunctional", "L1Norm", "L21Norm"]
