"""Functionals (``scico.functional`` surface for the hot path)."""

from ._functional import Functional, ScaledFunctional, SeparableFunctional, ZeroFunctional
from ._norm import L1Norm, L21Norm, SquaredL2Norm
from ._tvnorm import AnisotropicTVNorm, IsotropicTVNorm, TVNorm

__all__ = ["Functional", "ScaledFunctional", "SeparableFunctional", "ZeroFunctional",
           "L1Norm", "L21Norm", "SquaredL2Norm",
           "TVNorm", "IsotropicTVNorm", "AnisotropicTVNorm"]
