"""Functionals (``scico.functional`` surface for the hot path)."""

from ._functional import Functional, ScaledFunctional, ZeroFunctional
from ._norm import L1Norm, L21Norm
from ._tvnorm import AnisotropicTVNorm, IsotropicTVNorm, TVNorm

__all__ = ["Functional", "ScaledFunctional", "ZeroFunctional", "L1Norm", "L21Norm",
           "TVNorm", "IsotropicTVNorm", "AnisotropicTVNorm"]
