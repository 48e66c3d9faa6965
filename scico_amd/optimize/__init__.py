"""Optimizers (``scico.optimize`` surface for the hot path)."""

from ._admm import ADMM
from ._admmaux import GenericSubproblemSolver, LinearSubproblemSolver, SubproblemSolver
from ._admmauxconv import CircularConvolveSolver
from ._common import Optimizer
from ._padmm import ProximalADMM, ProximalADMMBase
from ._pgm import PGM, AcceleratedPGM, PGMStepSize

__all__ = ["ADMM", "LinearSubproblemSolver", "CircularConvolveSolver", "GenericSubproblemSolver", "SubproblemSolver",
           "Optimizer", "ProximalADMM", "ProximalADMMBase", "PGM", "AcceleratedPGM", "PGMStepSize"]
