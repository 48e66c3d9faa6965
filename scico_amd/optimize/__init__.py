"""Optimizers (``scico.optimize`` surface for the hot path)."""

from ._admm import ADMM
from ._admmaux import GenericSubproblemSolver, LinearSubproblemSolver, SubproblemSolver
from ._common import Optimizer
from ._pgm import PGM, AcceleratedPGM, PGMStepSize

__all__ = ["ADMM", "LinearSubproblemSolver", "GenericSubproblemSolver", "SubproblemSolver",
           "Optimizer", "PGM", "AcceleratedPGM", "PGMStepSize"]
