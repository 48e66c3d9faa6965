"""DFT-domain ADMM x-step (``scico/optimize/_admmauxconv.py:44-151``).

When ``f`` is ``None`` or a :class:`SquaredL2Loss` whose operator is a
:class:`CircularConvolve` or :class:`Identity`, and every ``C_i`` is shift
invariant (``Identity``, ``CircularConvolve``, ``FiniteDifference`` with
``circular=True``), the x-step operator

    A_lhs = sum_i rho_i C_i^T C_i + 2 scale A^T A

is diagonal in the DFT domain.  Its spectrum is built once from impulse
responses (``CircularConvolve.from_operator``, ``_circconv.py:258-307``)
and the x-step becomes ``x = real(ifftn(fftn(rhs) / A_lhs_dft))``: one
pass of the FFT convolution kernels with the reciprocal spectrum.
"""

from __future__ import annotations

from functools import reduce
from typing import Optional, Type

import numpy as np

from .. import _device as dev
from ..linop import CircularConvolve, Identity
from ..loss import SquaredL2Loss
from ._admmaux import LinearSubproblemSolver, SubproblemSolver


class CircularConvolveSolver(LinearSubproblemSolver):
    """x-step for operators diagonalized by the DFT (``_admmauxconv.py:44-151``).

    Attributes:
        admm: ADMM object the solver is attached to.
        A_lhs: :class:`CircularConvolve` whose spectrum is the DFT of the
            x-step operator.
    """

    def __init__(self, ndims: Optional[int] = None,
                 conv_class: Type[CircularConvolve] = CircularConvolve, device=None):
        self.ndims = ndims
        self.conv_class = conv_class
        self.device = device
        self.info = None

    def internal_init(self, admm):
        if admm.f is None:
            is_cc = [isinstance(C, self.conv_class) for C in admm.C_list]
            auto_ndims = admm.C_list[is_cc.index(True)].ndims if any(is_cc) else None
        else:
            if not isinstance(admm.f, SquaredL2Loss):
                raise TypeError(
                    "CircularConvolveSolver requires f to be a scico.loss.SquaredL2Loss; "
                    f"got {type(admm.f)}."
                )
            if not isinstance(admm.f.A, (self.conv_class, Identity)):
                raise TypeError(
                    "CircularConvoleSolver requires f.A to be a scico.linop.self.conv_class "
                    f"or scico.linop.Identity; got {type(admm.f.A)}."
                )
            auto_ndims = admm.f.A.ndims if isinstance(admm.f.A, self.conv_class) else None
        if self.ndims is None:
            self.ndims = auto_ndims
        SubproblemSolver.internal_init(self, admm)
        self.real_result = np.dtype(admm.C_list[0].input_dtype).kind == "f"

        # The C_i are assumed linear and shift invariant (not checked), as in
        # the reference.  Each term's spectrum comes from its impulse response.
        def spectrum(op):
            cc = self.conv_class.from_operator(op, ndims=self.ndims)
            return dev.to_numpy(cc.h_dft).astype(np.complex64)

        terms = [np.complex64(rho) * spectrum(C.gram_op)
                 for rho, C in zip(admm.rho_list, admm.C_list)]
        lhs = reduce(lambda a, b: a + b, terms)
        if admm.f is not None:
            lhs = lhs + np.complex64(2.0 * admm.f.scale) * spectrum(admm.f.A.gram_op)
        self.lhs_dft = lhs.astype(np.complex64)
        shape = admm.C_list[0].input_shape
        nd = self.ndims if self.ndims is not None else len(shape)
        self.A_lhs = self.conv_class(self.lhs_dft, shape, ndims=nd, h_is_dft=True)
        # The solve multiplies by the reciprocal spectrum (complex64).
        inv = (np.complex64(1.0) / self.lhs_dft).astype(np.complex64)
        self._solve_op = self.conv_class(inv, shape, ndims=nd, h_is_dft=True)

    def solve(self, x0):
        """``real(ifftn(fftn(rhs) / A_lhs_dft))``; ``x0`` is unused."""
        rhs = self.compute_rhs()
        return self._solve_op(rhs)
