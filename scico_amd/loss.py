"""Loss functionals (``scico/loss.py:40-267``): ``SquaredL2Loss``."""

from __future__ import annotations

import numpy as np

from . import _device as dev
from . import _vec as vec
from .functional import Functional
from .linop import Identity, LinearOperator


class Loss(Functional):
    """``scale * loss(y, A(x))`` (``loss.py:40-150``)."""

    def __init__(self, y, A=None, scale: float = 0.5):
        self.y = dev.as_device(y)
        self.A = Identity(tuple(self.y.shape)) if A is None else A
        self.scale = float(scale)

    def set_scale(self, new_scale: float):
        self.scale = float(new_scale)


class SquaredL2Loss(Loss):
    r"""``scale * ||y - A(x)||_2^2`` with identity weighting (``loss.py:154-267``)."""

    has_eval = True

    def __init__(self, y, A=None, scale: float = 0.5, W=None, prox_kwargs=None):
        if W is not None and not isinstance(W, Identity):
            raise NotImplementedError("SquaredL2Loss supports W=None (identity weighting).")
        super().__init__(y=y, A=A, scale=scale)
        self.W = Identity(tuple(self.y.shape))
        p = {"maxiter": 100, "tol": 1e-5}
        if prox_kwargs:
            p.update(prox_kwargs)
        self.prox_kwargs = p
        self.has_prox = isinstance(self.A, LinearOperator)

    def __call__(self, x):
        """``scale * sum |y - A x|^2`` (``loss.py:208-209``), device float64."""
        return vec.reduce(vec.RED_SQDIST, self.y, self.A(x), scale_=self.scale)

    def grad(self, x):
        """``2 scale A^T (A x - y)`` (the autodiff gradient of ``__call__``).

        For a :class:`CircularConvolve` the gradient is evaluated as
        ``2 scale (G x - A^T y)`` with ``G = A^T A`` and ``A^T y`` cached: one
        FFT convolution per gradient instead of two (equal up to float32
        rounding).  ``A x = real(ifftn(H fftn(x)))`` is the convolution by
        the Hermitian part ``Hs[k] = (H[k] + conj(H[-k])) / 2`` of the
        spectrum, so ``G`` is the convolution by ``|Hs|^2`` for any ``H``
        (``Hs = H`` for a real kernel).
        """
        from .linop import CircularConvolve

        c = 2.0 * self.scale
        if type(self.A) is CircularConvolve:
            G, ATy = self._conv_gram()
            g = vec.add(G(x), -1.0, ATy)
        else:
            r = vec.add(self.A(x), -1.0, self.y)
            g = self.A.adj(r)
        return g if c == 1.0 else vec.scale(c, g)

    def _conv_gram(self):
        """Cached (G, A^T y) for a circular-convolution operator; rebuilt
        when the operator or y changes (the cache holds the operator itself,
        so an identity test cannot match a different, later operator)."""
        from .linop import CircularConvolve

        A = self.A
        key = (self.y.data_ptr(), self.y._version)
        cache = getattr(self, "_gram_cache", None)
        if cache is None or cache[0] is not A or cache[1] != key:
            hd = dev.to_numpy(A.h_dft).astype(np.complex128)
            axes = tuple(range(hd.ndim))
            hneg = np.roll(np.flip(hd, axes), 1, axes)  # H[-k]
            hs = 0.5 * (hd + np.conj(hneg))
            g = hs.real**2 + hs.imag**2
            G = CircularConvolve(g.astype(np.complex64), A.input_shape, ndims=A.ndims,
                                 h_is_dft=True)
            cache = (A, key, G, A.adj(self.y))
            self._gram_cache = cache
        return cache[2], cache[3]

    @property
    def hessian(self) -> LinearOperator:
        """``2 scale A^T W A`` (``loss.py:246-262``)."""
        A = self.A
        c = 2.0 * self.scale

        def h(x):
            v = A.adj(A(x))
            return v if c == 1.0 else vec.scale(c, v)

        return LinearOperator(input_shape=A.input_shape, output_shape=A.input_shape,
                              eval_fn=h, adj_fn=h, input_dtype=A.input_dtype)

    def prox(self, v, lam=1.0, **kwargs):
        """``(I + lam 2 scale A^T A)^{-1} (v + lam 2 scale A^T y)``: closed form for
        the identity, CG otherwise (``loss.py:211-243``)."""
        from .solver import cg

        if not isinstance(self.A, LinearOperator):
            raise NotImplementedError("prox needs A to be a LinearOperator")
        v = dev.as_device(v)
        if type(self.A) is Identity:
            # diagonal A (``loss.py:220-226``): (c y + v) / (c + 1), c = 2 scale lam
            c = 2.0 * self.scale * float(lam)
            return vec.scale(1.0 / (c + 1.0), vec.axpby(c, self.y, 1.0, v))
        H = self.hessian
        c = float(lam)
        lhs = LinearOperator(input_shape=v.shape, output_shape=v.shape,
                             eval_fn=lambda x: vec.axpby(1.0, x, c, H(x)),
                             adj_fn=lambda x: vec.axpby(1.0, x, c, H(x)))
        rhs = vec.axpby(1.0, v, 2.0 * c * self.scale, self.A.adj(self.y))
        x0 = kwargs.get("x0")
        x0 = dev.zeros(v.shape) if x0 is None else x0
        x, _ = cg(lhs, rhs, x0, **self.prox_kwargs)
        return x
