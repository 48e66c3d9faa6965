"""LinearOperator protocol mirror.

Mirrors the behaviour of ``scico.linop.LinearOperator`` that callers of the
hot path rely on (``scico/operator/_operator.py:63-414``,
``scico/linop/_linop.py:123-497``): shape and dtype attributes, the
``__call__`` shape check (``_operator.py:219-223``), the ``adj`` dtype and
shape checks (``_linop.py:319-325``), ``.T`` / ``.H`` / ``gram_op`` /
``gram``, sums, differences, scalar multiples and composition.

Arrays are device tensors (``torch.Tensor`` on the HIP device, used only as
memory); the arithmetic of sums and scalings runs in the library's vector
kernels.  Host arrays passed in are copied to the device.
"""

from __future__ import annotations

from typing import Callable, Optional

import numpy as np

from .. import _device as dev
from .. import _vec as vec


def _shape_tuple(s) -> tuple:
    """A shape, or a tuple of block shapes (``BlockArray`` operands)."""
    if isinstance(s, (int, np.integer)):
        return (int(s),)
    return tuple(_shape_tuple(v) if isinstance(v, (tuple, list)) else int(v) for v in s)


def is_scalar(s) -> bool:
    return isinstance(s, (int, float, np.integer, np.floating)) and not isinstance(s, bool)


class LinearOperator:
    """Linear operator with an explicit adjoint (``_linop.py:123-185``)."""

    def __init__(
        self,
        input_shape,
        output_shape=None,
        eval_fn: Optional[Callable] = None,
        adj_fn: Optional[Callable] = None,
        input_dtype=np.float32,
        output_dtype=None,
        jit: bool = False,
    ):
        self.input_shape = _shape_tuple(input_shape)
        self.output_shape = _shape_tuple(output_shape if output_shape is not None else input_shape)
        self.input_dtype = np.dtype(input_dtype)
        self.output_dtype = np.dtype(output_dtype if output_dtype is not None else input_dtype)
        if eval_fn is not None:
            self._eval = eval_fn
        if callable(adj_fn):
            self._adj = adj_fn
        elif adj_fn is not None:
            raise TypeError(f"Argument 'adj_fn' must be either a Callable or None; got {adj_fn}.")
        elif not hasattr(self, "_adj"):
            self._adj = None

    # --- attributes (_operator.py:106-167)
    @property
    def shape(self):
        return (self.output_shape, self.input_shape)

    @property
    def input_size(self) -> int:
        return int(np.prod(self.input_shape))

    @property
    def output_size(self) -> int:
        return int(np.prod(self.output_shape))

    @property
    def matrix_shape(self):
        return (self.output_size, self.input_size)

    def __repr__(self):
        return (
            f"{type(self).__name__}\n  input_shape:  {self.input_shape}\n"
            f"  output_shape: {self.output_shape}\n  input_dtype:  {self.input_dtype}\n"
            f"  output_dtype: {self.output_dtype}\n"
        )

    # --- evaluation
    def _eval(self, x):
        raise NotImplementedError(f"{type(self).__name__} does not define _eval")

    def __call__(self, x):
        if isinstance(x, LinearOperator):
            return ComposedLinearOperator(self, x)
        if self.input_shape != dev.array_shape(x):
            raise ValueError(
                f"Cannot evaluate {type(self)} with input_shape={self.input_shape} "
                f"on array with shape={dev.array_shape(x)}."
            )
        return self._eval(dev.as_device(x, self.input_dtype))

    def __matmul__(self, other):
        return self(other)

    def adj(self, y):
        """Adjoint applied to ``y`` (``_linop.py:296-325``)."""
        if isinstance(y, LinearOperator):
            return ComposedLinearOperator(self.H, y)
        if self._adj is None:
            raise NotImplementedError(f"{type(self).__name__} has no adjoint")
        if self.output_dtype != dev.array_dtype(y):
            raise ValueError(f"Dtype error: expected {self.output_dtype}, got {dev.array_dtype(y)}.")
        if self.output_shape != dev.array_shape(y):
            raise ValueError(
                f"""Shapes do not conform: input array with shape {dev.array_shape(y)} does not match
                LinearOperator output_shape {self.output_shape}."""
            )
        return self._adj(dev.as_device(y, self.output_dtype))

    @property
    def T(self) -> "LinearOperator":
        """Transpose (= adjoint for real operators, ``_linop.py:327-360``)."""
        return LinearOperator(
            input_shape=self.output_shape,
            output_shape=self.input_shape,
            eval_fn=self.adj,
            adj_fn=self.__call__,
            input_dtype=self.output_dtype,
            output_dtype=self.input_dtype,
        )

    @property
    def H(self) -> "LinearOperator":
        return self.T

    def conj(self) -> "LinearOperator":
        return self

    @property
    def gram_op(self) -> "LinearOperator":
        """``A^H A`` as an operator (``_linop.py:404-440``)."""
        return LinearOperator(
            input_shape=self.input_shape,
            output_shape=self.input_shape,
            eval_fn=self.gram,
            adj_fn=self.gram,
            input_dtype=self.input_dtype,
            output_dtype=self.input_dtype,
        )

    def gram(self, x):
        return self.adj(self(x))

    # --- algebra (_linop.py:212-258)
    def _check_conform(self, other):
        if self.shape != other.shape:
            raise ValueError(f"Shapes {self.shape} and {other.shape} do not conform.")

    def __add__(self, other):
        if not isinstance(other, LinearOperator):
            return NotImplemented
        self._check_conform(other)
        return _SumOperator(self, other, +1.0)

    def __sub__(self, other):
        if not isinstance(other, LinearOperator):
            return NotImplemented
        self._check_conform(other)
        return _SumOperator(self, other, -1.0)

    def __mul__(self, scalar):
        if not is_scalar(scalar):
            return NotImplemented
        return _ScaledOperator(self, float(scalar))

    __rmul__ = __mul__

    def __truediv__(self, scalar):
        if not is_scalar(scalar):
            return NotImplemented
        return _ScaledOperator(self, 1.0 / float(scalar))

    def __neg__(self):
        return _ScaledOperator(self, -1.0)


class _SumOperator(LinearOperator):
    """``A + B`` / ``A - B``: both operands evaluated, then added."""

    def __init__(self, A: LinearOperator, B: LinearOperator, sign: float):
        self.A, self.B, self.sign = A, B, sign
        super().__init__(
            input_shape=A.input_shape,
            output_shape=A.output_shape,
            input_dtype=A.input_dtype,
            output_dtype=A.output_dtype,
        )

    def _eval(self, x):
        return vec.add(self.A(x), self.sign, self.B(x))

    def _adj(self, y):
        return vec.add(self.A.adj(y), self.sign, self.B.adj(y))


class _ScaledOperator(LinearOperator):
    """``c * A``."""

    def __init__(self, A: LinearOperator, scalar: float):
        self.A, self.scalar = A, scalar
        super().__init__(
            input_shape=A.input_shape,
            output_shape=A.output_shape,
            input_dtype=A.input_dtype,
            output_dtype=A.output_dtype,
        )

    def _eval(self, x):
        return vec.scale(self.scalar, self.A(x))

    def _adj(self, y):
        return vec.scale(self.scalar, self.A.adj(y))


class ComposedLinearOperator(LinearOperator):
    """``A @ B`` (``_linop.py:443-497``)."""

    def __init__(self, A: LinearOperator, B: LinearOperator):
        if A.input_shape != B.output_shape:
            raise ValueError(f"Incompatible shapes {A.shape}, {B.shape}.")
        self.A, self.B = A, B
        super().__init__(
            input_shape=B.input_shape,
            output_shape=A.output_shape,
            input_dtype=B.input_dtype,
            output_dtype=A.output_dtype,
        )

    def _eval(self, x):
        return self.A(self.B(x))

    def _adj(self, y):
        return self.B.adj(self.A.adj(y))


class Identity(LinearOperator):
    """Identity (``scico/linop/_diag.py:268-290``); ``diagonal`` is 1.0."""

    def __init__(self, input_shape, input_dtype=np.float32):
        super().__init__(input_shape=input_shape, input_dtype=input_dtype)
        self.diagonal = 1.0

    def _eval(self, x):
        return x

    def _adj(self, y):
        return y


def valid_adjoint(A: LinearOperator, AT: Optional[LinearOperator] = None, eps: Optional[float] = 1e-7,
                  x=None, y=None, key=None):
    """Dot test ``|y^T A x - (A^T y)^T x| / max(...)`` (``scico/linop/_util.py:113-183``).

    Random test vectors are standard normal float32 drawn with NumPy
    (``key`` is an integer seed); returns a bool when ``eps`` is given,
    else the relative error.
    """
    if AT is None:
        AT = A.T
    rng = np.random.default_rng(0 if key is None else key)
    if x is None:
        x = rng.standard_normal(A.input_shape).astype(np.float32)
    if y is None:
        y = rng.standard_normal(AT.input_shape).astype(np.float32)
    Ax = dev.to_numpy(A(x)).astype(np.float64)
    ATy = dev.to_numpy(AT(y)).astype(np.float64)
    u = float(np.sum(dev.to_numpy(y).astype(np.float64) * Ax))
    v = float(np.sum(ATy * dev.to_numpy(x).astype(np.float64)))
    err = abs(u - v) / max(abs(u), abs(v))
    if eps is None:
        return err
    return err < eps
