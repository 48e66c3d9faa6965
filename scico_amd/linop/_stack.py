"""Vertical stack of linear operators (``scico/linop/_stack.py:25-72``,
``scico/operator/_stack.py:26-140``).

``H x = (A_1 x, ..., A_N x)``: a :class:`BlockArray` of the outputs, or,
when every output has the same shape and ``collapse_output`` is set, one
array of shape ``(N, *shape)``.  The adjoint sums ``A_i^T y_i`` in block
order.
"""

from __future__ import annotations

from typing import Optional, Sequence

from .. import _device as dev
from .. import _vec as vec
from ..blockarray import BlockArray
from ._linop import LinearOperator


class VerticalStack(LinearOperator):
    def __init__(self, ops: Sequence[LinearOperator], collapse_output: Optional[bool] = True,
                 jit: bool = True, **kwargs):
        if not all(isinstance(op, LinearOperator) for op in ops):
            raise TypeError("All elements of 'ops' must be of type LinearOperator.")
        ops = list(ops)
        if any(op.input_shape != ops[0].input_shape for op in ops):
            raise ValueError(
                "Operators must have the same input_shape; got "
                f"{[op.input_shape for op in ops]}."
            )
        shapes = [op.output_shape for op in ops]
        if any(isinstance(s[0], tuple) for s in shapes if len(s)):
            raise ValueError(
                "Combining these shapes would result in a twice-nested BlockArray, which is "
                "not supported."
            )
        self.ops = ops
        self.collapse_output = bool(collapse_output) and all(s == shapes[0] for s in shapes)
        out_shape = (len(ops), *shapes[0]) if self.collapse_output else tuple(shapes)
        super().__init__(input_shape=ops[0].input_shape, output_shape=out_shape,
                         input_dtype=ops[0].input_dtype, output_dtype=ops[0].output_dtype)

    def _eval(self, x):
        outs = [op(x) for op in self.ops]
        if not self.collapse_output:
            return BlockArray(outs)
        full = dev.empty(self.output_shape)
        for i, o in enumerate(outs):
            vec.copy(o, out=full[i])
        return full

    def _adj(self, y):
        acc = None
        for i, op in enumerate(self.ops):
            t = op.adj(y[i])
            acc = t if acc is None else vec.add(acc, 1.0, t)
        return acc
