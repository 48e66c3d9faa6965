"""Conjugate gradient (``scico/solver.py:326-405``) with device scalars.

Iterations are enqueued with a device-side stopping test (each kernel of
iteration k is predicated on ``flags[k]``), so no host synchronisation is
needed; for long runs the host checks the flag every ``check_every``
iterations to stop enqueueing no-op work.
"""

from __future__ import annotations

import torch

from . import _device as dev
from . import _vec as vec
from ._lib import call, load
from ._pred import predicated
from .linop import LinearOperator

# Layout of the float64 CG state block (see include/scico_amd.h).
ST_BN, ST_TOL2, ST_NITER, ST_RELRES, ST_NUM = 0, 1, 2, 3, 8


class CGWork:
    """Device buffers of one CG solve of size ``n`` with ``maxiter`` steps."""

    def __init__(self, n: int, maxiter: int):
        self.n, self.maxiter = n, maxiter
        self.G = int(load().scxr_elem_blocks(n))
        self.state = torch.zeros(ST_NUM + 2 * (maxiter + 1), dtype=torch.float64, device=dev.DEVICE)
        self.flags = torch.zeros(maxiter + 1, dtype=torch.int32, device=dev.DEVICE)
        self.ws = torch.empty(2 * self.G, dtype=torch.float64, device=dev.DEVICE)
        self.pap = torch.empty(1, dtype=torch.float64, device=dev.DEVICE)

    def start(self, b, Ax, r, p, tol, atol):
        call("scxr_cg_start", self.n, b.data_ptr(), Ax.data_ptr(), r.data_ptr(), p.data_ptr(),
             self.state.data_ptr(), self.flags.data_ptr(), self.maxiter, float(tol), float(atol),
             self.ws.data_ptr(), dev.stream())

    def iterate(self, k, x, r, p, Ap, pap_parts, pap_n):
        call("scxr_cg_iter", self.n, k, self.maxiter, x.data_ptr(), r.data_ptr(), p.data_ptr(),
             Ap.data_ptr(), self.state.data_ptr(), self.flags.data_ptr(), pap_parts.data_ptr(),
             pap_n, self.ws.data_ptr(), dev.stream())

    def iterate_t(self, k, x, r, p, pT, Ap, pap_parts, pap_n):
        """``iterate`` for an M x N image ``p`` that also writes its transpose
        ``pT`` (N x M) for ``scxr_xr2d_forward_t``; bitwise the same x, r, p."""
        M, N = p.shape
        call("scxr_cg_iter_t", M, N, k, self.maxiter, x.data_ptr(), r.data_ptr(), p.data_ptr(),
             pT.data_ptr(), Ap.data_ptr(), self.state.data_ptr(), self.flags.data_ptr(),
             pap_parts.data_ptr(), pap_n, self.ws.data_ptr(), dev.stream())

    def info(self):
        return {"num_iter": self.state[ST_NITER], "rel_res": self.state[ST_RELRES]}


def cg(A, b, x0=None, *, tol: float = 1e-5, atol: float = 0.0, maxiter: int = 1000,
       info: bool = True, M=None, check_every: int = 16):
    """Solve ``A x = b`` for symmetric positive definite ``A``.

    ``A`` is a callable on device arrays whose kernels honour the
    predicate context (every operator of this package does).  Returns
    ``(x, info)`` with ``info["num_iter"]`` and ``info["rel_res"]`` as 0-d
    device tensors.
    """
    if M is not None:
        raise NotImplementedError("Preconditioned CG is not supported.")
    b = dev.as_device(b)
    if x0 is None:
        if isinstance(A, LinearOperator):
            x0 = torch.zeros(A.input_shape, dtype=b.dtype, device=b.device)
        else:
            raise ValueError("Argument 'x0' must be specified if argument 'A' is not a LinearOperator.")
    x = dev.as_device(x0).clone()
    n = x.numel()
    w = CGWork(n, int(maxiter))
    r = torch.empty_like(x)
    p = torch.empty_like(x)
    Ax = A(x)
    w.start(b, Ax, r, p, tol, atol)
    for k in range(int(maxiter)):
        if check_every and maxiter > check_every and k and k % check_every == 0:
            if int(w.flags[k].item()) == 0:
                break
        with predicated(w.flags[k : k + 1]):
            Ap = A(p)
            vec.dot(p, Ap, out=w.pap[0])
        w.iterate(k, x, r, p, Ap, w.pap, 1)
    return (x, w.info()) if info else x
