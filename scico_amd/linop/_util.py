"""Operator utilities (``scico/linop/_util.py:27-110``)."""

from __future__ import annotations

import numpy as np
import torch

from .. import _device as dev
from .. import _vec as vec


def power_iteration(A, maxiter: int = 100, key=None, device=None):
    """Largest eigenvalue of a diagonalisable operator by power iteration.

    The start vector is standard normal from NumPy (``key`` is an integer
    seed).  Returns ``(mu, v)`` with ``mu`` a 0-d float64 device tensor.
    """
    rng = np.random.default_rng(0 if key is None else key)
    v = dev.as_device(rng.standard_normal(A.input_shape).astype(np.float32))
    v = vec.div_dev(v, vec.norm(v))
    mu = torch.zeros((), dtype=torch.float64, device=dev.DEVICE)
    for _ in range(maxiter):
        Av = A(v)
        nAv = vec.norm(Av)
        mu = vec.dot(v, Av) / (vec.norm(v) ** 2)
        v = vec.div_dev(Av, nAv)
    if float(vec.norm(v)) == 0.0:
        mu = torch.zeros((), dtype=torch.float64, device=dev.DEVICE)
    return mu, v


def operator_norm(A, maxiter: int = 100, key=None, device=None):
    """``sqrt(lambda_max(A^H A))`` (``_util.py:75-110``)."""
    return torch.sqrt(power_iteration(A.H @ A, maxiter=maxiter, key=key)[0])
