"""Proximal gradient methods (``scico/optimize/_pgm.py:30-252``,
``scico/optimize/_pgmaux.py:23-62``).

``PGM`` / ``AcceleratedPGM`` follow the reference algorithm for any
``f`` with a gradient and ``g`` with a prox.  For the deconvolution
structure

    f = SquaredL2Loss(y, CircularConvolve (one real 1024 x 1024 image),
                      scale=0.5),
    g = lam * IsotropicTVNorm() or lam * AnisotropicTVNorm()  (2D),
    constant step size,

the iteration runs through fused kernels (:class:`_FusedDeconv`): the
prox writes x, the partials of the residual and of TV(x) and the
extrapolated point in one pass, and one FFT convolution by ``G = A^T A`` per iteration gives both
the next gradient (``G v = (1 + c) G x - c G x_prev`` by linearity) and,
on the way through the spectrum, the objective (Parseval).  The reference
evaluates three convolutions per iteration (``A v``, ``A^T r`` and
``A x`` for the objective).
"""

from __future__ import annotations

from typing import List

import numpy as np
import torch

from .. import _device as dev
from .. import _vec as vec
from .._lib import call, load
from ._common import Optimizer


class PGMStepSize:
    """Constant reciprocal step size ``L`` (``_pgmaux.py:23-62``)."""

    def internal_init(self, pgm):
        self.pgm = pgm

    def update(self, v):
        return self.pgm.L


class PGM(Optimizer):
    def __init__(self, f, g, L0: float, x0, step_size=None, **kwargs):
        self.f = f
        if g.has_prox is not True:
            raise ValueError(f"Functional 'g' ({type(g)}) must have a prox method.")
        self.g = g
        if step_size is None:
            step_size = PGMStepSize()
        if type(step_size) is not PGMStepSize:
            raise NotImplementedError("Only the constant PGMStepSize is supported.")
        self.step_size = step_size
        self.step_size.internal_init(self)
        self.L = float(L0)
        self.fixed_point_residual = float("inf")
        self.x = dev.as_device(x0).clone()
        super().__init__(**kwargs)

    def x_step(self, v, L):
        """``g.prox(v - (1/L) grad f(v), 1/L)`` (``_pgm.py:106-117``)."""
        u = vec.axpby(1.0, v, -(1.0 / L), self.f.grad(v))
        return self.g.prox(u, 1.0 / L)

    def _working_vars_finite(self) -> bool:
        return bool(torch.isfinite(self.x).all().item())

    def _objective_evaluatable(self):
        return self.f.has_eval and self.g.has_eval

    def _itstat_extra_fields(self):
        return {"L": "%9.3e", "Residual": "%9.3e"}, ["L", "norm_residual()"]

    def _state_variable_names(self) -> List[str]:
        return ["x", "L"]

    def minimizer(self):
        return self.x

    def objective(self, x=None):
        """``f(x) + g(x)`` (``_pgm.py:151-156``), device float64."""
        fused = getattr(self, "_fused", None)
        if x is None and fused is not None and fused.n_steps:
            return fused.stat(0)
        x = self.x if x is None else x
        return self.f(x) + self.g(x)

    def norm_residual(self):
        fused = getattr(self, "_fused", None)
        if fused is not None and fused.n_steps:
            return fused.stat(1)
        return self.fixed_point_residual

    def _momentum(self):
        """Extrapolation coefficient of the step (0 for PGM)."""
        return 0.0

    def step(self):
        if getattr(self, "_fused", None) is not None:
            self._fused.step(self._momentum())
            return
        self.L = self.step_size.update(self.x)
        x = self.x_step(self.x, self.L)
        self.fixed_point_residual = vec.reduce(vec.RED_SQDIST, self.x, x, post=vec.POST_SQRT)
        self.x = x


class AcceleratedPGM(PGM):
    """FISTA (``_pgm.py:183-252``)."""

    def __init__(self, f, g, L0: float, x0, step_size=None, **kwargs):
        super().__init__(f=f, g=g, L0=L0, x0=x0, step_size=step_size, **kwargs)
        self.v = self.x.clone()
        self.t = 1.0
        self._fused = _FusedDeconv.build(self)

    def _momentum(self):
        # t_{k+1} = (1 + sqrt(1 + 4 t_k^2)) / 2 in float32, as the reference
        # evaluates it after the first step (data independent: host side)
        t_old = self.t
        f32 = np.float32
        self.t = float(f32(0.5) * (f32(1) + np.sqrt(f32(1) + f32(4) * f32(t_old) ** 2)))
        return float(f32(f32(t_old) - f32(1)) / f32(self.t))

    def _after_load_state(self):
        if getattr(self, "_fused", None) is not None:
            self._fused.adopt_state()

    def step(self):
        if self._fused is not None:
            self._fused.step(self._momentum())
            return
        x_old = self.x
        self.L = self.step_size.update(self.v)
        self.x = self.x_step(self.v, self.L)
        self.fixed_point_residual = vec.reduce(vec.RED_SQDIST, self.x, self.v,
                                               post=vec.POST_SQRT)
        # t_{k+1} = (1 + sqrt(1 + 4 t_k^2)) / 2 in float32, as the reference
        # evaluates it after the first step (the sequence does not depend on
        # the data, so it is computed on the host).
        t_old = self.t
        f32 = np.float32
        self.t = float(f32(0.5) * (f32(1) + np.sqrt(f32(1) + f32(4) * f32(t_old) ** 2)))
        c = float(f32(f32(t_old) - f32(1)) / f32(self.t))
        self.v = vec.axpby(1.0, self.x, c, vec.add(self.x, -1.0, x_old))

    def _working_vars_finite(self) -> bool:
        return bool(torch.isfinite(self.x).all().item()) and bool(torch.isfinite(self.v).all().item())

    def _state_variable_names(self) -> List[str]:
        return ["x", "v", "t", "L"]


class _FusedDeconv:
    """Fused device implementation of AcceleratedPGM-TV deconvolution."""

    enabled = True  # tests switch this off to run the generic path

    @staticmethod
    def build(pgm: AcceleratedPGM):
        from ..functional import ScaledFunctional
        from ..functional._tvnorm import TVNorm
        from ..linop import CircularConvolve
        from ..loss import SquaredL2Loss

        if not _FusedDeconv.enabled or not torch.cuda.is_available():
            return None
        f, g = pgm.f, pgm.g
        if not (type(f) is SquaredL2Loss and f.scale == 0.5 and type(f.A) is CircularConvolve):
            return None
        A = f.A
        if not (A.ndims == 2 and A.batch == 1 and A.input_shape == (1024, 1024)):
            return None
        scale, base = (g.scale, g.functional) if isinstance(g, ScaledFunctional) else (1.0, g)
        if not isinstance(base, TVNorm) or base.axes is not None:
            return None
        if tuple(pgm.x.shape) != A.input_shape or pgm.x.dtype != torch.float32:
            return None
        return _FusedDeconv(pgm, float(scale), base)

    def __init__(self, pgm: AcceleratedPGM, lam: float, tv):
        self.pgm = pgm
        self.A = pgm.f.A
        self.lam = lam
        self.iso = int(tv.isotropic)
        self.circ = int(tv.circular)
        self.M, self.N = self.A.input_shape
        n = self.M * self.N
        lib = load()
        f32 = dict(dtype=torch.float32, device=dev.DEVICE)
        f64 = dict(dtype=torch.float64, device=dev.DEVICE)
        y = dev.as_device(pgm.f.y)
        # G = A^T A on the half grid: |Hs|^2 with Hs the Hermitian part of
        # the spectrum (float64, rounded once), as SquaredL2Loss's gram
        hd = dev.to_numpy(self.A.h_dft).astype(np.complex128)
        hneg = np.roll(np.flip(hd, (0, 1)), 1, (0, 1))
        hs = 0.5 * (hd + np.conj(hneg))
        H1 = self.N // 2 + 1
        # half-grid arrays transposed (column k1 contiguous), as the kernels
        # keep the half spectrum
        self.gs = dev.as_device(np.ascontiguousarray((hs.real**2 + hs.imag**2)[:, :H1].T,
                                                     dtype=np.float32))
        yh = np.fft.rfft2(dev.to_numpy(y).astype(np.float64))
        self.Yh = torch.as_tensor(np.ascontiguousarray(yh.T.astype(np.complex64)),
                                  device=dev.DEVICE)
        self.aty = self.A.adj(y)
        self.n_obj = H1
        self.n_res = int(lib.scxr_tv_prox_apgm_parts(self.M, self.N))
        self.n_tv = self.n_res  # the prox kernel's TV partials, one per workgroup
        self.ws_obj = torch.empty(self.n_obj, **f64)
        self.ws_res = torch.empty(self.n_res, **f64)
        self.ws_tv = torch.empty(self.n_tv, **f64)
        self.u = torch.empty((self.M, self.N), **f32)
        self.gx = torch.zeros((self.M, self.N), **f32)
        self.obj_scale = 0.5 / float(n)
        self.hist_cap = 0
        self.hist = None
        self.counter = torch.zeros(1, dtype=torch.int32, device=dev.DEVICE)
        self.n_steps = 0
        self._reserve(max(64, pgm.maxiter))
        self.adopt_state()

    def _reserve(self, rows: int):
        if rows <= self.hist_cap:
            return
        new = torch.zeros((rows, 2), dtype=torch.float64, device=dev.DEVICE)
        if self.hist is not None:
            new[: self.hist_cap] = self.hist
        self.hist, self.hist_cap = new, rows

    def stat(self, col: int):
        return self.hist[self.n_steps - 1, col]

    def _consts(self):
        L = float(self.pgm.L)
        b = -(1.0 / L)
        lam_eff = (1.0 / L) * self.lam  # ScaledFunctional.prox(v, 1/L)
        tau = float(np.float32(np.float32(np.sqrt(np.float32(2.0))) * np.float32(4)) *
                    np.float32(lam_eff))
        return b, tau

    def adopt_state(self):
        """Own contiguous x and v, and the first gradient step from them."""
        p = self.pgm
        self.x = p.x.contiguous().clone()
        self.v = p.v.contiguous().clone() if hasattr(p, "v") else self.x.clone()
        self._publish()
        self._prime()

    def _publish(self):
        self.pgm.x = self.x
        if hasattr(self.pgm, "v"):
            self.pgm.v = self.v

    def _prime(self):
        """The state a step starts from: u = v - (1/L) grad f(v) and gx = G x
        (the step forms G v_next = (1 + c) G x_next - c gx)."""
        b, _ = self._consts()
        P = lambda t: t.data_ptr()
        s = dev.stream()
        vec.fill(self.gx, 0.0)
        call("scxr_cconv_apgm_grad", self.A.plan, P(self.v), P(self.gs), None, None, P(self.gx),
             P(self.v), P(self.aty), 0.0, b, P(self.u), None, s)
        if not torch.equal(self.v, self.x):  # a restored mid-run state: gx = G x as well
            scratch = torch.empty_like(self.u)
            call("scxr_cconv_apgm_grad", self.A.plan, P(self.x), P(self.gs), None, None,
                 P(self.gx), P(self.x), P(self.aty), 0.0, b, P(scratch), None, s)
            torch.cuda.current_stream().synchronize()

    def step(self, c: float):
        if self.n_steps >= self.hist_cap:
            self._reserve(2 * self.hist_cap)
        self.enqueue(c)
        self.n_steps += 1
        self._publish()

    def enqueue(self, c: float):
        """One iteration (no host synchronisation; capturable)."""
        b, tau = self._consts()
        P = lambda t: t.data_ptr()
        s = dev.stream()
        call("scxr_tv_prox_apgm", self.M, self.N, P(self.u), P(self.x), P(self.v), tau, self.iso,
             self.circ, c, P(self.ws_res), P(self.ws_tv), None, s)
        call("scxr_cconv_apgm_grad_itstat", self.A.plan, P(self.x), P(self.gs), P(self.Yh),
             P(self.ws_obj), P(self.gx), P(self.v), P(self.aty), c, b, P(self.u), self.obj_scale,
             P(self.ws_tv), self.n_tv, self.lam, P(self.ws_res), self.n_res, P(self.hist),
             P(self.counter), self.hist_cap, None, s)
