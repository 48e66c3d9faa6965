"""ADMM (``scico/optimize/_admm.py:35-378``).

``ADMM.step`` follows the reference algorithm for any combination of the
package's operators and functionals.  When the problem has the
tomographic ADMM-TV structure

    f = SquaredL2Loss(y, XRayTransform2D, scale=0.5),
    g = lam * L21Norm() or lam * L1Norm(),   C = FiniteDifference (2D),
    LinearSubproblemSolver (scico CG),       alpha = 1,

the same iteration runs through fused kernels (:class:`_FusedTV`): the
x-step operator ``rho C^T C + A^T A`` and the right-hand side are fused
into the back projection, the CG scalars and stopping test stay on the
device, and the z/u update and iteration statistics are single passes.
With ``XRayTransform3D`` (3D, FiniteDifference over all axes), or the
identity of a denoising loss ``SquaredL2Loss(y)`` (1D to 3D), the same
structure runs with the FD terms in separate nd passes.
"""

from __future__ import annotations

import ctypes
from typing import List, Optional

import numpy as np
import torch

from .. import _device as dev
from .. import _vec as vec
from .._lib import call, load
from ..functional import L1Norm, L21Norm, ScaledFunctional
from ..linop import FiniteDifference, Identity
from ..linop.xray import XRayTransform2D, XRayTransform3D
from ..loss import SquaredL2Loss
from ..solver import CGWork
from ._admmaux import GenericSubproblemSolver, LinearSubproblemSolver
from ._common import Optimizer


class ADMM(Optimizer):
    def __init__(self, f, g_list, C_list, rho_list, alpha: float = 1.0, x0=None,
                 subproblem_solver=None, **kwargs):
        N = len(g_list)
        if len(C_list) != N:
            raise ValueError(f"len(C_list)={len(C_list)} not equal to len(g_list)={N}.")
        if len(rho_list) != N:
            raise ValueError(f"len(rho_list)={len(rho_list)} not equal to len(g_list)={N}.")
        self.f = f
        self.g_list = list(g_list)
        self.C_list = list(C_list)
        self.rho_list = [float(r) for r in rho_list]
        self.alpha = float(alpha)
        if subproblem_solver is None:
            subproblem_solver = GenericSubproblemSolver()
        self.subproblem_solver = subproblem_solver
        self.subproblem_solver.internal_init(self)
        if x0 is None:
            x0 = torch.zeros(C_list[0].input_shape, dtype=torch.float32, device=dev.DEVICE)
        self.x = dev.as_device(x0).clone()
        self.z_list, self.z_list_old = self.z_init(self.x)
        self.u_list = self.u_init(self.x)
        super().__init__(**kwargs)
        self._fused = _FusedTV.build(self)

    # --- reference API
    def _working_vars_finite(self) -> bool:
        for v in [self.x] + self.z_list + self.u_list:
            if not bool(torch.isfinite(v).all().item()):
                return False
        return True

    def _objective_evaluatable(self):
        return (not self.f or self.f.has_eval) and all([g.has_eval for g in self.g_list])

    def _itstat_extra_fields(self):
        fields = {"Prml Rsdl": "%9.3e", "Dual Rsdl": "%9.3e"}
        attrib = ["norm_primal_residual()", "norm_dual_residual()"]
        if type(self.subproblem_solver) is LinearSubproblemSolver and \
                self.subproblem_solver.cg_function == "scico":
            fields.update({"CG It": "%5d", "CG Res": "%9.3e"})
            attrib.extend(["subproblem_solver.info['num_iter']",
                           "subproblem_solver.info['rel_res']"])
        return fields, attrib

    def _state_variable_names(self) -> List[str]:
        return ["x", "z_list", "z_list_old", "u_list"]

    def minimizer(self):
        return self.x

    def objective(self, x=None, z_list=None):
        """``f(x) + sum_i g_i(z_i)`` (``_admm.py:209-252``), device float64."""
        if (x is None) != (z_list is None):
            raise ValueError("Both or neither of arguments 'x' and 'z_list' must be supplied.")
        if x is None:
            if self._fused is not None and self._fused.has_stats:
                return self._fused.stat(0)
            x, z_list = self.x, self.z_list
        out = torch.zeros((), dtype=torch.float64, device=dev.DEVICE)
        if self.f:
            out = out + self.f(x)
        for g, z in zip(self.g_list, z_list):
            out = out + g(z)
        return out

    def norm_primal_residual(self, x=None):
        """``sqrt(sum_i rho_i ||C_i x - z_i||^2)``; uses ``self.x`` (``_admm.py:254-277``)."""
        if x is None and self._fused is not None and self._fused.has_stats:
            return self._fused.stat(1)
        s = torch.zeros((), dtype=torch.float64, device=dev.DEVICE)
        for rho, C, z in zip(self.rho_list, self.C_list, self.z_list):
            s = s + rho * vec.reduce(vec.RED_SQDIST, C(self.x), z)
        return torch.sqrt(s)

    def norm_dual_residual(self):
        """``||sum_i rho_i C_i^T (z_i - z_i_old)||`` (``_admm.py:279-295``)."""
        if self._fused is not None and self._fused.has_stats:
            return self._fused.stat(2)
        acc = None
        for rho, z, zo, C in zip(self.rho_list, self.z_list, self.z_list_old, self.C_list):
            t = vec.scale(rho, C.adj(vec.add(z, -1.0, zo)))
            acc = t if acc is None else vec.add(acc, 1.0, t)
        return vec.norm(acc)

    def z_init(self, x0):
        z = [C(x0) for C in self.C_list]
        return z, list(z)

    def u_init(self, x0):
        return [torch.zeros_like(z) for z in self.z_list]

    def step(self):
        """One ADMM iteration (``_admm.py:334-378``)."""
        if self._fused is not None:
            self._fused.step()
            return
        self.x = self.subproblem_solver.solve(self.x)
        self.z_list_old = list(self.z_list)
        for i, (rho, g, C, z, u) in enumerate(
                zip(self.rho_list, self.g_list, self.C_list, self.z_list, self.u_list)):
            if self.alpha == 1.0:
                Cx = C(self.x)
            else:
                Cx = vec.axpby(self.alpha, C(self.x), 1.0 - self.alpha, z)
            z = g.prox(vec.add(Cx, 1.0, u), 1 / rho, v0=z)
            u = vec.add(vec.add(u, 1.0, Cx), -1.0, z)
            self.z_list[i] = z
            self.u_list[i] = u

    def _after_load_state(self):
        if self._fused is not None:
            self._fused.adopt_state()


def _tv_functional(g):
    """(lam, isotropic) if g is lam * L21Norm(0) / lam * L1Norm, else None."""
    scale = 1.0
    base = g
    if isinstance(g, ScaledFunctional):
        scale, base = g.scale, g.functional
    if type(base) is L21Norm and base.l2_axis == 0:
        return scale, True
    if type(base) is L1Norm:
        return scale, False
    return None


class _FusedTV:
    """Fused device implementation of ADMM-TV with an X-ray transform."""

    enabled = True  # tests switch this off to run the generic path

    @staticmethod
    def build(admm: ADMM):
        if not _FusedTV.enabled:
            return None
        ss = admm.subproblem_solver
        if type(ss) is not LinearSubproblemSolver or ss.cg_function != "scico":
            return None
        f = admm.f
        if not (isinstance(f, SquaredL2Loss) and f.scale == 0.5 and
                (isinstance(f.A, (XRayTransform2D, XRayTransform3D)) or type(f.A) is Identity)):
            return None
        if len(admm.C_list) != 1 or admm.alpha != 1.0:
            return None
        C = admm.C_list[0]
        if not (type(C) is FiniteDifference and C.input_shape == f.A.input_shape):
            return None
        tv = _tv_functional(admm.g_list[0])
        if tv is None:
            return None
        if float(ss.cg_kwargs.get("atol", 0.0)) < 0 or set(ss.cg_kwargs) - {"tol", "atol", "maxiter"}:
            return None
        return _FusedTV(admm, tv[0], tv[1])

    def __init__(self, admm: ADMM, lam: float, isotropic: bool):
        self.admm = admm
        self.A = admm.f.A
        self.C: FiniteDifference = admm.C_list[0]
        self.lam = float(lam)
        self.iso = bool(isotropic)
        self.rho = float(admm.rho_list[0])
        self.circ = int(self.C.circular)
        # 2D projector: FD terms fused into the projector epilogues (the
        # general-geometry kernels have none: nd passes instead)
        self.epi = isinstance(self.A, XRayTransform2D) and self.A.tiled
        # forward / back projection with out= and pred= (for the identity
        # of a denoising loss, a predicated copy)
        if type(self.A) is Identity:
            self._fwd = self._bwd = lambda v, out, pred=None: vec.copy(v, out=out, pred=pred)
        else:
            self._fwd, self._bwd = self.A.project, self.A.back_project
        cgk = admm.subproblem_solver.cg_kwargs
        self.tol = float(cgk.get("tol", 1e-4))
        self.atol = float(cgk.get("atol", 0.0))
        self.cg_maxiter = int(cgk.get("maxiter", 100))
        shape = self.A.input_shape
        self.ndim = len(shape)
        self.shape_c = (ctypes.c_int32 * self.ndim)(*shape)
        n = int(np.prod(shape))
        lib = load()
        self.G = int(lib.scxr_elem_blocks(n))
        if self.epi:
            self.M, self.N = shape
            info = (ctypes.c_int32 * 4)()
            call("scxr_xr2d_info", self.A.plan, info)
            self.n_fwd, self.n_adj = int(info[0]), int(info[1])
        else:
            self.n_fwd, self.n_adj = 1, self.G
        f64 = dict(dtype=torch.float64, device=dev.DEVICE)
        self.cg = CGWork(n, self.cg_maxiter)
        self.r = torch.empty(shape, dtype=torch.float32, device=dev.DEVICE)
        self.p = torch.empty_like(self.r)
        # p transposed, written by the CG direction update for the 2D
        # projector's column-fast views (saves its transpose pass)
        self.pT = torch.empty(tuple(reversed(shape)), dtype=torch.float32,
                              device=dev.DEVICE) if self.epi else None
        self.Ap = torch.empty_like(self.r)
        self.rhs = torch.empty_like(self.r)
        self.sino = torch.empty(self.A.output_shape, dtype=torch.float32, device=dev.DEVICE)
        # A x of the current iterate: written by the residual projection that
        # ends an iteration and read by the x-step operator of the next (x
        # does not change in between), so an iteration projects x once
        self.ax = torch.empty_like(self.sino)
        # A^T y: y is fixed during a solve, so the x-step right-hand side
        # A^T y + rho C^T (z - u) takes it from here instead of back
        # projecting y every iteration (refreshed by adopt_state / reset)
        self.aty = torch.empty_like(self.r)
        self.pap = torch.empty(self.n_adj, **f64)
        self.ws_prml = torch.empty(self.G, **f64)
        self.ws_g = torch.empty(self.G, **f64)
        self.ws_dual = torch.empty(self.G, **f64)
        self.ws_res = torch.empty(self.n_fwd, **f64)
        self.hist_cap = 0
        self.hist = None
        self.counter = torch.zeros(1, dtype=torch.int32, device=dev.DEVICE)
        self.n_steps = 0
        self.has_stats = False
        self.thr = float(np.float32((1.0 / self.rho) * self.lam))
        self.adopt_state()
        self._reserve(max(64, admm.maxiter))

    def adopt_state(self):
        """Own contiguous working copies of x, z, u (updated in place)."""
        a = self.admm
        self.x = a.x.contiguous().clone()
        self.z = a.z_list[0].contiguous().clone()
        self.u = a.u_list[0].contiguous().clone()
        self.zold = a.z_list_old[0].contiguous().clone()
        self._publish()
        self._project_x()

    def _project_x(self):
        """ax = A x for the iterate just set and aty = A^T y (device work only)."""
        if self.epi:
            call("scxr_xr2d_forward", self.A.plan, self.x.data_ptr(), self.ax.data_ptr(), None,
                 dev.stream())
        else:
            self._fwd(self.x, out=self.ax)
        self._ax_ver = self.x._version  # the kernels' own writes do not count
        self._back_project_y()

    def _back_project_y(self):
        """aty = A^T y, noting which y (and which version of it) it came from."""
        y = self.admm.f.y
        if self.epi:
            call("scxr_xr2d_adjoint", self.A.plan, y.data_ptr(), self.aty.data_ptr(), None,
                 dev.stream())
        else:
            self._bwd(y, out=self.aty)
        self._aty_of = (y, y._version)

    def _check_inputs(self):
        """Host-side checks before an iteration is enqueued (a captured
        iteration replays what it was captured with): a new iterate assigned
        to ``admm.x`` is adopted, one modified in place by torch operations
        is projected again, and y is back projected again if ``f.y`` was
        replaced or modified in place since aty was computed."""
        if self.admm.x is not self.x:
            self.adopt_state()
        elif self.x._version != self._ax_ver:
            self._project_x()
        y = self.admm.f.y
        src, ver = self._aty_of
        if y is not src or y._version != ver:
            self._back_project_y()

    def _publish(self):
        a = self.admm
        a.x = self.x
        a.z_list = [self.z]
        a.u_list = [self.u]
        a.z_list_old = [self.zold]

    def reset(self, x0=None):
        """Restart from ``x0`` (zeros by default) with z = C x0, u = 0 and an
        empty statistics history; device kernels only (no allocation), so it
        may run between replays of a captured iteration."""
        if x0 is None:
            vec.fill(self.x, 0.0)
        else:
            vec.copy(dev.as_device(x0), out=self.x)
        self.C._eval(self.x, out=self.z)
        vec.copy(self.z, out=self.zold)
        vec.fill(self.u, 0.0)
        self._project_x()
        vec.fill(self.counter.view(torch.float32), 0.0)
        self.n_steps = 0
        self.has_stats = False

    def _reserve(self, rows: int):
        if rows <= self.hist_cap:
            return
        new = torch.zeros((rows, 5), dtype=torch.float64, device=dev.DEVICE)
        if self.hist is not None:
            new[: self.hist_cap] = self.hist
        self.hist, self.hist_cap = new, rows

    def stat(self, col: int):
        return self.hist[self.n_steps - 1, col]

    def step(self):
        if self.n_steps >= self.hist_cap:
            self._reserve(2 * self.hist_cap)
        self.enqueue()
        self.n_steps += 1
        self.has_stats = True
        self._publish()
        a = self.admm
        a.subproblem_solver.info = {"num_iter": self.hist[self.n_steps - 1, 3],
                                    "rel_res": self.hist[self.n_steps - 1, 4]}

    def enqueue(self):
        """Enqueue one full ADMM iteration (no host synchronisation; capturable).
        It starts from ``ax`` = A x, left by the previous iteration (or by
        ``adopt_state`` / ``reset``), and leaves A x of the new iterate there;
        A^T y comes from ``aty``, back projected again when ``f.y`` has been
        replaced or modified in place since (checked on the host here, so a
        captured graph keeps the A^T y it was captured with)."""
        self._check_inputs()
        if not self.epi:
            self._enqueue_nd()
            return
        A, plan, s = self.A, self.A.plan, dev.stream()
        P = lambda t: t.data_ptr()
        y = self.admm.f.y
        # x-step: rhs = A^T y + rho C^T (z - u), A^T y kept from the start
        # (the same operations as the adjoint's right-hand-side epilogue)
        call("scxr_admm_rhs2d", self.M, self.N, P(self.aty), P(self.z), P(self.u), self.rho,
             self.circ, P(self.rhs), s)
        # A_lhs x0 = rho C^T C x + A^T A x, with A x from the last iteration
        call("scxr_xr2d_adjoint_epi", plan, P(self.ax), P(self.Ap), 1, self.rho, self.circ,
             P(self.x), None, None, None, None, s)
        cg = self.cg
        cg.start(self.rhs, self.Ap, self.r, self.p, self.tol, self.atol)
        last = self.cg_maxiter - 1
        for k in range(self.cg_maxiter):
            flag = cg.flags[k:].data_ptr()
            if k == 0:
                call("scxr_xr2d_forward", plan, P(self.p), P(self.sino), flag, s)
            else:  # p transposed by the previous step's direction update
                call("scxr_xr2d_forward_t", plan, P(self.p), P(self.pT), P(self.sino), flag, s)
            call("scxr_xr2d_adjoint_epi", plan, P(self.sino), P(self.Ap), 1, self.rho,
                 self.circ, P(self.p), None, None, P(self.pap), flag, s)
            if k < last:
                cg.iterate_t(k, self.x, self.r, self.p, self.pT, self.Ap, self.pap, self.n_adj)
            else:
                cg.iterate(k, self.x, self.r, self.p, self.Ap, self.pap, self.n_adj)
        # z / u update and iteration statistics
        call("scxr_admm_zu2d", self.M, self.N, P(self.x), P(self.z), P(self.u), P(self.zold),
             self.thr, self.circ, int(self.iso), P(self.ws_prml), P(self.ws_g), s)
        call("scxr_admm_dual2d", self.M, self.N, P(self.z), P(self.zold), self.rho, self.circ,
             P(self.ws_dual), s)
        call("scxr_xr2d_forward_res", plan, P(self.x), P(self.ax), P(y), P(self.ws_res), None,
             s)
        self._itstat()

    def _itstat(self):
        P = lambda t: t.data_ptr()
        call("scxr_admm_itstat", P(self.ws_res), self.n_fwd, P(self.ws_g), self.G,
             P(self.ws_prml), self.G, P(self.ws_dual), self.G, P(self.cg.state), self.lam,
             self.rho, P(self.hist), P(self.counter), self.hist_cap, dev.stream())

    def _enqueue_nd(self):
        """The same iteration for any projector with ``project`` /
        ``back_project(out=, pred=)``; FD terms in nd passes."""
        A, s, nd, sh = self.A, dev.stream(), self.ndim, self.shape_c
        P = lambda t: t.data_ptr()
        y = self.admm.f.y
        call("scxr_admm_rhs_nd", nd, sh, P(self.aty), P(self.z), P(self.u), self.rho, self.circ,
             P(self.rhs), s)
        self._bwd(self.ax, out=self.Ap)  # A x from the last iteration
        call("scxr_admm_lhs_nd", nd, sh, P(self.x), P(self.Ap), self.rho, self.circ, P(self.Ap),
             None, None, s)
        cg = self.cg
        cg.start(self.rhs, self.Ap, self.r, self.p, self.tol, self.atol)
        for k in range(self.cg_maxiter):
            flag = cg.flags[k : k + 1]
            self._fwd(self.p, out=self.sino, pred=flag)
            self._bwd(self.sino, out=self.Ap, pred=flag)
            call("scxr_admm_lhs_nd", nd, sh, P(self.p), P(self.Ap), self.rho, self.circ,
                 P(self.Ap), P(self.pap), P(flag), s)
            cg.iterate(k, self.x, self.r, self.p, self.Ap, self.pap, self.n_adj)
        call("scxr_admm_zu_nd", nd, sh, P(self.x), P(self.z), P(self.u), P(self.zold), self.thr,
             self.circ, int(self.iso), P(self.ws_prml), P(self.ws_g), s)
        call("scxr_admm_dual_nd", nd, sh, P(self.z), P(self.zold), self.rho, self.circ,
             P(self.ws_dual), s)
        self._fwd(self.x, out=self.ax)
        vec.reduce(vec.RED_SQDIST, self.ax, y, out=self.ws_res[0])
        self._itstat()
