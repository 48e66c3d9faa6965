"""The reference's ADMM-TV <-> AcceleratedPGM-TV pin (``test_tvnorm.py:26-120``)
run on the GPU at the reference's own bar, on the generic ADMM path and on
the fused device path: SNR(ADMM, APGM) > 50 dB and the TV value equal to
``lam * norm(FD(x))`` to rel_res < 1e-6.  Each GPU solution is also
compared with the oracle's solution of the same problem."""

import numpy as np
import pytest

from oracle import solvers as S
from scico_amd import functional, linop, loss, metric
from scico_amd.optimize import ADMM, AcceleratedPGM, LinearSubproblemSolver
from scico_amd.optimize._admm import _FusedTV

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _solve(y, lam, isotropic, circular, fused, admm_its, L0, apgm_its):
    f = loss.SquaredL2Loss(y=y)
    g = lam * (functional.L21Norm() if isotropic else functional.L1Norm())
    C = linop.FiniteDifference(y.shape, circular=circular, append=None if circular else 0)
    old = _FusedTV.enabled
    _FusedTV.enabled = fused
    try:
        solver = ADMM(f=f, g_list=[g], C_list=[C], rho_list=[1e1], x0=y, maxiter=admm_its,
                      subproblem_solver=LinearSubproblemSolver(
                          cg_kwargs={"tol": 1e-4, "maxiter": 25}))
        assert (solver._fused is not None) == fused
        x_tvdn = solver.solve()
        cg_its = [int(v) for v in solver.history(transpose=True).CG_It]
    finally:
        _FusedTV.enabled = old
    tv = functional.IsotropicTVNorm if isotropic else functional.AnisotropicTVNorm
    h = lam * tv(circular=circular, input_shape=y.shape)
    x_aprx = AcceleratedPGM(f=f, g=h, L0=L0, x0=y, maxiter=apgm_its).solve()
    gv, hv = float(g(C(x_tvdn))), float(h(x_tvdn))
    return x_tvdn, x_aprx, cg_its, abs(gv - hv) / max(abs(gv), abs(hv))


def _check(y, lam, isotropic, circular, fused, admm_its, L0, apgm_its, label):
    x_tvdn, x_aprx, cg_its, rr = _solve(y, lam, isotropic, circular, fused, admm_its, L0,
                                        apgm_its)
    snr = metric.snr(x_tvdn, x_aprx)
    xo_admm, xo_apgm, hist = S.tv_denoise_pair(y, lam, isotropic, circular, admm_its, L0,
                                               apgm_its)
    e_admm = rel_l2(x_tvdn.cpu().numpy(), xo_admm)
    e_apgm = rel_l2(x_aprx.cpu().numpy(), xo_apgm)
    cg_oracle = [int(h[4]) for h in hist]
    cg_diff = sum(a != b for a, b in zip(cg_its, cg_oracle))
    print(f"  {label} iso={isotropic} circ={circular} fused={fused}: SNR {snr:.2f} dB, "
          f"TV rel_res {rr:.1e}, rel-L2 vs oracle ADMM {e_admm:.1e} APGM {e_apgm:.1e}, "
          f"CG counts differing {cg_diff}/{len(cg_its)}", flush=True)
    assert snr > 50
    assert rr < 1e-6
    # Same problem, same arithmetic: the iterates agree with the oracle's to
    # float32 summation-order noise (the denoising system is well conditioned).
    assert e_admm <= 1e-5 and e_apgm <= 1e-5
    # CG counts may differ only where the stopping test sits within rounding
    # of its threshold
    assert cg_diff <= max(2, len(cg_its) // 25)


@pytest.mark.parametrize("fused", [False, True])
@pytest.mark.parametrize("isotropic", [True, False])
@pytest.mark.parametrize("circular", [False, True])
def test_tv_admm_apgm_pin_2d(gpu, fused, isotropic, circular):
    """``test_tvnorm.py:74-120``: 32^2, rho 10, 150 ADMM / 400 APGM iterations."""
    _, y = S.tv_denoise_problem(2)
    _check(y, 5e-2, isotropic, circular, fused, 150, 1e3, 400, "2D")


@pytest.mark.parametrize("fused", [False, True])
@pytest.mark.parametrize("circular", [True, False])
def test_tv_admm_apgm_pin_1d(gpu, fused, circular):
    """``test_tvnorm.py:26-58``: 128 samples, anisotropic, 50 ADMM / 100 APGM."""
    _, y = S.tv_denoise_problem(1)
    _check(y, 5e-2, False, circular, fused, 50, 5e2, 100, "1D")
