"""Pins for the solver-side oracle (CPU)."""

import json
import os

import numpy as np
import pytest
from scipy.optimize import minimize

from oracle import solvers as S

GOLD = os.path.join(os.path.dirname(__file__), "golden")
KAT = json.load(open(os.path.join(GOLD, "reference_kat.json")))


def test_fd_reference_outputs():
    k = KAT["fd_append0_1d"]
    np.testing.assert_array_equal(S.fd(np.array(k["x"], np.float32))[0], k["out"])
    k = KAT["fd_2d_doc"]
    g = S.fd(np.array(k["x"], np.float32))
    # append=0 adds a trailing zero row/column to the doctest's plain diff
    np.testing.assert_array_equal(g[0][:-1], k["axis0"])
    np.testing.assert_array_equal(g[1][:, :-1], k["axis1"])
    assert np.all(g[0][-1] == 0) and np.all(g[1][:, -1] == 0)


@pytest.mark.parametrize("shape", [(7,), (5, 6), (3, 4, 5)])
@pytest.mark.parametrize("circular", [False, True])
def test_fd_adjoint(shape, circular):
    rng = np.random.default_rng(0)
    x = rng.standard_normal(shape).astype(np.float32)
    g = rng.standard_normal((len(shape), *shape)).astype(np.float32)
    a = np.vdot(S.fd(x, circular).astype(np.float64), g)
    b = np.vdot(x.astype(np.float64), S.fd_adj(g, circular))
    assert abs(a - b) <= 1e-5 * max(abs(a), abs(b))
    if circular:
        for ax in range(len(shape)):
            np.testing.assert_allclose(S.fd(x, True)[ax], np.roll(x, -1, ax) - x, atol=1e-6)


def test_l21_closed_form():
    k = KAT["l21_ones"]
    x = np.ones(k["shape"], np.float32)
    assert abs(S.l21_norm(x) - k["norm"]) < 1e-4
    np.testing.assert_allclose(S.l21_prox(x, 1.0), k["prox_scale"] * x, rtol=1e-5)


def _brute_prox(f, v, lam):
    r = minimize(lambda z: 0.5 * np.sum((z - v) ** 2) + lam * f(z), v, method="Nelder-Mead",
                 options={"xatol": 1e-9, "fatol": 1e-12, "maxiter": 20000})
    return r.x


def test_prox_brute_force():
    """Prox optimality (reference pattern functional/prox.py)."""
    rng = np.random.default_rng(1)
    v = rng.standard_normal(4)
    lam = 0.3
    np.testing.assert_allclose(S.l1_prox(v.astype(np.float32), lam),
                               _brute_prox(lambda z: np.sum(np.abs(z)), v, lam), atol=1e-4)
    v2 = rng.standard_normal((2, 2))
    l21 = lambda z: np.sum(np.sqrt(np.sum(z.reshape(2, 2) ** 2, axis=0)))
    np.testing.assert_allclose(S.l21_prox(v2.astype(np.float32), lam),
                               _brute_prox(l21, v2.ravel(), lam).reshape(2, 2), atol=1e-4)
    z = S.l21_prox(np.zeros((2, 3), np.float32), 1.0)
    assert np.all(z == 0)  # no_nan_divide: zero columns stay zero


def test_haar_gram():
    k = KAT["haar_gram"]
    x = np.random.default_rng(2).standard_normal((3, 4)).astype(np.float32)
    for ax in (0, 1):
        w = S.haar(x, axes=[ax])
        np.testing.assert_allclose(S.haar_adj(w, axes=[ax]), k["single_axis"] * x, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(S.haar_adj(S.haar(x)), k["two_axes"] * x, rtol=1e-5, atol=1e-6)


def test_tv_value_identity():
    x = np.random.default_rng(3).standard_normal((8, 9)).astype(np.float32)
    g = S.fd(x)
    assert abs(S.tv_norm(x) - np.sum(np.sqrt(g[0].astype(np.float64) ** 2 + g[1] ** 2))) < 1e-6


def _snr(ref, cmp):
    """``scico.metric.snr`` (``metric.py:48-62``) in float64."""
    ref = np.asarray(ref, np.float64)
    return 10 * np.log10(np.var(ref) / np.mean((ref - np.asarray(cmp, np.float64)) ** 2))


@pytest.mark.parametrize("isotropic", [True, False])
@pytest.mark.parametrize("circular", [False, True])
def test_tv_admm_apgm_reference_pin_2d(isotropic, circular):
    """``test_tvnorm.py:74-120`` at the reference's own bar: ADMM (rho 10,
    x0 = y, 150 iterations, CG tol 1e-4 / maxiter 25) with lam * L21 (L1)
    of FiniteDifference, and AcceleratedPGM (L0 1e3, 400 iterations) with
    the approximate Haar TV prox, agree to SNR > 50 dB; and the TV value
    equals lam * norm(FD(x)) to rel_res < 1e-6."""
    _, y = S.tv_denoise_problem(2)
    lam = 5e-2
    x_admm, x_apgm, _ = S.tv_denoise_pair(y, lam, isotropic, circular, 150, 1e3, 400)
    assert _snr(x_admm, x_apgm) > 50
    g = S.fd(x_admm, circular)
    gv = lam * (S.l21_norm(g) if isotropic else S.l1_norm(g))
    hv = lam * S.tv_norm(x_admm, isotropic, circular)
    assert abs(gv - hv) / max(abs(gv), abs(hv)) < 1e-6


@pytest.mark.parametrize("circular", [True, False])
def test_tv_admm_apgm_reference_pin_1d(circular):
    """``test_tvnorm.py:26-58``: 1D anisotropic TV, ADMM 50 iterations
    against APGM (L0 5e2, 100 iterations), SNR > 50 dB."""
    _, y = S.tv_denoise_problem(1)
    x_admm, x_apgm, _ = S.tv_denoise_pair(y, 5e-2, False, circular, 50, 5e2, 100)
    assert _snr(x_admm, x_apgm) > 50


@pytest.mark.parametrize("factor", [1 / np.sqrt(2.0), 1.1])
def test_tv_pin_catches_a_threshold_slip(factor):
    """The 50 dB pin is sharp: a Haar prox whose threshold misses its
    sqrt(2) factor (36 dB), or is off by 10 % (43 dB), fails it."""
    _, y = S.tv_denoise_problem(2)
    x_admm, x_apgm, _ = S.tv_denoise_pair(y, 5e-2, True, False, 150, 1e3, 400, thr_factor=factor)
    assert _snr(x_admm, x_apgm) < 48


def test_cg_direct_solve():
    """test_solver.py:28-114 pattern: CG on an SPD system vs direct solve."""
    rng = np.random.default_rng(5)
    n = 40
    Am = rng.standard_normal((n, n))
    Am = (Am.T @ Am + n * np.eye(n)).astype(np.float32)
    b = rng.standard_normal(n).astype(np.float32)
    x, info = S.cg(lambda v: (Am @ v).astype(np.float32), b, np.zeros(n, np.float32),
                   tol=1e-6, maxiter=200)
    xd = np.linalg.solve(Am.astype(np.float64), b)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-5
    assert info["rel_res"] <= 1e-6 * 1.01


def test_admm_quadratic_closed_form():
    """test_admm.py:111-258 pattern: with lam = 0 the minimiser of
    0.5||y - x||^2 is y (ADMM contracts at rate 8 rho / (1 + 8 rho))."""
    y = np.random.default_rng(6).standard_normal((10, 12)).astype(np.float32)
    I = lambda v: v
    x, hist = S.admm_tv(y, I, I, 0.0, 1e-2, np.zeros_like(y), 20, cg_tol=1e-7, cg_maxiter=100)
    np.testing.assert_allclose(x, y, atol=1e-5)


def test_circconv_direct_and_single_pixel():
    """test_circconv.py:35-56 (single pixel = flipped kernel inner product)
    and a direct circular sum, both in float64."""
    rng = np.random.default_rng(7)
    x = rng.standard_normal((12, 8)).astype(np.float32)
    h = rng.standard_normal((3, 2)).astype(np.float32)
    hd = S.circconv_hdft(h, x.shape)
    Ax = S.circconv(x, hd)
    pix = np.sum(np.flip(h, (0, 1)) * x[-3:, -2:])
    np.testing.assert_allclose(Ax[-1, -1], pix, rtol=1e-5)
    np.testing.assert_allclose(Ax, S.circconv_direct(x, h), atol=1e-5)
    c = (1, 1)
    hdc = S.circconv_hdft(h, x.shape, center=c)
    np.testing.assert_allclose(S.circconv(x, hdc), S.circconv_direct(x, h, center=c), atol=1e-5)
    yv = rng.standard_normal(x.shape).astype(np.float32)
    a = np.vdot(S.circconv(x, hdc).astype(np.float64), yv)
    b = np.vdot(x.astype(np.float64), S.circconv_adj(yv, hdc))
    assert abs(a - b) < 1e-5 * abs(a)


def test_apgm_fixed_point():
    """x* = prox(x* - grad f(x*)/L, 1/L) at convergence (test_pgm.py pattern)."""
    rng = np.random.default_rng(8)
    y = rng.standard_normal(20).astype(np.float32)
    lam = 0.5
    grad = lambda v: (v - y).astype(np.float32)
    prox = lambda v, s: S.l1_prox(v, lam * s)
    x, res = S.apgm(grad, prox, 1.0, np.zeros_like(y), 50)
    np.testing.assert_allclose(x, S.l1_prox(y, lam), atol=1e-6)


@pytest.mark.parametrize("shape", [(9, 7), (5, 6, 4), (11,)])
@pytest.mark.parametrize("iso", [True, False])
@pytest.mark.parametrize("circular", [False, True])
def test_tv_prox_weighted_pins(shape, iso, circular):
    """The weighted prox is the reference prox at w = 1 (float32, bitwise) and
    a constant weight c acts as lam * c (float64)."""
    rng = np.random.default_rng(0)
    v = rng.standard_normal(shape).astype(np.float32)
    ref = S.tv_prox(v, 0.13, iso, circular)
    np.testing.assert_array_equal(S.tv_prox_weighted(v, 0.13, None, iso, circular), ref)
    np.testing.assert_array_equal(
        S.tv_prox_weighted(v, 0.13, np.ones(shape, np.float32), iso, circular), ref)
    a = S.tv_prox_weighted(v, 0.13, np.full(shape, 1.7), iso, circular, dtype=np.float64)
    b = S.tv_prox_weighted(v, 0.13 * 1.7, None, iso, circular, dtype=np.float64)
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12)
    # the float64 twin agrees with the float32 restatement to float32 rounding
    c = S.tv_prox_weighted(v, 0.13, None, iso, circular, dtype=np.float64)
    np.testing.assert_allclose(c, ref, rtol=1e-5, atol=1e-6)


def test_unrolled_apgm_oracle_matches_apgm():
    """unrolled_apgm_tv at w = 1 is apgm with the reference TV prox (the two
    restatements round the momentum and step scalars differently)."""
    rng = np.random.default_rng(1)
    A = rng.standard_normal((40, 64)).astype(np.float32) / 8
    AT = A.T.copy()
    y = rng.standard_normal(40).astype(np.float32)
    fA = lambda v: (A @ v.ravel()).astype(np.float32)
    fAT = lambda r: (AT @ r).reshape(8, 8).astype(np.float32)
    L = float(np.linalg.norm(A.astype(np.float64), 2) ** 2) * 1.01
    lam = 0.05
    x = S.unrolled_apgm_tv(y, fA, fAT, L, lam, None, 12, np.zeros((8, 8), np.float32))
    grad = lambda v: fAT((fA(v) - y).astype(np.float32))
    xr, _ = S.apgm(grad, lambda v, s: S.tv_prox(v, lam * s), L, np.zeros((8, 8), np.float32), 12)
    np.testing.assert_allclose(x, xr, rtol=1e-4, atol=1e-5)


def test_padmm_oracle_quadratic_closed_form():
    """Proximal ADMM (oracle) on argmin 1/2 ||y - M x||^2 + (lam/2) ||D x||^2
    written as f = 0, g = (SquaredL2Loss(y), (lam/2) SquaredL2Norm),
    A = (M; D), B = -I, converges to the solution of
    (M^T M + lam D^T D) x = M^T y (the setting of ``test_padmm.py:180-200``)."""
    rng = np.random.default_rng(12345)
    M = rng.standard_normal((12, 8))
    D = rng.standard_normal((10, 8))
    y = rng.standard_normal(12)
    lam = 1.0
    A = lambda x: [M @ x, D @ x]
    AT = lambda w: M.T @ w[0] + D.T @ w[1]
    f_prox = lambda v, l: v
    g_prox = lambda v, l: [S.sql2_identity_prox(v[0], l, y, dtype=np.float64),
                           v[1] / (1.0 + l * lam)]
    mu = 1.01 * np.linalg.norm(np.vstack([M, D]), 2) ** 2
    x, z, u = S.padmm(f_prox, g_prox, A, AT, np.zeros(8), [np.zeros(12), np.zeros(10)],
                      [np.zeros(12), np.zeros(10)], rho=1.0, mu=mu, nu=1.0, maxiter=4000,
                      dtype=np.float64)
    xs = np.linalg.solve(M.T @ M + lam * D.T @ D, M.T @ y)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-6
    # a sign slip in one block of A (adjoint left unchanged) breaks it
    bad = lambda w: [-(M @ w), D @ w]
    xb, _, _ = S.padmm(f_prox, g_prox, bad, AT, np.zeros(8), [np.zeros(12), np.zeros(10)],
                       [np.zeros(12), np.zeros(10)], rho=1.0, mu=mu, nu=1.0, maxiter=400,
                       dtype=np.float64)
    assert np.linalg.norm(xb - xs) / np.linalg.norm(xs) > 1e-2


def test_sql2_identity_prox_optimality():
    """prox of 1/2 ||y - x||^2 at v minimises 1/2||x - v||^2 + lam/2 ||y - x||^2."""
    rng = np.random.default_rng(3)
    y, v = rng.standard_normal(50), rng.standard_normal(50)
    lam = 0.7
    p = S.sql2_identity_prox(v, lam, y, dtype=np.float64)
    g = (p - v) + lam * (p - y)
    assert np.abs(g).max() < 1e-12
