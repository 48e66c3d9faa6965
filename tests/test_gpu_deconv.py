"""GPU parity of the FFT, CircularConvolve, TV prox / norm and APGM-TV
deconvolution against the oracle (numpy.fft and oracle.solvers)."""

import numpy as np
import pytest
import torch

from oracle import solvers as S
from scico_amd import functional, linop, loss
from scico_amd.linop import CircularConvolve, valid_adjoint
from scico_amd.optimize import AcceleratedPGM
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def gaussian_psf(n=15, sigma=2.0):
    r = np.arange(n) - n // 2
    g = np.exp(-(r[:, None] ** 2 + r[None, :] ** 2) / (2 * sigma**2))
    return (g / g.sum()).astype(np.float32)


@pytest.mark.parametrize("shape,ndims", [((16,), 1), ((12, 8), 2), ((6, 8, 12), 3),
                                         ((2, 12, 8), 2), ((1024, 1024), 2), ((64, 48), 2),
                                         ((4, 256), 1)])
def test_fftn_vs_numpy(gpu, shape, ndims):
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    A = CircularConvolve(np.ones((1,) * ndims, np.float32), shape, ndims=ndims)
    axes = tuple(range(len(shape) - ndims, len(shape)))
    for inverse in (False, True):
        got = A.fftn(dev(x), inverse=inverse).cpu().numpy()
        f = np.fft.ifftn if inverse else np.fft.fftn
        ref64 = f(x.astype(np.complex128), axes=axes)
        e = rel_l2(got, ref64)
        print(f"  fftn {shape} inv={inverse}: rel-L2 vs float64 {e:.2e}", flush=True)
        assert e <= 1e-5


@pytest.mark.parametrize("x_shape,ndims,h_shape", [((12,), None, (3,)), ((12, 8), None, (3, 2)),
                                                   ((6, 8, 12), None, (3, 2, 4)),
                                                   ((2, 12, 8), 2, (3, 2))])
def test_circconv_reference_patterns(gpu, x_shape, ndims, h_shape):
    """test_circconv.py:35-70: one output pixel equals the inner product with
    the flipped kernel; adjoint dot test; direct circular sum."""
    rng = np.random.default_rng(1)
    h = rng.standard_normal(h_shape).astype(np.float32)
    x = rng.standard_normal(x_shape).astype(np.float32)
    A = CircularConvolve(h, x_shape, ndims)
    Ax = A(dev(x)).cpu().numpy()
    nd = A.ndims
    hf = np.flip(h, tuple(range(-nd, 0)))
    xi = (...,) + tuple(slice(-h.shape[a], None) for a in range(-nd, 0))
    Ai = (...,) + tuple(-1 for _ in range(nd))
    np.testing.assert_allclose(np.sum(hf * x[xi], axis=tuple(-(a + 1) for a in range(nd))),
                               Ax[Ai], rtol=1e-5, atol=1e-6)
    assert valid_adjoint(A, A.T, eps=1e-5)
    if len(x_shape) == nd:
        np.testing.assert_allclose(Ax, S.circconv_direct(x, h), atol=1e-5)


def test_circconv_center_and_from_operator(gpu):
    """test_circconv.py:151-192: h_center <=> roll; from_operator."""
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, 16).astype(np.float32)
    h = np.array([0.5, 1.0, 0.25], np.float32)
    A = CircularConvolve(h, x.shape, h_center=1)
    B = CircularConvolve(h, x.shape)
    np.testing.assert_allclose(A(dev(x)).cpu().numpy(), np.roll(B(dev(x)).cpu().numpy(), -1),
                               atol=1e-5)
    x2 = rng.standard_normal((12, 8)).astype(np.float32)
    h2 = rng.standard_normal((3, 2)).astype(np.float32)
    A2 = CircularConvolve(h2, (12, 8))
    B2 = CircularConvolve.from_operator(A2)
    np.testing.assert_allclose(A2(dev(x2)).cpu().numpy(), B2(dev(x2)).cpu().numpy(), atol=1e-5)
    with pytest.raises(ValueError):
        CircularConvolve(h2, (12, 8), h_is_dft=True, h_center=[1, 1])


@pytest.mark.parametrize("x_shape,kind", [((4, 4), "real"), ((64, 256), "real"),
                                          ((256, 64), "centre"), ((3, 16, 32), "real"),
                                          ((8, 4096), "complex"), ((2048, 8), "complex"),
                                          ((128, 128), "complex"), ((32, 64), "centre")])
def test_circconv_real_2d_path(gpu, x_shape, kind):
    """Two power-of-two axes run on half spectra (two real rows per complex
    row transform, Hermitian part of the spectrum): real(ifft2(H fft2(x)))
    for a real kernel, a kernel with a non-integer centre (a spectrum that
    is not Hermitian) and an arbitrary complex spectrum (h_is_dft), forward
    and adjoint, against the oracle (complex64 numpy FFTs) and float64."""
    rng = np.random.default_rng(11)
    F = x_shape[-2:]
    x = rng.standard_normal(x_shape).astype(np.float32)
    if kind == "complex":
        h = (rng.standard_normal(F) + 1j * rng.standard_normal(F)).astype(np.complex64)
        A = CircularConvolve(h, x_shape, ndims=2, h_is_dft=True)
        hd = h
    else:
        h = rng.standard_normal((min(3, F[0]), min(5, F[1]))).astype(np.float32)
        center = [-1.0, 2.5] if kind == "centre" else None
        A = CircularConvolve(h, x_shape, ndims=2, h_center=center)
        hd = S.circconv_hdft(h, F, center=center)
    y = A(dev(x)).cpu().numpy()
    yt = A.T(dev(x)).cpu().numpy()
    xs = x.reshape(-1, *F)
    ref = np.stack([S.circconv(v, hd) for v in xs]).reshape(x_shape)
    reft = np.stack([S.circconv_adj(v, hd) for v in xs]).reshape(x_shape)
    hd64 = hd.astype(np.complex128)
    r64 = np.stack([np.fft.ifft2(hd64 * np.fft.fft2(v.astype(np.float64))).real
                    for v in xs]).reshape(x_shape)
    e, et, e64 = rel_l2(y, ref), rel_l2(yt, reft), rel_l2(y, r64)
    print(f"  real 2D path {x_shape} {kind}: fwd {e:.2e} adj {et:.2e} (vs float64 {e64:.2e})",
          flush=True)
    assert e <= 1e-5 and et <= 1e-5 and e64 <= 1e-5
    # dot test, normalised by the norms (with an arbitrary complex spectrum
    # y^T A x is a sum of terms of random sign, far below ||y|| ||A x||)
    v = rng.standard_normal(x_shape).astype(np.float32)
    u1 = float(np.sum(v.astype(np.float64) * y))
    u2 = float(np.sum(A.T(dev(v)).cpu().numpy().astype(np.float64) * x))
    assert abs(u1 - u2) <= 1e-6 * np.linalg.norm(v) * np.linalg.norm(y.astype(np.float64))
    if kind != "complex":
        assert valid_adjoint(A, A.T, eps=1e-5)


def test_circconv_1024_psf(gpu):
    """1024^2 image, 15x15 Gaussian PSF (sigma 2) centred at (7, 7):
    forward / adjoint against the oracle (complex64 FFTs) and float64."""
    rng = np.random.default_rng(3)
    x = rng.random((1024, 1024)).astype(np.float32)
    h = gaussian_psf()
    A = CircularConvolve(h, x.shape, h_center=[7, 7])
    hd = S.circconv_hdft(h, x.shape, center=[7, 7])
    y = A(dev(x)).cpu().numpy()
    yt = A.T(dev(x)).cpu().numpy()
    e, et = rel_l2(y, S.circconv(x, hd)), rel_l2(yt, S.circconv_adj(x, hd))
    e64 = rel_l2(y, S.circconv_f64(x, h, center=[7, 7]))
    print(f"  1024^2 PSF conv: fwd {e:.2e} adj {et:.2e} (vs float64 {e64:.2e})", flush=True)
    assert e <= 1e-5 and et <= 1e-5
    np.testing.assert_allclose(A.h_dft.cpu().numpy(), hd, rtol=0, atol=2e-6)


@pytest.mark.parametrize("shape", [(33, 31), (8, 9, 10), (40,)])
@pytest.mark.parametrize("iso", [True, False])
@pytest.mark.parametrize("circular", [False, True])
def test_tv_prox_and_norm(gpu, shape, iso, circular):
    rng = np.random.default_rng(4)
    v = rng.standard_normal(shape).astype(np.float32)
    F = functional.IsotropicTVNorm(circular=circular) if iso else \
        functional.AnisotropicTVNorm(circular=circular)
    lam = 0.13
    got = (0.7 * F).prox(dev(v), lam).cpu().numpy()
    ref = S.tv_prox(v, 0.7 * lam, isotropic=iso, circular=circular)
    assert rel_l2(got, ref) <= 1e-6
    val = float(F(dev(v)))
    assert abs(val - S.tv_norm(v, isotropic=iso, circular=circular)) <= 1e-9 * abs(val)


def _deconv_problem(N, seed=0):
    from scico_amd import phantom

    xgt = phantom.ellipse_phantom(N, 12, seed=seed)
    h = gaussian_psf()
    hd = S.circconv_hdft(h, (N, N), center=[7, 7])
    y = S.circconv(xgt, hd)
    y = phantom.noisy(y, 0.01, seed=seed + 1)
    L0 = float(np.max(np.abs(hd.astype(np.complex128)) ** 2))
    return xgt, h, hd, y, L0


def test_apgm_tv_deconvolution_parity(gpu):
    """AcceleratedPGM with isotropic TV (test_tvnorm.py pattern) against the
    oracle, 20 iterations at 128^2."""
    N = 128
    xgt, h, hd, y, L0 = _deconv_problem(N)
    lam = 2e-2
    A = CircularConvolve(h, (N, N), h_center=[7, 7])
    f = loss.SquaredL2Loss(y=dev(y), A=A)
    g = lam * functional.IsotropicTVNorm()
    solver = AcceleratedPGM(f=f, g=g, L0=L0, x0=dev(np.zeros((N, N), np.float32)), maxiter=20)
    x = solver.solve().cpu().numpy()
    grad = lambda v: S.circconv_adj((S.circconv(v, hd) - y).astype(np.float32), hd)
    prox = lambda v, s: S.tv_prox(v, lam * s)
    xr, res = S.apgm(grad, prox, L0, np.zeros((N, N), np.float32), 20)
    e = rel_l2(x, xr)
    hist = solver.history(transpose=True)
    print(f"  APGM-TV 128^2 x20: rel-L2 vs oracle {e:.2e}", flush=True)
    assert e <= 1e-5
    np.testing.assert_allclose(hist.Residual, res, rtol=1e-4)
    assert all(np.isfinite(hist.Objective))


@pytest.mark.parametrize("extra_axis", (False, True))
@pytest.mark.parametrize("center", (None, [-1.0, 2.5]))
def test_circular_convolve_solver(gpu, extra_axis, center):
    """DFT-domain ADMM x-step (``test_admm.py:416-468``): 50 ADMM
    iterations with CircularConvolveSolver agree with the CG x-step to
    atol 1e-4 and mse < 1e-9, and (centre None) the x-step spectrum equals
    the closed form rho * sum_k (2 - 2 cos(2 pi f_k)) + |H(f)|^2 (circular
    FD gram plus 2 * scale * A^T A with scale 1/2)."""
    from scico_amd import metric
    from scico_amd.optimize import ADMM, CircularConvolveSolver, LinearSubproblemSolver
    Nx = 8
    x = np.pad(np.ones((Nx, Nx), np.float32), Nx)
    psf = np.ones((3, 3), np.float32) / 9.0
    if extra_axis:
        x, psf = x[np.newaxis], psf[np.newaxis]
    A = linop.CircularConvolve(h=psf, input_shape=x.shape, ndims=2, input_dtype=np.float32,
                               h_center=center)
    y = A(dev(x))
    lam, rho, maxiter = 1e-2, 1e-1, 50

    def run(solver):
        s = ADMM(f=loss.SquaredL2Loss(y=y, A=A), g_list=[lam * functional.L1Norm()],
                 C_list=[linop.FiniteDifference(input_shape=x.shape, circular=True)],
                 rho_list=[rho], maxiter=maxiter, itstat_options={"display": False},
                 x0=A.adj(y), subproblem_solver=solver)
        return s, s.solve()

    _, x_lin = run(LinearSubproblemSolver())
    s_dft, x_dft = run(CircularConvolveSolver())
    assert s_dft.subproblem_solver.A_lhs.ndims == 2
    np.testing.assert_allclose(x_dft.cpu().numpy(), x_lin.cpu().numpy(), atol=1e-4, rtol=0)
    assert metric.mse(x_lin, x_dft) < 1e-9

    if center is not None:
        # A half-sample shift zeroes the Nyquist bins of A, so A^T A is not
        # exactly shift invariant there; the closed form below is for the
        # plain (centre None) convolution.
        return
    n = x.shape[-1]
    f = np.fft.fftfreq(n)
    fd = (2 - 2 * np.cos(2 * np.pi * f))[:, None] + (2 - 2 * np.cos(2 * np.pi * f))[None, :]
    hp = np.zeros((n, n))
    hp[:3, :3] = psf.reshape(3, 3)
    H2 = np.abs(np.fft.fft2(hp)) ** 2
    ref = rho * fd + H2
    got = s_dft.subproblem_solver.lhs_dft
    assert np.abs(got.imag).max() < 1e-5
    np.testing.assert_allclose(got.real, ref, rtol=0, atol=2e-6)


@pytest.mark.parametrize("kind", ["real", "complex"])
def test_squared_l2_grad_circconv(gpu, kind):
    """SquaredL2Loss.grad with a CircularConvolve (evaluated as G x - A^T y
    with G the convolution by |Hs|^2, Hs the Hermitian part of the
    spectrum) equals A^T (A x - y), also for an arbitrary complex spectrum;
    a new operator assigned to the loss rebuilds the cached G."""
    rng = np.random.default_rng(5)
    F = (64, 32)
    if kind == "complex":
        h = (rng.standard_normal(F) + 1j * rng.standard_normal(F)).astype(np.complex64)
        A = CircularConvolve(h, F, h_is_dft=True)
    else:
        A = CircularConvolve(rng.standard_normal((3, 4)).astype(np.float32), F)
    y = dev(rng.standard_normal(F).astype(np.float32))
    x = dev(rng.standard_normal(F).astype(np.float32))
    f = loss.SquaredL2Loss(y=y, A=A)
    ref = A.adj(A(x) - y)
    assert rel_l2(f.grad(x).cpu().numpy(), ref.cpu().numpy()) <= 1e-5
    B = CircularConvolve(rng.standard_normal((2, 2)).astype(np.float32), F)
    f.A = B
    refB = B.adj(B(x) - y)
    assert rel_l2(f.grad(x).cpu().numpy(), refB.cpu().numpy()) <= 1e-5


@pytest.mark.parametrize("fused", [True, False])
def test_apgm_deconvolution_config_length(gpu, fused):
    """The GPU-scale configuration's deconvolution at full length: 200
    AcceleratedPGM-TV iterations at 1024^2 (15 x 15 Gaussian PSF, 1 % noise,
    lam 2e-2) against the oracle's run (``tests/golden/apgm_deconv_1024.npz``,
    written by ``tests/golden/make_deconv_fixture.py``).  APGM does not
    amplify rounding the way ten-step CG on the CT normal equations does:
    the other valid float32 runs (float64 FFTs, FFT-rounding-sized noise)
    end 1.4e-7 (rel-L2) from the oracle, so the reconstruction is held to
    the parity bar of 1e-5 (rel-L2) directly; the objective and residual statistics
    to 1e-5 of their value (residual: of its largest value so far).  Both
    the fused iteration (one convolution per iteration: the gradient by
    linearity, the objective by Parseval) and the generic one."""
    import json
    import os
    import sys

    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    if gold not in sys.path:
        sys.path.insert(0, gold)
    import make_deconv_fixture as D

    fx = np.load(os.path.join(gold, "apgm_deconv_1024.npz"), allow_pickle=False)
    meta = json.loads(str(fx["meta"]))
    xgt, psf, hd, y, _ = D.problem()
    assert D.sha256(y) == meta["y_sha256"], "regenerated blurred image differs from the fixture's"
    N = D.N
    A = CircularConvolve(psf, (N, N), h_center=[7, 7])
    f = loss.SquaredL2Loss(y=dev(y), A=A)
    g = D.LAM * functional.IsotropicTVNorm()
    from scico_amd.optimize import _pgm

    old = _pgm._FusedDeconv.enabled
    _pgm._FusedDeconv.enabled = fused
    try:
        solver = AcceleratedPGM(f=f, g=g, L0=meta["L0"], x0=dev(np.zeros((N, N), np.float32)),
                                maxiter=D.ITERS)
    finally:
        _pgm._FusedDeconv.enabled = old
    assert (solver._fused is not None) == fused
    x = solver.solve().cpu().numpy()
    h = solver.history(transpose=True)
    hr = fx["hist"]
    d_obj = np.abs(np.asarray(h.Objective, np.float64) - hr[:, 0]) / np.abs(hr[:, 0])
    d_res = np.abs(np.asarray(h.Residual, np.float64) - hr[:, 1]) / \
        np.maximum.accumulate(np.abs(hr[:, 1]))
    e_crop, e_pool = rel_l2(D.crop(x), fx["x_crop"]), rel_l2(D.pool(x), fx["x_pool"])
    snr = D.snr(x, xgt)
    fin = meta["floor_final"]
    print(f"  APGM-TV deconvolution 1024^2 x{D.ITERS} ({'fused' if fused else 'generic'}): central crop rel-L2 {e_crop:.2e} (floor "
          f"{fin['x_crop_rel_l2']:.2e}), pooled {e_pool:.2e} (floor {fin['x_pool_rel_l2']:.2e}); "
          f"objective dev max {d_obj.max():.2e}, residual dev max {d_res.max():.2e}; SNR "
          f"{snr:.4f} vs {meta['snr_db']:.4f} dB", flush=True)
    assert e_crop <= 1e-5 and e_pool <= 1e-5
    assert d_obj.max() <= 1e-5 and d_res.max() <= 1e-5
    assert abs(snr - meta["snr_db"]) <= 1e-3


def _deconv_1024(fused, maxiter, x0=None):
    from scico_amd.optimize import _pgm

    xgt, h, hd, y, L0 = _deconv_problem(1024)
    A = CircularConvolve(h, (1024, 1024), h_center=[7, 7])
    f = loss.SquaredL2Loss(y=dev(y), A=A)
    g = 2e-2 * functional.AnisotropicTVNorm()
    old = _pgm._FusedDeconv.enabled
    _pgm._FusedDeconv.enabled = fused
    try:
        x0 = dev(np.zeros((1024, 1024), np.float32)) if x0 is None else x0
        s = AcceleratedPGM(f=f, g=g, L0=L0, x0=x0, maxiter=maxiter)
    finally:
        _pgm._FusedDeconv.enabled = old
    assert (s._fused is not None) == fused
    return s


def test_fused_apgm_first_step_and_state(gpu):
    """The fused deconvolution iteration (anisotropic TV here): its first
    step is the generic one bit for bit (the extrapolation coefficient is 0
    there, so G v is the convolution itself), the statistics agree to
    float32 resolution; after 30 steps the iterates agree to 1e-6; a state
    saved after 10 steps and restored into a new solver continues to the
    same iterate (1e-6); 4 steps captured in a HIP graph and replayed equal
    4 eager steps bit for bit."""
    import tempfile

    sf, sg = _deconv_1024(True, 1), _deconv_1024(False, 1)
    sf.solve()
    sg.solve()
    assert torch.equal(sf.x, sg.x) and torch.equal(sf.v, sg.v)
    hf, hg = sf.history(transpose=True), sg.history(transpose=True)
    assert abs(float(hf.Residual[0]) - float(hg.Residual[0])) <= 1e-12 * float(hg.Residual[0])
    assert abs(float(hf.Objective[0]) - float(hg.Objective[0])) <= 1e-6 * float(hg.Objective[0])
    sf, sg = _deconv_1024(True, 30), _deconv_1024(False, 30)
    xf, xg = sf.solve().cpu().numpy(), sg.solve().cpu().numpy()
    e30 = rel_l2(xf, xg)
    # save after 10, restore, continue 20
    s1 = _deconv_1024(True, 10)
    s1.solve()
    with tempfile.TemporaryDirectory() as d:
        path = d + "/state.npz"
        s1.save_state(path)
        s2 = _deconv_1024(True, 20)
        s2.load_state(path)
    s2.t = s1.t
    xr = s2.solve().cpu().numpy()
    e_restore = rel_l2(xr, xf)
    # graph capture of 4 steps against 4 eager steps
    se, sc = _deconv_1024(True, 4), _deconv_1024(True, 4)
    for _ in range(4):
        se.step()
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        for _ in range(4):
            sc.step()
    torch.cuda.synchronize()
    graph.replay()
    torch.cuda.synchronize()
    same = torch.equal(sc.x, se.x) and torch.equal(sc._fused.hist[:4], se._fused.hist[:4])
    print(f"  fused APGM 1024^2: 30 steps vs generic rel-L2 {e30:.2e}; restored run "
          f"{e_restore:.2e}; graph replay bitwise {same}", flush=True)
    assert e30 <= 1e-6 and e_restore <= 1e-6 and same


@pytest.mark.parametrize("iso,circular", [(True, True), (False, True), (True, False)])
def test_fused_apgm_tv_variants(gpu, iso, circular):
    """The fused deconvolution iteration with isotropic / anisotropic and
    circular / non-circular TV against the generic iteration: the first
    step bit for bit, 10 steps to 1e-6, statistics to 1e-6."""
    from scico_amd.optimize import _pgm

    xgt, h, hd, y, L0 = _deconv_problem(1024)
    A = CircularConvolve(h, (1024, 1024), h_center=[7, 7])
    f = loss.SquaredL2Loss(y=dev(y), A=A)
    tv = functional.IsotropicTVNorm(circular=circular) if iso else \
        functional.AnisotropicTVNorm(circular=circular)
    g = 2e-2 * tv
    sols = {}
    for fused in (True, False):
        old = _pgm._FusedDeconv.enabled
        _pgm._FusedDeconv.enabled = fused
        try:
            s1 = AcceleratedPGM(f=f, g=g, L0=L0, x0=dev(np.zeros((1024, 1024), np.float32)),
                                maxiter=1)
            s10 = AcceleratedPGM(f=f, g=g, L0=L0, x0=dev(np.zeros((1024, 1024), np.float32)),
                                 maxiter=10)
        finally:
            _pgm._FusedDeconv.enabled = old
        assert (s1._fused is not None) == fused
        s1.solve()
        x10 = s10.solve().cpu().numpy()
        sols[fused] = (s1.x.clone(), x10, s10.history(transpose=True))
    assert torch.equal(sols[True][0], sols[False][0])
    e = rel_l2(sols[True][1], sols[False][1])
    hf, hg = sols[True][2], sols[False][2]
    d_obj = max(abs(float(a) - float(b)) / abs(float(b)) for a, b in zip(hf.Objective, hg.Objective))
    d_res = max(abs(float(a) - float(b)) / float(hg.Residual[0])
                for a, b in zip(hf.Residual, hg.Residual))
    print(f"  fused APGM iso={iso} circ={circular}: 10 steps rel-L2 {e:.2e}, objective "
          f"{d_obj:.2e}, residual {d_res:.2e}", flush=True)
    assert e <= 1e-6 and d_obj <= 1e-6 and d_res <= 1e-6
