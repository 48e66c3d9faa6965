"""Pins for the X-ray oracle (CPU).  The oracle is checked against closed
forms, a brute-force float64 overlap model, adjointness and the
reference's known-answer arrays before it is trusted as the checker."""

import json
import os

import numpy as np
import pytest

from oracle import xray as X

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def geom(shape, angles, **kw):
    return X.xr2d_defaults(shape, angles, **kw)


def proj(g, im):
    return X.project(im, g["x0"], g["dx"], g["y0"], g["ny"], g["angles"])


def bproj(g, y):
    return X.back_project(y, g["x0"], g["dx"], g["nx"], g["y0"], g["angles"])


def test_angle_zero_sums_rows():
    """_xray2d.py:66-70: angle 0 sums rows; unit pixels align with bins."""
    rng = np.random.default_rng(0)
    im = rng.random((7, 5)).astype(np.float32)
    g = geom(im.shape, [0.0], dx=1.0, det_count=7)
    y = proj(g, im)
    np.testing.assert_array_equal(y[0], im.sum(axis=1, dtype=np.float32))
    inds, w = X.calc_weights(g["x0"], g["dx"], g["nx"], g["angles"], g["y0"])
    np.testing.assert_array_equal(inds[0], np.repeat(np.arange(7)[:, None], 5, axis=1))
    np.testing.assert_array_equal(w[0], np.ones((7, 5), np.float32))


def test_angle_half_pi_sums_columns():
    rng = np.random.default_rng(1)
    im = rng.random((6, 6)).astype(np.float32)
    g = geom(im.shape, [np.pi / 2], dx=1.0, det_count=6)
    y = proj(g, im)
    np.testing.assert_allclose(y[0], im.sum(axis=0), rtol=1e-5)


@pytest.mark.parametrize(
    "shape,nang,kw",
    [((12, 13), 10, {}), ((16, 16), 3, {"det_count": 11, "dx": 1 / np.sqrt(2)}),
     ((9, 7), 17, {"dx": 0.5}), ((8, 8), 5, {"det_count": 6})],
)
def test_dense_overlap_matrix(shape, nang, kw):
    """Brute-force float64 interval-overlap matrix equals the oracle."""
    ang = np.linspace(0, np.pi, nang, endpoint=False)
    g = geom(shape, ang, **kw)
    H = X.dense_matrix_overlap(g["x0"], g["dx"], g["nx"], g["angles"], g["y0"], g["ny"])
    rng = np.random.default_rng(2)
    x = rng.random(shape).astype(np.float32)
    y = rng.random((nang, g["ny"])).astype(np.float32)
    ys = proj(g, x)
    np.testing.assert_allclose(ys.ravel(), H @ x.ravel(), rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(bproj(g, y).ravel(), H.T @ y.ravel(), rtol=1e-5, atol=1e-5)


def test_weights_invariants():
    ang = np.linspace(0, np.pi, 37, endpoint=False)
    g = geom((20, 23), ang)
    inds, w = X.calc_weights(g["x0"], g["dx"], g["nx"], g["angles"], g["y0"])
    assert w.dtype == np.float32 and inds.dtype == np.int32
    assert np.all(w > 0) and np.all(w <= 1)
    assert np.all((w + (np.float32(1) - w)) == 1)


def test_mass_conservation():
    """No pixel-area factor: sum over bins = sum of image when all in range."""
    ang = np.linspace(0, np.pi, 11, endpoint=False)
    im = np.random.default_rng(3).random((10, 10)).astype(np.float32)
    g = geom(im.shape, ang)
    y = proj(g, im)
    np.testing.assert_allclose(y.sum(axis=1), np.full(11, im.sum()), rtol=1e-5)


def test_adjointness_and_f64_twin():
    ang = np.linspace(0, np.pi, 10, endpoint=False)
    g = geom((12, 13), ang)
    rng = np.random.default_rng(4)
    x = rng.random((12, 13)).astype(np.float32)
    y = rng.random((10, g["ny"])).astype(np.float32)
    a = float(np.sum(proj(g, x).astype(np.float64) * y))
    b = float(np.sum(bproj(g, y).astype(np.float64) * x))
    assert abs(a - b) / max(abs(a), abs(b)) < 1e-6
    y64 = X.project_f64(x, g["x0"], g["dx"], g["y0"], g["ny"], g["angles"])
    assert np.linalg.norm(y64 - proj(g, x)) / np.linalg.norm(y64) < 1e-5


@pytest.mark.parametrize("dx,factor", [(0.5, 1.0), (1 / np.sqrt(2), 1.02 / np.sqrt(2.0))])
def test_fbp_psnr_pin(dx, factor):
    """test_xray_2d.py:88-102: 256^2 square, 360 views, PSNR > 28 dB."""
    N = 256
    x = np.zeros((N, N), np.float32)
    x[N // 4 : -N // 4, N // 4 : -N // 4] = 1
    ang = np.linspace(0, np.pi, 360, endpoint=False)
    g = geom(x.shape, ang, det_count=int(factor * N), dx=dx)
    y = proj(g, x)
    xr = X.fbp(y, g["x0"], g["dx"], g["nx"], g["y0"], g["angles"])
    mse = np.mean((xr.astype(np.float64) - x) ** 2)
    psnr = 10 * np.log10(1.0 / mse)
    assert psnr > 28


def test_xray3d_reference_kat():
    kat = json.load(open(os.path.join(GOLD, "reference_kat.json")))["xray3d_scaling"]
    x = np.zeros(kat["input_shape"], np.float32)
    s = kat["ones_slice"]
    x[s[0][0] : s[0][1], s[1][0] : s[1][1], s[2][0] : s[2][1]] = 1
    for case in kat["cases"]:
        M = X.matrices_from_euler_angles(x.shape, x.shape[:2], kat["seq"], kat["angles"],
                                         voxel_spacing=case["voxel_spacing"])
        np.testing.assert_allclose(X.project_3d(x, M, x.shape[:2]), np.array(case["truth"]),
                                   atol=1e-6)


def test_xray3d_dot_test():
    """test_xray_3d.py:9-26 (matched adjoint, eps 1e-5)."""
    N = 16
    dc = int(N * 1.05 / np.sqrt(2.0))
    M = X.matrices_from_euler_angles((N, N, N), (dc, dc), "X",
                                     np.linspace(0, np.pi, 3, endpoint=False)[:, None])
    r = np.random.default_rng(0)
    v = r.random((N, N, N)).astype(np.float32)
    y = r.random((3, dc, dc)).astype(np.float32)
    a = np.vdot(X.project_3d(v, M, (dc, dc)).astype(np.float64), y)
    b = np.vdot(v.astype(np.float64), X.back_project_3d(y, M, (N, N, N)))
    assert abs(a - b) / max(abs(a), abs(b)) < 1e-5


def test_xray3d_single_axis_slices():
    """Rotation about volume axis 1: each slice vol[:, j, :] maps only to
    proj[:, :, j] (closed form of _xray3d.py:214-230 with t1 = 0)."""
    shape, det = (6, 5, 7), (10, 5)
    M = X.matrices_from_euler_angles(shape, det, "Y", np.linspace(0, np.pi, 4, endpoint=False))
    v = np.zeros(shape, np.float32)
    v[:, 2, :] = np.random.default_rng(5).random((6, 7))
    p = X.project_3d(v, M, det)
    assert np.abs(p[:, :, [0, 1, 3, 4]]).max() == 0
    assert np.abs(p[:, :, 2]).max() > 0


def _flip_rates(N, A, step=1):
    """Fraction of (pixel, view) pairs whose detector bin changes when
    (a) Px = (Pxmin + P0 i) + P1 j is evaluated with two fused multiply-adds
    (one rounding each) and (b) cos/sin are float32 library results instead
    of the correctly rounded float32 of the float64 values."""
    g = X.xr2d_defaults((N, N), np.linspace(0, np.pi, A, endpoint=False))
    views = np.arange(0, A, step)
    th = g["angles"][views]
    Pxmin, P0, P1, _ = X.view_params(g["x0"], g["dx"], th, g["y0"])
    i = np.arange(N, dtype=np.float32).reshape(1, -1, 1)
    j = np.arange(N, dtype=np.float32).reshape(1, 1, -1)
    e = lambda v: v.reshape(-1, 1, 1)
    ref = np.floor((e(Pxmin) + e(P0) * i) + e(P1) * j)
    f64 = np.float64
    t = (e(Pxmin).astype(f64) + e(P0).astype(f64) * i.astype(f64)).astype(np.float32)
    fma = np.floor((t.astype(f64) + e(P1).astype(f64) * j.astype(f64)).astype(np.float32))
    u0, u1 = np.cos(th), np.sin(th)  # float32 in, float32 library result out
    x00, x01 = np.float32(g["x0"][0]), np.float32(g["x0"][1])
    dx0, dx1 = np.float32(g["dx"][0]), np.float32(g["dx"][1])
    Px0 = (x00 * u0 + x01 * u1) - g["y0"]
    q0, q1 = dx0 * u0, dx1 * u1
    c = np.stack([Px0, Px0 + q0, Px0 + q1, (Px0 + q0) + q1])
    lib = np.floor((e(c.min(axis=0)) + e(q0) * i) + e(q1) * j)
    n = ref.size
    return float(np.count_nonzero(fma != ref)) / n, float(np.count_nonzero(lib != ref)) / n


@pytest.mark.parametrize("N,A,step", [(256, 180, 1), (512, 720, 4)])
def test_bin_flip_rates(N, A, step):
    """Residual-risk measurement for the two XLA choices no fixture pins:
    FMA contraction of the Px expression and float32 cos/sin rounding."""
    r_fma, r_lib = _flip_rates(N, A, step)
    print(f"\n  {N}^2/{A} views: bin flips with FMA {r_fma:.2e}, with float32 cos/sin {r_lib:.2e}")
    assert r_fma < 1e-3 and r_lib < 1e-2


@pytest.mark.parametrize("N", [2, 3, 4, 5, 64, 91, 184, 363, 724])
def test_fbp_filter_window_keeps_n_bins(N):
    """The reference's slice [(N-1)//2 : -(N-1)//2] of the length 2N-1
    convolution (_xray2d.py:192-194) keeps N bins for even and odd N; the
    product's window and the oracle's filtered rows have that length."""
    from scico_amd.linop.xray._fbp import filter_window

    s, L = filter_window(N)
    assert L == N and s == (N - 1) // 2
    full = np.arange(2 * N - 1)
    assert np.array_equal(full[s : s + L], full[(N - 1) // 2 : -(N - 1) // 2])
    ang = np.linspace(0, np.pi, 3, endpoint=False)
    g = geom((8, 8), ang, det_count=N)
    y = np.ones((3, N), np.float32)
    # the oracle's FBP back projects the filtered rows with the N-bin geometry
    xr = X.fbp(y, g["x0"], g["dx"], g["nx"], g["y0"], g["angles"])
    assert xr.shape == (8, 8)
