"""GPU parity of the 2D projector against the oracle (through the C ABI).

Bar: detector bins and weights bit-exact; projections and back
projections within rel-L2 1e-5 of the float32 oracle (the error against
the float64 oracle is printed beside it); dot test <= 1e-5; bitwise
determinism.  Full-size cases compare a subset of views (views are
independent) plus size-independent properties.
"""

import numpy as np
import pytest
import torch

from oracle import xray as X
from scico_amd import phantom
from scico_amd.linop import valid_adjoint
from scico_amd.linop.xray import XRayTransform2D
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

CASES = [
    # (shape, n_angles, kwargs)
    ((12, 13), 10, {}),
    ((16, 16), 3, {"det_count": int(16 * 1.05 / np.sqrt(2.0)), "dx": 1 / np.sqrt(2)}),
    ((9, 7), 17, {"dx": 0.5}),
    ((33, 65), 31, {"det_count": 20}),  # many bins out of range
    ((64, 48), 45, {"det_count": 90, "y0": -40.0}),
    ((256, 256), 180, {"det_count": 363}),  # 256^2, 180 views, 363 bins
    ((40, 24), 1100, {}),  # two views per forward workgroup
    ((20, 20), 2100, {"dx": 0.9}),  # four views per forward workgroup
    ((30, 30), 7, {"det_count": 1100}),  # wide detector: one chain per wave
    ((16, 16), 5, {"dx": 1.8, "det_count": 60}),  # one pixel per adjoint thread
    ((24, 20), 13, {"x0": (-7.0, -15.0), "dx": (0.8, 1.25)}),  # offset, anisotropic
]


def _op(shape, nang, kw):
    ang = np.linspace(0, np.pi, nang, endpoint=False)
    H = XRayTransform2D(shape, ang, **kw)
    g = X.xr2d_defaults(shape, ang, x0=kw.get("x0"), dx=kw.get("dx"), y0=kw.get("y0"),
                        det_count=kw.get("det_count"))
    return H, g


@pytest.mark.parametrize("shape,nang,kw", CASES)
def test_view_table_bit_exact(gpu, shape, nang, kw):
    H, g = _op(shape, nang, kw)
    vt = H.view_table()
    ref = np.stack(X.view_params(g["x0"], g["dx"], g["angles"], g["y0"]), axis=1)
    np.testing.assert_array_equal(vt, ref)


@pytest.mark.parametrize("shape,nang,kw", CASES)
def test_weights_bit_exact(gpu, shape, nang, kw):
    print(f"weights {shape} A={nang} {kw}", flush=True)
    H, g = _op(shape, nang, kw)
    inds, w = H.weights()
    ri, rw = X.calc_weights(g["x0"], g["dx"], g["nx"], g["angles"], g["y0"])
    np.testing.assert_array_equal(inds.cpu().numpy(), ri)
    np.testing.assert_array_equal(w.cpu().numpy(), rw)


@pytest.mark.parametrize("shape,nang,kw", CASES)
def test_forward_adjoint_parity(gpu, shape, nang, kw):
    print(f"fwd/adj {shape} A={nang} {kw}", flush=True)
    H, g = _op(shape, nang, kw)
    rng = np.random.default_rng(7)
    x = rng.random(shape).astype(np.float32)
    y = rng.standard_normal(H.output_shape).astype(np.float32)
    ys = H(x).cpu().numpy()
    bp = H.T(y).cpu().numpy()
    ref_y = X.project(x, g["x0"], g["dx"], g["y0"], g["ny"], g["angles"])
    ref_bp = X.back_project(y, g["x0"], g["dx"], g["nx"], g["y0"], g["angles"])
    e_y, e_bp = rel_l2(ys, ref_y), rel_l2(bp, ref_bp)
    y64 = X.project_f64(x, g["x0"], g["dx"], g["y0"], g["ny"], g["angles"])
    bp64 = X.back_project_f64(y, g["x0"], g["dx"], g["nx"], g["y0"], g["angles"])
    print(f"  rel-L2 vs fp32 oracle: fwd {e_y:.2e} adj {e_bp:.2e}; "
          f"vs fp64: fwd {rel_l2(ys, y64):.2e} adj {rel_l2(bp, bp64):.2e}", flush=True)
    assert e_y <= 1e-5 and e_bp <= 1e-5


@pytest.mark.parametrize("shape,nang,kw", CASES[:3] + CASES[-1:])
def test_dot_test(gpu, shape, nang, kw):
    H, _ = _op(shape, nang, kw)
    assert valid_adjoint(H, H.T, eps=1e-5)


def test_determinism(gpu):
    H = XRayTransform2D((256, 256), phantom.angles(180), det_count=363)
    x = torch.as_tensor(phantom.ellipse_phantom(256, 10, seed=0), device="cuda")
    y1, y2 = H(x), H(x)
    b1, b2 = H.T(y1), H.T(y1)
    assert torch.equal(y1, y2) and torch.equal(b1, b2)


@pytest.mark.parametrize("shape,nang,dc", [((256, 256), 90, 363), ((300, 260), 75, 400)])
def test_forward_split(gpu, shape, nang, dc):
    """Forward line split: every part count gives the oracle's projection
    (rel-L2 1e-5) and the unsplit result up to float32 rounding (rel-L2
    1e-6), deterministically; the residual partials of the split forward
    sum to the oracle's residual; the automatic choice is 4 parts from
    256 x 256 pixels up with fewer than two forward groups per CU, else 1."""
    import ctypes

    from scico_amd import _lib

    ang = phantom.angles(nang)
    H = XRayTransform2D(shape, ang, det_count=dc)
    assert H.set_forward_split(0) == 4
    assert XRayTransform2D((96, 80), ang).set_forward_split(0) == 1
    assert XRayTransform2D((512, 512), phantom.angles(720), det_count=724).set_forward_split(0) == 1
    rng = np.random.default_rng(5)
    x = rng.random(shape).astype(np.float32)
    y = rng.standard_normal((nang, dc)).astype(np.float32)
    g = X.xr2d_defaults(shape, ang, det_count=dc)
    ref = X.project(x, g["x0"], g["dx"], g["y0"], g["ny"], g["angles"])
    ref_res = float(np.sum((y.astype(np.float64) - ref) ** 2))
    xd, yd = torch.as_tensor(x, device="cuda"), torch.as_tensor(y, device="cuda")
    out = (ctypes.c_int32 * 4)()
    _lib.call("scxr_xr2d_info", H.plan, out)
    n_fwd = int(out[0])
    assert n_fwd > 0
    base = None
    for parts in (1, 2, 3, 4, 8):
        assert H.set_forward_split(parts) == parts
        s1, s2 = H(xd), H(xd)
        assert torch.equal(s1, s2), parts
        e = rel_l2(s1.cpu().numpy(), ref)
        assert e <= 1e-5, (parts, e)
        if base is None:
            base = s1.clone()
        else:
            assert rel_l2(s1.cpu().numpy(), base.cpu().numpy()) <= 1e-6, parts
        ws = torch.empty(n_fwd, dtype=torch.float64, device="cuda")
        sr = torch.empty_like(s1)
        _lib.call("scxr_xr2d_forward_res", H.plan, xd.data_ptr(), sr.data_ptr(), yd.data_ptr(),
                  ws.data_ptr(), None, None)
        assert torch.equal(sr, s1), parts
        own = float(((yd - sr).double() ** 2).sum())  # float32 differences, as the kernel
        assert abs(float(ws.sum()) - own) <= 1e-12 * own, parts
        assert abs(float(ws.sum()) - ref_res) <= 1e-5 * ref_res, parts
    with pytest.raises(Exception):
        H.set_forward_split(9)


def test_full_size_512_720(gpu):
    """512^2, 720 views, 724 bins: view-subset parity + mass + dot test."""
    N, A, D = 512, 720, 724
    ang = phantom.angles(A)
    H = XRayTransform2D((N, N), ang, det_count=D)
    x = phantom.ellipse_phantom(N, 20, seed=0)
    xd = torch.as_tensor(x, device="cuda")
    y = H(xd)
    sub = np.arange(0, A, 37)
    g = X.xr2d_defaults((N, N), ang[sub], det_count=D)
    ref = X.project(x, g["x0"], g["dx"], g["y0"], g["ny"], g["angles"])
    e = rel_l2(y.cpu().numpy()[sub], ref)
    print(f"  512^2/720: view-subset rel-L2 {e:.2e}", flush=True)
    assert e <= 1e-5
    # no pixel-area factor and every pixel in range: each view sums to sum(x)
    sums = y.double().sum(dim=1).cpu().numpy()
    np.testing.assert_allclose(sums, np.full(A, x.astype(np.float64).sum()), rtol=1e-5)
    # adjoint of a sinogram supported on the subset = oracle adjoint of the subset
    ys = np.zeros((A, D), np.float32)
    ys[sub] = np.random.default_rng(3).standard_normal((len(sub), D))
    bp = H.T(ys).cpu().numpy()
    ref_bp = X.back_project(ys[sub], g["x0"], g["dx"], g["nx"], g["y0"], g["angles"])
    e = rel_l2(bp, ref_bp)
    print(f"  512^2/720: adjoint subset rel-L2 {e:.2e}", flush=True)
    assert e <= 1e-5
    assert valid_adjoint(H, H.T, eps=1e-5)


def test_shape_errors(gpu):
    H = XRayTransform2D((8, 8), phantom.angles(4))
    with pytest.raises(ValueError):
        H(np.zeros((8, 9), np.float32))
    with pytest.raises(ValueError):
        H.adj(np.zeros((4, 5), np.float32))
    with pytest.raises(ValueError):
        H.adj(np.zeros(H.output_shape, np.float64))


@pytest.mark.parametrize("N,nang,dc,dx", [(64, 90, 64, None), (64, 45, 91, None),
                                          (256, 360, 256, 0.5),
                                          (256, 360, int(1.02 / np.sqrt(2.0) * 256), 1 / np.sqrt(2))])
def test_fbp(gpu, N, nang, dc, dx):
    """fbp vs the oracle (FFT row filter) and the reference's PSNR > 28 dB
    pin for a 256^2 square with 360 views (test_xray_2d.py:88-102)."""
    x = np.zeros((N, N), np.float32)
    x[N // 4 : -N // 4, N // 4 : -N // 4] = 1.0
    ang = np.linspace(0, np.pi, nang, endpoint=False)
    kw = {} if dx is None else {"dx": dx}
    H = XRayTransform2D((N, N), ang, det_count=dc, **kw)
    y = H(x)
    xr = H.fbp(y).cpu().numpy()
    g = X.xr2d_defaults((N, N), ang, det_count=dc, dx=dx)
    ref = X.fbp(y.cpu().numpy(), g["x0"], g["dx"], g["nx"], g["y0"], g["angles"])
    e = rel_l2(xr, ref)
    mse = np.mean((xr.astype(np.float64) - x) ** 2)
    psnr = 10 * np.log10(1.0 / mse)
    print(f"  fbp {N}^2/{nang}/{dc}: rel-L2 vs oracle {e:.2e}, PSNR {psnr:.2f} dB", flush=True)
    assert e <= 1e-5
    if N == 256:
        assert psnr > 28


def test_operator_norm(gpu):
    """operator_norm (linop/_util.py:75-110) vs the dense float64 matrix."""
    from scico_amd.linop import operator_norm

    ang = np.linspace(0, np.pi, 10, endpoint=False)
    H = XRayTransform2D((12, 13), ang)
    g = X.xr2d_defaults((12, 13), ang)
    Hd = X.dense_matrix_overlap(g["x0"], g["dx"], g["nx"], g["angles"], g["y0"], g["ny"])
    ref = np.linalg.norm(Hd, 2)
    val = float(operator_norm(H, maxiter=200))
    assert abs(val - ref) <= 1e-4 * ref


def test_angles_outside_half_turn(gpu):
    """Negative angles and angles beyond 2 pi (the reference takes cos/sin of
    whatever it is given): bins and weights bit-exact, projections within
    the parity bar."""
    shape = (24, 28)
    ang = np.array([-2.5, -0.3, 0.0, 1.1, 3.3, 6.5, 9.0, 4.0 * np.pi + 0.2], np.float64)
    H = XRayTransform2D(shape, ang)
    g = X.xr2d_defaults(shape, ang)
    inds, w = H.weights()
    ri, rw = X.calc_weights(g["x0"], g["dx"], g["nx"], g["angles"], g["y0"])
    assert np.array_equal(inds.cpu().numpy(), ri) and np.array_equal(w.cpu().numpy(), rw)
    rng = np.random.default_rng(4)
    x = rng.random(shape).astype(np.float32)
    y = rng.standard_normal(H.output_shape).astype(np.float32)
    ref_f = X.project(x, g["x0"], g["dx"], g["y0"], g["ny"], g["angles"])
    ref_b = X.back_project(y, g["x0"], g["dx"], g["nx"], g["y0"], g["angles"])
    assert rel_l2(H(x).cpu().numpy(), ref_f) <= 1e-5
    assert rel_l2(H.T(y).cpu().numpy(), ref_b) <= 1e-5


def test_wide_detector_2048(gpu):
    """2048^2 with the default det_count 2897: rows wider than 64 KiB of LDS
    take the device's full LDS (or the general-geometry forward); parity on
    12 views against the C restatement, mass per view and the dot test."""
    from oracle import cpu

    N, A = 2048, 12
    ang = phantom.angles(A)
    H = XRayTransform2D((N, N), ang)
    assert H.det_count == 2897
    info = H.launch_info()
    x = phantom.ellipse_phantom(N, 12, seed=2)
    y = H(torch.as_tensor(x, device="cuda")).cpu().numpy()
    g = X.xr2d_defaults((N, N), ang)
    ry = cpu.project(x, g)
    e = rel_l2(y, ry)
    ys = np.random.default_rng(1).standard_normal((A, H.det_count)).astype(np.float32)
    eb = rel_l2(H.T(ys).cpu().numpy(), cpu.back_project(ys, g))
    print(f"  2048^2/2897 bins: {info}; forward rel-L2 {e:.2e}, adjoint {eb:.2e}", flush=True)
    assert e <= 1e-5 and eb <= 1e-5
    np.testing.assert_allclose(y.astype(np.float64).sum(axis=1),
                               np.full(A, x.astype(np.float64).sum()), rtol=1e-5)
    assert valid_adjoint(H, H.T, eps=1e-5)


@pytest.mark.parametrize("dx,dc", [(3.0, 180), (0.02, 5)])
def test_general_geometry_kernels(gpu, dx, dc):
    """Pixels of 3 detector units (adjoint tile over 64 bins) and of 0.02
    (forward lane stride over 64) run the general-geometry gathers: parity
    with the oracle, dot test, and ADMM-TV through the fused solver's nd
    passes against the generic ADMM path."""
    shape, A = (40, 36), 20
    ang = np.linspace(0, np.pi, A, endpoint=False)
    H = XRayTransform2D(shape, ang, dx=dx, det_count=dc)
    info = H.launch_info()
    print(f"  dx {dx}: {info}", flush=True)
    assert not H.tiled
    g = X.xr2d_defaults(shape, ang, dx=dx, det_count=dc)
    rng = np.random.default_rng(3)
    x = rng.random(shape).astype(np.float32)
    y = rng.standard_normal(H.output_shape).astype(np.float32)
    assert rel_l2(H(x).cpu().numpy(), X.project(x, g["x0"], g["dx"], g["y0"], g["ny"],
                                                  g["angles"])) <= 1e-5
    assert rel_l2(H.T(y).cpu().numpy(), X.back_project(y, g["x0"], g["dx"], g["nx"], g["y0"],
                                                        g["angles"])) <= 1e-5
    assert valid_adjoint(H, H.T, eps=1e-5)
    from scico_amd import functional, linop, loss
    from scico_amd.optimize import ADMM, LinearSubproblemSolver
    from scico_amd.optimize import _admm as admm_mod

    yy = H(x)
    xs = []
    for fused in (True, False):
        old = admm_mod._FusedTV.enabled
        admm_mod._FusedTV.enabled = fused
        try:
            s = ADMM(f=loss.SquaredL2Loss(y=yy, A=H), g_list=[1e-2 * functional.L21Norm()],
                     C_list=[linop.FiniteDifference(shape, append=0)], rho_list=[1.0], maxiter=1,
                     subproblem_solver=LinearSubproblemSolver(cg_kwargs={"tol": 0.0,
                                                                         "maxiter": 5}))
            assert (s._fused is not None) == fused
            xs.append(s.solve().cpu().numpy())
        finally:
            admm_mod._FusedTV.enabled = old
    assert rel_l2(xs[0], xs[1]) <= 1e-4
