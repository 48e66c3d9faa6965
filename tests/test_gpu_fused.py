"""Parity of each fused ADMM-TV operation against the oracle, from
identical inputs (the full iteration is too ill-conditioned in float32 for
a 1e-5 bar; see test_gpu_solvers.py)."""

import ctypes

import numpy as np
import pytest
import torch

from oracle import solvers as S
from oracle import xray as X
from scico_amd import _lib
from scico_amd.linop.xray import XRayTransform2D
from scico_amd.solver import CGWork
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

N, A_, D = 96, 120, 140
RHO = 1.7


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


@pytest.fixture(scope="module")
def setup(gpu):
    ang = np.linspace(0, np.pi, A_, endpoint=False).astype(np.float32)
    H = XRayTransform2D((N, N), ang, det_count=D)
    g = X.xr2d_defaults((N, N), ang, det_count=D)
    rng = np.random.default_rng(11)
    return dict(
        H=H, g=g,
        A=lambda v: X.project(v, g["x0"], g["dx"], g["y0"], g["ny"], g["angles"]),
        AT=lambda v: X.back_project(v, g["x0"], g["dx"], g["nx"], g["y0"], g["angles"]),
        x=rng.standard_normal((N, N)).astype(np.float32),
        y=rng.standard_normal((A_, D)).astype(np.float32),
        z=rng.standard_normal((2, N, N)).astype(np.float32),
        u=rng.standard_normal((2, N, N)).astype(np.float32) * 0.1,
    )


def info(H):
    out = (ctypes.c_int32 * 4)()
    _lib.call("scxr_xr2d_info", H.plan, out)
    return list(out)


@pytest.mark.parametrize("circ", [0, 1])
def test_rhs_epilogue(setup, circ):
    d = setup
    out = torch.empty((N, N), device="cuda")
    y, z, u = dev(d["y"]), dev(d["z"]), dev(d["u"])  # keep the buffers alive over the call
    _lib.call("scxr_xr2d_adjoint_epi", d["H"].plan, y.data_ptr(), out.data_ptr(), 2, RHO, circ,
              None, z.data_ptr(), u.data_ptr(), None, None, None)
    ref = d["AT"](d["y"]) + np.float32(RHO) * S.fd_adj((d["z"] - d["u"]).astype(np.float32),
                                                        bool(circ))
    assert rel_l2(out.cpu().numpy(), ref) <= 1e-6


@pytest.mark.parametrize("circ", [0, 1])
def test_lhs_epilogue_and_curvature(setup, circ):
    d = setup
    H = d["H"]
    p = dev(d["x"])
    s = H.project(p)
    out = torch.empty((N, N), device="cuda")
    n_adj = info(H)[1]
    parts = torch.empty(n_adj, dtype=torch.float64, device="cuda")
    _lib.call("scxr_xr2d_adjoint_epi", H.plan, s.data_ptr(), out.data_ptr(), 1, RHO, circ,
              p.data_ptr(), None, None, parts.data_ptr(), None, None)
    ref = (np.float32(RHO) * S.fd_adj(S.fd(d["x"], bool(circ)), bool(circ))
           + d["AT"](d["A"](d["x"]))).astype(np.float32)
    assert rel_l2(out.cpu().numpy(), ref) <= 1e-5
    pap = float(parts.sum())
    ref_pap = float(np.sum(d["x"].astype(np.float64) * ref))
    assert abs(pap - ref_pap) <= 1e-5 * abs(ref_pap)


def test_forward_residual_partials(setup):
    d = setup
    H = d["H"]
    n_fwd = info(H)[0]
    parts = torch.empty(n_fwd, dtype=torch.float64, device="cuda")
    s = torch.empty(H.output_shape, device="cuda")
    x, y = dev(d["x"]), dev(d["y"])
    _lib.call("scxr_xr2d_forward_res", H.plan, x.data_ptr(), s.data_ptr(), y.data_ptr(),
              parts.data_ptr(), None, None)
    ref = float(np.sum((d["y"] - d["A"](d["x"])).astype(np.float64) ** 2))
    assert abs(float(parts.sum()) - ref) <= 1e-5 * ref


@pytest.mark.parametrize("circ,iso", [(0, 1), (1, 1), (0, 0)])
def test_zu_update_and_dual(setup, circ, iso):
    d = setup
    lam = 0.3
    thr = np.float32((1.0 / RHO) * lam)
    x, z, u = dev(d["x"]), dev(d["z"]), dev(d["u"])
    zold = torch.empty_like(z)
    G = int(_lib.load().scxr_elem_blocks(N * N))
    wp = torch.empty(G, dtype=torch.float64, device="cuda")
    wg = torch.empty(G, dtype=torch.float64, device="cuda")
    wd = torch.empty(G, dtype=torch.float64, device="cuda")
    _lib.call("scxr_admm_zu2d", N, N, x.data_ptr(), z.data_ptr(), u.data_ptr(), zold.data_ptr(),
              float(thr), circ, iso, wp.data_ptr(), wg.data_ptr(), None)
    _lib.call("scxr_admm_dual2d", N, N, z.data_ptr(), zold.data_ptr(), RHO, circ, wd.data_ptr(),
              None)
    Cx = S.fd(d["x"], bool(circ))
    v = (Cx + d["u"]).astype(np.float32)
    zr = S.l21_prox(v, thr) if iso else S.l1_prox(v, thr)
    ur = ((d["u"] + Cx) - zr).astype(np.float32)
    np.testing.assert_allclose(z.cpu().numpy(), zr, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(u.cpu().numpy(), ur, rtol=1e-6, atol=1e-6)
    np.testing.assert_array_equal(zold.cpu().numpy(), d["z"])
    prml = float(np.sum((Cx.astype(np.float64) - zr) ** 2))
    gval = S.l21_norm(zr) if iso else S.l1_norm(zr)
    dual = float(np.sum((np.float32(RHO) * S.fd_adj((zr - d["z"]).astype(np.float32),
                                                    bool(circ))).astype(np.float64) ** 2))
    assert abs(float(wp.sum()) - prml) <= 1e-6 * prml
    assert abs(float(wg.sum()) - gval) <= 1e-6 * gval
    assert abs(float(wd.sum()) - dual) <= 1e-5 * dual


def test_one_cg_step(setup):
    """cg_start + one CG iteration with the x-step operator, vs the oracle's
    cg with maxiter=1 (one step is well conditioned)."""
    d = setup
    H = d["H"]
    x0 = d["x"] * np.float32(0.5)
    b = (d["AT"](d["y"])).astype(np.float32)
    lhs = lambda v: (np.float32(RHO) * S.fd_adj(S.fd(v)) + d["AT"](d["A"](v))).astype(np.float32)
    xr, inf = S.cg(lhs, b, x0, tol=0.0, maxiter=1)
    n = N * N
    w = CGWork(n, 1)
    x = dev(x0).clone()
    r, p, Ap = (torch.empty((N, N), device="cuda") for _ in range(3))
    n_adj = info(H)[1]
    parts = torch.empty(n_adj, dtype=torch.float64, device="cuda")

    def apply(v, out, dot):
        s = H.project(v)
        _lib.call("scxr_xr2d_adjoint_epi", H.plan, s.data_ptr(), out.data_ptr(), 1, RHO, 0,
                  v.data_ptr(), None, None, parts.data_ptr() if dot else None, None, None)

    apply(x, Ap, False)
    w.start(dev(b), Ap, r, p, 0.0, 0.0)
    apply(p, Ap, True)
    w.iterate(0, x, r, p, Ap, parts, n_adj)
    assert int(w.info()["num_iter"].item()) == 1
    assert rel_l2(x.cpu().numpy(), xr) <= 1e-6
    assert abs(float(w.info()["rel_res"].item()) - inf["rel_res"]) <= 1e-5 * inf["rel_res"]


def test_cg_iter_t_and_forward_t(gpu):
    """The CG direction update that also writes p transposed gives x, r, p,
    the CG state and flags bitwise equal to scxr_cg_iter's, and pT equal to
    p.T; the forward fed that transpose equals the plain forward bitwise
    (edge tiles: 70 x 45 image; views of both streaming classes)."""
    from scico_amd.solver import CGWork

    M_, N_ = 70, 45
    rng = np.random.default_rng(17)
    b, Ax, Ap = (dev(rng.standard_normal((M_, N_)).astype(np.float32)) for _ in range(3))
    out = {}
    for fused in (False, True):
        cg = CGWork(M_ * N_, 3)
        x = torch.zeros((M_, N_), device="cuda")
        r, p = torch.empty_like(x), torch.empty_like(x)
        pT = torch.full((N_, M_), -7.0, device="cuda")
        cg.start(b, Ax, r, p, 0.0, 0.0)
        pap = torch.tensor([2.5 * float((b.double() ** 2).sum())], dtype=torch.float64,
                           device="cuda")
        for k in range(2):
            if fused:
                cg.iterate_t(k, x, r, p, pT, Ap, pap, 1)
            else:
                cg.iterate(k, x, r, p, Ap, pap, 1)
        out[fused] = (x.clone(), r.clone(), p.clone(), cg.state.clone(), cg.flags.clone(), pT)
    for a, b2 in zip(out[False][:5], out[True][:5]):
        assert torch.equal(a, b2)
    assert torch.equal(out[True][5], out[True][2].t().contiguous())
    ang = np.linspace(0, np.pi, 30, endpoint=False).astype(np.float32)
    H = XRayTransform2D((M_, N_), ang)
    im = dev(rng.random((M_, N_)).astype(np.float32))
    imT = im.t().contiguous()
    s1 = torch.empty(H.output_shape, device="cuda")
    s2 = torch.full(H.output_shape, -1.0, device="cuda")
    _lib.call("scxr_xr2d_forward", H.plan, im.data_ptr(), s1.data_ptr(), None, None)
    _lib.call("scxr_xr2d_forward_t", H.plan, im.data_ptr(), imT.data_ptr(), s2.data_ptr(), None,
              None)
    assert torch.equal(s1, s2)


@pytest.mark.parametrize("circ", [0, 1])
def test_rhs2d_equals_rhs_epilogue(setup, circ):
    """A^T y kept and rho C^T (z - u) added in the elementwise pass equals
    the adjoint's right-hand-side epilogue bit for bit."""
    d = setup
    H = d["H"]
    y, z, u = dev(d["y"]), dev(d["z"]), dev(d["u"])
    epi = torch.empty((N, N), device="cuda")
    _lib.call("scxr_xr2d_adjoint_epi", H.plan, y.data_ptr(), epi.data_ptr(), 2, RHO, circ, None,
              z.data_ptr(), u.data_ptr(), None, None, None)
    aty = torch.empty((N, N), device="cuda")
    _lib.call("scxr_xr2d_adjoint", H.plan, y.data_ptr(), aty.data_ptr(), None, None)
    out = torch.full((N, N), -3.0, device="cuda")
    _lib.call("scxr_admm_rhs2d", N, N, aty.data_ptr(), z.data_ptr(), u.data_ptr(), RHO, circ,
              out.data_ptr(), None)
    assert torch.equal(out, epi)
