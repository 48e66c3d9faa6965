"""ADMM-TV with XRayTransform3D (the fused nd path) and HIP-graph replay of
the fused iteration (2D and 3D).

Bars: the fused nd passes (rhs, lhs with dot partials, z/u update with
statistics, dual residual) against the oracle's FiniteDifference and proxes
to 1e-6 / 1e-12; whole iterations against the oracle ADMM within 10x the
oracle's own drift under +-1 ulp projector noise (the CG x-step amplifies
last-bit differences); graph replays bitwise equal to eager iterations.
"""

import ctypes

import numpy as np
import pytest
import torch

from oracle import solvers as S
from oracle import xray as X
from scico_amd import _device, functional, linop, loss, phantom
from scico_amd._lib import call
from scico_amd.linop.xray import XRayTransform2D, XRayTransform3D
from scico_amd.optimize import ADMM, LinearSubproblemSolver
from scico_amd.optimize import _admm as admm_mod
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

f32 = np.float32


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def _shape_c(shape):
    return (ctypes.c_int32 * len(shape))(*shape)


# (64, 128, 160): more elements than one pass of the 1024-block grid, so
# threads take several batches and the coordinates carry on every axis.
@pytest.mark.parametrize("shape", [(7, 9, 11), (12, 10), (33,), (64, 128, 160)])
@pytest.mark.parametrize("circular", [False, True])
@pytest.mark.parametrize("iso", [True, False])
def test_nd_fused_ops(gpu, shape, circular, iso):
    nd = len(shape)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(shape).astype(f32)
    base = rng.standard_normal(shape).astype(f32)
    z = rng.standard_normal((nd, *shape)).astype(f32)
    u = rng.standard_normal((nd, *shape)).astype(f32)
    rho, thr = f32(0.7), f32(0.3)
    sh, s = _shape_c(shape), _device.stream()
    G = torch.empty(1024, dtype=torch.float64, device="cuda")
    P = lambda t: t.data_ptr()
    # rhs (device copies are held in variables: a temporary's memory may be
    # reused by the next allocation before the kernel runs)
    out, z_in, u_in = dev(base), dev(z), dev(u)
    call("scxr_admm_rhs_nd", nd, sh, P(out), P(z_in), P(u_in), float(rho), int(circular),
         P(out), s)
    ref = (base + rho * S.fd_adj((z - u).astype(f32), circular)).astype(f32)
    assert rel_l2(out.cpu().numpy(), ref) <= 1e-6
    # lhs with dot partials
    ata = dev(base)
    xd = dev(x)
    G.zero_()
    call("scxr_admm_lhs_nd", nd, sh, P(xd), P(ata), float(rho), int(circular), P(ata), P(G),
         None, s)
    ref = (rho * S.fd_adj(S.fd(x, circular), circular) + base).astype(f32)
    got = ata.cpu().numpy()
    assert rel_l2(got, ref) <= 1e-6
    dot = float(G.sum())
    assert abs(dot - np.vdot(x.astype(np.float64), got.astype(np.float64))) <= 1e-12 * np.abs(
        x.astype(np.float64) * got).sum()
    # z / u update with statistics
    zd, ud, zo = dev(z), dev(u), torch.empty((nd, *shape), device="cuda")
    wp, wg = torch.zeros(1024, dtype=torch.float64, device="cuda"), torch.zeros(
        1024, dtype=torch.float64, device="cuda")
    call("scxr_admm_zu_nd", nd, sh, P(xd), P(zd), P(ud), P(zo), float(thr), int(circular),
         int(iso), P(wp), P(wg), s)
    Cx = S.fd(x, circular)
    v = (Cx + u).astype(f32)
    zr = S.l21_prox(v, thr) if iso else S.l1_prox(v, thr)
    ur = ((u + Cx) - zr).astype(f32)
    assert rel_l2(zd.cpu().numpy(), zr) <= 1e-6
    assert rel_l2(ud.cpu().numpy(), ur) <= 1e-6
    np.testing.assert_array_equal(zo.cpu().numpy(), z)
    prml = float(wp.sum())
    pref = float(np.sum((Cx - zd.cpu().numpy()).astype(f32).astype(np.float64) ** 2))
    assert abs(prml - pref) <= 1e-10 * pref
    gref = S.l21_norm(zd.cpu().numpy()) if iso else S.l1_norm(zd.cpu().numpy())
    assert abs(float(wg.sum()) - gref) <= 1e-10 * gref
    # dual residual
    wd = torch.zeros(1024, dtype=torch.float64, device="cuda")
    call("scxr_admm_dual_nd", nd, sh, P(zd), P(z_in), float(rho), int(circular), P(wd), s)
    dref = (rho * S.fd_adj((zd.cpu().numpy() - z).astype(f32), circular)).astype(np.float64)
    assert abs(float(wd.sum()) - float(np.sum(dref**2))) <= 1e-10 * float(np.sum(dref**2))


def _problem3d(N=20, V=30, seed=0):
    det = (int(np.ceil(N * 1.42)), N)
    ang = phantom.angles(V)
    M = XRayTransform3D.matrices_from_euler_angles((N, N, N), det, "Y", ang)
    vol = phantom.ellipsoid_phantom((N, N, N), 8, seed=seed)
    M32 = M.astype(f32)
    y0 = X.project_3d(vol, M32, det)
    y = phantom.noisy(y0, 0.01, seed=seed + 1)
    A = lambda v: X.project_3d(v, M32, det)
    AT = lambda v: X.back_project_3d(v, M32, (N, N, N))
    return M, det, vol, y, A, AT


def _solver3d(M, det, y, N, lam, rho, iters, fused=True, x0=None, iso=False):
    old = admm_mod._FusedTV.enabled
    admm_mod._FusedTV.enabled = fused
    try:
        H = XRayTransform3D((N, N, N), M, det)
        g = lam * (functional.L21Norm() if iso else functional.L1Norm())
        s = ADMM(f=loss.SquaredL2Loss(y=dev(y), A=H), g_list=[g],
                 C_list=[linop.FiniteDifference((N, N, N), append=0)], rho_list=[rho],
                 maxiter=iters, x0=None if x0 is None else dev(x0),
                 subproblem_solver=LinearSubproblemSolver(cg_kwargs={"tol": 1e-4, "maxiter": 10}))
        assert (s._fused is not None) == fused
        return s
    finally:
        admm_mod._FusedTV.enabled = old


# Whole iterations are held to twice the largest drift of the oracle runs
# whose projector outputs carry +-1 ulp noise, over JITTER_SEEDS (adjoint
# noise seeded seed + 10); F32_RES absorbs floors that round to zero.
JITTER_SEEDS = (1, 2, 3, 4, 5)
F32_RES = 1e-6


def _noisy(A, seed):
    rng = np.random.default_rng(seed)
    eps = f32(6e-8)
    return lambda v: (lambda r: (r * (1 + eps * rng.choice([-1, 0, 1], r.shape))).astype(f32))(A(v))


@pytest.mark.parametrize("iso", [False, True])
def test_admm_tv_3d_one_step(gpu, iso):
    N, lam, rho = 20, 2e-2, 1.0
    M, det, vol, y, A, AT = _problem3d(N)
    x1, _, z1, u1 = S.admm_tv(y, A, AT, lam, rho, np.zeros((N, N, N), f32), 2, cg_maxiter=10,
                              isotropic=iso, state=True)
    xr, hr, _, _ = S.admm_tv(y, A, AT, lam, rho, x1, 1, cg_maxiter=10, isotropic=iso, z0=z1,
                             u0=u1, state=True)
    floor = ofloor = 0.0
    for seed in JITTER_SEEDS:
        xn, hn, _, _ = S.admm_tv(y, _noisy(A, seed), _noisy(AT, seed + 10), lam, rho, x1, 1,
                                 cg_maxiter=10, isotropic=iso, z0=z1, u0=u1, state=True)
        floor = max(floor, rel_l2(xn, xr))
        ofloor = max(ofloor, abs(hn[0][1] - hr[0][1]) / hr[0][1])
    for fused in (True, False):
        s = _solver3d(M, det, y, N, lam, rho, 1, fused, x0=x1, iso=iso)
        s.z_list, s.u_list, s.z_list_old = [dev(z1)], [dev(u1)], [dev(z1)]
        s._after_load_state()
        x = s.solve().cpu().numpy()
        h = s.history(transpose=True)
        ex = rel_l2(x, xr)
        eo = abs(h.Objective[0] - hr[0][1]) / hr[0][1]
        print(f"  3D ADMM step iso={iso} fused={fused}: x {ex:.2e} (largest floor {floor:.2e}), "
              f"objective {eo:.2e} (floor {ofloor:.2e})", flush=True)
        assert ex <= 2 * floor + F32_RES
        assert eo <= 2 * ofloor + F32_RES
        assert h.CG_It[0] == hr[0][4]


def test_admm_tv_3d_multi_step(gpu):
    N, lam, rho, it = 20, 2e-2, 1.0, 5
    M, det, vol, y, A, AT = _problem3d(N)
    x0 = np.zeros((N, N, N), f32)
    xr, hr = S.admm_tv(y, A, AT, lam, rho, x0, it, cg_maxiter=10, isotropic=False)
    floor = 0.0
    for seed in JITTER_SEEDS:
        xn, _ = S.admm_tv(y, _noisy(A, seed), _noisy(AT, seed + 10), lam, rho, x0, it,
                          cg_maxiter=10, isotropic=False)
        floor = max(floor, rel_l2(xn, xr))
    s = _solver3d(M, det, y, N, lam, rho, it)
    x = s.solve().cpu().numpy()
    h = s.history(transpose=True)
    snr = lambda a: 10 * np.log10(np.var(vol) / np.mean((a.astype(np.float64) - vol) ** 2))
    e = rel_l2(x, xr)
    print(f"  3D ADMM x{it}: vs oracle {e:.2e} (floor {floor:.2e}); SNR {snr(x):.3f} vs "
          f"{snr(xr):.3f} dB", flush=True)
    assert e <= 2 * floor + F32_RES
    assert abs(snr(x) - snr(xr)) < 0.1
    assert list(h.CG_It) == [r[4] for r in hr]


def _graph_replay_matches_eager(solver_factory, steps=4):
    """Capture one fused iteration, replay it ``steps`` times with eager
    work and synchronisation between replays, and compare every statistics
    row and the final x bitwise with ``steps`` eager iterations."""
    se = solver_factory()
    fe = se._fused
    fe._reserve(steps + 2)
    fe.reset()
    for _ in range(steps):
        fe.enqueue()
    torch.cuda.synchronize()
    xe, he = fe.x.clone(), fe.hist[:steps].clone()

    sg = solver_factory()
    fg = sg._fused
    fg._reserve(steps + 2)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        fg.enqueue()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        fg.enqueue()
    fg.reset()
    torch.cuda.synchronize()
    for k in range(steps):
        graph.replay()
        torch.cuda.synchronize()
        _ = torch.ones(1000, device="cuda").sum().item()  # unrelated eager work
        assert torch.equal(fg.hist[k], he[k]), (k, fg.hist[k], he[k])
    assert torch.equal(fg.x, xe)


def test_graph_replay_3d(gpu):
    N = 16
    M, det, vol, y, _, _ = _problem3d(N, 24)
    _graph_replay_matches_eager(lambda: _solver3d(M, det, y, N, 2e-2, 1.0, 8))


def test_graph_replay_2d(gpu):
    N, V, D = 64, 60, 91
    ang = phantom.angles(V)
    H = XRayTransform2D((N, N), ang, det_count=D)
    y = H(phantom.ellipse_phantom(N, 6, seed=0)).cpu().numpy()

    def make():
        return ADMM(f=loss.SquaredL2Loss(y=dev(y), A=H), g_list=[2e-2 * functional.L21Norm()],
                    C_list=[linop.FiniteDifference((N, N), append=0)], rho_list=[1.0], maxiter=8,
                    subproblem_solver=LinearSubproblemSolver(cg_kwargs={"tol": 1e-4,
                                                                        "maxiter": 10}))

    _graph_replay_matches_eager(make)


def test_admm_tv3d_slab_config_length(gpu):
    """The 3D configuration at full iteration count on its central 16-slice
    slab (256 x 16 x 256 volume, 360 views about axis 1, detector 368 x 16,
    anisotropic TV, CG 10 steps with tol 1e-4, 50 iterations) against the
    oracle's run (``tests/golden/admm3d_slab.npz``, written by
    ``tests/golden/make_admm3d_fixture.py``): every statistics column within
    twice the per-iteration floor envelope (running maximum; residual
    columns scaled by their running maximum; plus 1e-6 of the statistic,
    float32 resolution), the pooled final volume and the SNR within twice
    their floors, CG counts differing on at most twice as many iterations
    as the other valid runs'."""
    import json
    import os
    import sys

    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    if gold not in sys.path:
        sys.path.insert(0, gold)
    import make_admm3d_fixture as F

    d = np.load(os.path.join(gold, "admm3d_slab.npz"), allow_pickle=False)
    fx = {k: d[k] for k in d.files}
    meta = json.loads(str(fx["meta"]))
    M, xgt, y = F.problem()
    assert F.sha256(y) == meta["y_sha256"], "regenerated projections differ from the fixture's"
    shape = (F.N, F.SLAB, F.N)
    H = XRayTransform3D(shape, M, F.DET)
    assert H.info()["mode"] == "slice"
    s = ADMM(f=loss.SquaredL2Loss(y=dev(y), A=H), g_list=[F.LAM * functional.L1Norm()],
             C_list=[linop.FiniteDifference(shape, append=0)], rho_list=[F.RHO],
             maxiter=F.ITERS, subproblem_solver=LinearSubproblemSolver(
                 cg_kwargs={"tol": F.CG_TOL, "maxiter": F.CG_MAXITER}))
    assert s._fused is not None
    x = s.solve().cpu().numpy()
    h = s.history(transpose=True)
    hr = fx["hist"]
    sc = F.scales(hr)
    worst = {}
    for col, key, field in ((1, "objective", "Objective"), (2, "prml", "Prml_Rsdl"),
                            (3, "dual", "Dual_Rsdl"), (5, "cg_res", "CG_Res")):
        got = np.asarray(getattr(h, field), dtype=np.float64)
        dev_ = np.abs(got - hr[:, col]) / sc[key]
        bar = 2 * np.maximum.accumulate(fx[f"floor_{key}"]) + 1e-6 * np.abs(hr[:, col]) / sc[key]
        worst[key] = float(np.max(dev_ / bar))
        over = np.nonzero(dev_ > bar)[0]
        assert len(over) == 0, (key, [(int(k), dev_[k], bar[k]) for k in over[:8]])
    cg_diff = int(np.sum(np.asarray(h.CG_It, np.int64) != hr[:, 4].astype(np.int64)))
    fin = meta["floor_final"]
    snr = F.snr(x, xgt)
    e_pool = rel_l2(F.pool(x), fx["x_pool"])
    print(f"  3D slab x{F.ITERS}: worst deviation / bar {worst}; CG counts differ on {cg_diff} "
          f"(variants up to {fin['cg_count_diffs']}); SNR {snr:.3f} vs {meta['snr_db']:.3f} dB "
          f"(floor {fin['snr_abs']:.3f}); pooled volume rel-L2 {e_pool:.2e} (floor "
          f"{fin['x_pool_rel_l2']:.2e})", flush=True)
    assert cg_diff <= 2 * fin["cg_count_diffs"]
    assert abs(snr - meta["snr_db"]) <= 2 * fin["snr_abs"] + 1e-3
    assert e_pool <= 2 * fin["x_pool_rel_l2"] + 1e-7
