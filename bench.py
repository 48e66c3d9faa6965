#!/usr/bin/env python
"""ADMM-TV reconstruction throughput on one MI355X.

Workload: 512 x 512 float32 ellipse phantom (20 ellipses, seed 0), 720 views
on [0, pi), 724 detector bins, 1 % Gaussian noise; ADMM with
SquaredL2Loss + 2e-2 * L21(FiniteDifference(append=0)), rho = 1 and a
10-step CG x-step, starting from x0 = 0.

value: iterations/s with fixed work per iteration (CG tol 0, so exactly 10
CG steps; 11 forward + 11 back projections, since an iteration keeps A x of
its residual projection and A^T y instead of recomputing them, plus the
iteration statistics),
timed over --steps iterations replayed from one captured HIP graph.  The
configured-tolerance run (CG tol 1e-4, 100 iterations, where warm-started
CG stops early) is reported beside it with its mean CG iteration count.

Prints one JSON line with the roofline of the dominant projector kernel
(HIP events) and the CPU baseline of the C/OpenMP restatement.  "extras"
reports the other workloads of the same hot path on the same GPU (skip
with --no-extras): the 1024^2 CircularConvolve + AcceleratedPGM-TV
deconvolution (200 iterations), 3D ADMM-TV on a 256^3 volume (360 views,
detector 368 x 256), 256 ragged slices batched in one launch per
projection (autotuned, graph-captured ADMM-TV), and the forward + backward
pass of the 20-step unrolled APGM-TV reconstruction (512^2, 360 views).

``--kernels`` runs only the projector launches (2D headline shape, 3D
slice shape) plus a 64 MiB copy for counter calibration: the workload of
the rocprofv3 counter runs.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
# lane-instructions / s: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz (a wave64 fp32,
# fp64 or conversion instruction issues every 4 cycles on its SIMD)
VALU_PEAK = 256 * 64 * 2.4e9

# HBM traffic per launch of the 2D projector at 512^2 / 720 views / 724 bins
# from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (KiB per dispatch), scaled
# by the counters of a 64 MiB copy (FETCH_SIZE reads 0.5x the bytes moved,
# WRITE_SIZE 1.0x): forward = image transpose + projection.
TRAFFIC_BYTES = {
    "forward": 1024 * ((524.1 + 9084.7) * 2.0 + (1024.0 + 2047.3) * 1.0),
    "adjoint": 1024 * (4936.2 * 2.0 + 1024.0 * 1.0),
}
# VALU instructions (lane-instructions) per (pixel, view) interaction,
# rocprofv3 --pmc SQ_INSTS_VALU x 64 / interactions, same workload.
VALU_PER_INTERACTION = {"forward": 13.27, "adjoint": 13.25}
# the profile these counter readings come from (commands and raw figures)
COUNTERS_FROM = "profiles/xr2d_projector_512x512_720views_det724.md"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_problem(N, A, D, n_ell, seed):
    import torch

    from scico_amd import phantom
    from scico_amd.linop.xray import XRayTransform2D

    ang = phantom.angles(A)
    xgt = phantom.ellipse_phantom(N, n_ell, seed=seed)
    H = XRayTransform2D((N, N), ang, det_count=D)
    y0 = H(torch.as_tensor(xgt, device="cuda")).cpu().numpy()
    y = phantom.noisy(y0, 0.01, seed=seed + 1)
    return ang, xgt, y, H


def make_solver(H, y, N, lam, rho, cg_maxiter, maxiter, cg_tol=1e-4):
    import torch

    from scico_amd import functional, linop, loss
    from scico_amd.optimize import ADMM, LinearSubproblemSolver

    f = loss.SquaredL2Loss(y=torch.as_tensor(y, device="cuda"), A=H)
    C = linop.FiniteDifference((N, N), append=0)
    g = lam * functional.L21Norm()
    s = ADMM(f=f, g_list=[g], C_list=[C], rho_list=[rho], maxiter=maxiter,
             subproblem_solver=LinearSubproblemSolver(cg_kwargs={"tol": cg_tol,
                                                                 "maxiter": cg_maxiter}))
    assert s._fused is not None, "fused ADMM-TV path not selected"
    return s


def time_kernels(H, reps=50, isolated=False):
    """Average duration of one forward (transpose + projection) and one
    adjoint launch, from HIP events on the launch stream: back to back
    (``reps`` launches between two events), or with ``isolated`` each
    launch between its own events and waited for, as a profiler times a
    dispatch."""
    import torch

    x = torch.rand(H.input_shape, device="cuda")
    y = torch.rand(H.output_shape, device="cuda")
    s = torch.empty(H.output_shape, device="cuda")
    b = torch.empty(H.input_shape, device="cuda")
    for _ in range(3):
        H.project(x, out=s)
        H.back_project(y, out=b)
    out = {}
    for name, fn in (("forward", lambda: H.project(x, out=s)),
                     ("adjoint", lambda: H.back_project(y, out=b))):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        if isolated:
            tot = 0.0
            for _ in range(reps):
                e0.record()
                fn()
                e1.record()
                e1.synchronize()
                tot += e0.elapsed_time(e1)
            out[name] = tot / reps / 1000.0
            continue
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        e1.synchronize()
        out[name] = e0.elapsed_time(e1) / reps / 1000.0
    return out


def run_timed(solver, steps, warmup, use_graph):
    """Warm up (capturing one iteration in a HIP graph), restart from x0 = 0,
    then time ``steps`` iterations.  Returns (seconds, itstat rows)."""
    import torch

    fused = solver._fused
    fused._reserve(steps + warmup + 8)
    if use_graph:
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            fused.enqueue()  # eager first iteration (also creates the plan)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            fused.enqueue()
        run = graph.replay
        for _ in range(max(0, warmup - 1)):
            run()
    else:
        run = fused.enqueue
        for _ in range(warmup):
            run()
    fused.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return dt, fused.hist[:steps].cpu().numpy()


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(y, N, A, D, lam, rho, cg_maxiter, iters):
    """The C/OpenMP restatement (projector) + NumPy (FD, prox, CG) of the
    same fixed-work iteration, with the projector compiled -march=native on
    this host (the portable x86-64-v3 build if that compile fails)."""
    import tempfile

    from oracle import _cbuild, cpu
    from oracle import xray as X
    from scico_amd import phantom

    march = "native"
    try:
        path = os.path.join(tempfile.mkdtemp(prefix="scxr_cpu_"), "libxray_cpu_native.so")
        cpu.use_library(_cbuild.build(native=True, lib=path))
    except Exception as e:  # noqa: BLE001 - fall back to the portable build
        log(f"bench: -march=native build of the CPU baseline failed ({e}); using x86-64-v3")
        march = "x86-64-v3"
    g = X.xr2d_defaults((N, N), phantom.angles(A), det_count=D)
    cpu.project(np.zeros((N, N), np.float32), g)  # load + warm the library
    t0 = time.perf_counter()
    cpu.admm_tv(y, g, lam, rho, np.zeros((N, N), np.float32), iters, cg_maxiter=cg_maxiter,
                cg_tol=0.0)
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", 0) or len(os.sched_getaffinity(0)))
    return {"value": iters / dt, "unit": "ADMM-TV iters/s", "cores": threads, "kind": "port",
            "cpu_model": _cpu_model(), "host_cpus": os.cpu_count(),
            "build": f"gcc -O3 -march={march} -fno-fast-math -ffp-contract=off -fopenmp",
            "sample": f"{iters} fixed-work ADMM-TV iterations (CG tol 0, 10 steps) of the "
                      f"same 512^2/720-view problem (C/OpenMP projector on {threads} threads + "
                      f"single-threaded NumPy FD/prox/CG), {dt:.1f} s"}


def extra_deconv(iters=200):
    """1024^2 ellipse phantom, 15 x 15 Gaussian PSF (sigma 2, centred at
    (7, 7)), 1 % noise; AcceleratedPGM with 2e-2 * IsotropicTVNorm, L0 =
    max |H|^2."""
    import torch

    from scico_amd import functional, loss, phantom
    from scico_amd.linop import CircularConvolve
    from scico_amd.optimize import AcceleratedPGM

    N = 1024
    r = np.arange(15) - 7
    psf = np.exp(-(r[:, None] ** 2 + r[None, :] ** 2) / (2 * 2.0**2))
    psf = (psf / psf.sum()).astype(np.float32)
    A = CircularConvolve(psf, (N, N), h_center=[7, 7])
    xgt = torch.as_tensor(phantom.ellipse_phantom(N, 20, seed=0), device="cuda")
    y = torch.as_tensor(phantom.noisy(A(xgt).cpu().numpy(), 0.01, seed=1), device="cuda")
    L0 = float(torch.max(torch.abs(A.h_dft.to(torch.complex128)) ** 2))
    f = loss.SquaredL2Loss(y=y, A=A)
    g = 2e-2 * functional.IsotropicTVNorm()

    def make(n):
        return AcceleratedPGM(f=f, g=g, L0=L0, x0=torch.zeros((N, N), device="cuda"), maxiter=n)

    make(3).solve()  # warm up
    torch.cuda.synchronize()
    s = make(iters)
    t0 = time.perf_counter()
    x = s.solve()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # The same iterations (step and objective, the device work of solve())
    # captured once in a HIP graph and replayed: the device rate without
    # the Python loop's per-launch host work.
    sg = make(iters)
    sg.step()
    sg.objective()  # eager first step (builds the cached A^T y)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        for _ in range(iters):
            sg.step()
            sg.objective()
    torch.cuda.synchronize()
    g0, g1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    g0.record()
    graph.replay()
    g1.record()
    g1.synchronize()
    gdt = g0.elapsed_time(g1) / 1000.0
    del graph, sg
    out = torch.empty_like(xgt)
    for _ in range(3):
        A._eval(xgt, out=out)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        A._eval(xgt, out=out)
    e1.record()
    e1.synchronize()
    # the same 20 applies from a captured graph (no per-launch host work)
    ag = torch.cuda.CUDAGraph()
    with torch.cuda.graph(ag):
        for _ in range(20):
            A._eval(xgt, out=out)
    ag.replay()
    torch.cuda.synchronize()
    e2, e3 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e2.record()
    ag.replay()
    e3.record()
    e3.synchronize()
    del ag
    snr = 10 * np.log10(float(torch.var(xgt.double())) /
                        float(torch.mean((x.double() - xgt.double()) ** 2)))
    return {"workload": "apgm_tv_deconv_1024x1024_psf15_sigma2", "iters": iters,
            "seconds": round(dt, 4), "iters_per_s": round(iters / dt, 2),
            "graph_iters_per_s": round(iters / gdt, 2),
            "cconv_apply_us": round(e0.elapsed_time(e1) / 20 * 1000, 2),
            "cconv_apply_graph_us": round(e2.elapsed_time(e3) / 20 * 1000, 2),
            "snr_db": round(snr, 3), "L0": L0, "lambda": 2e-2}


def extra_3d(steps=50, warmup=2):
    """256^3 ellipsoid phantom (30), 360 views about axis 1, detector
    368 x 256, 1 % noise; ADMM with 2e-2 * L1Norm (anisotropic TV) of
    FiniteDifference(append=0), rho 1, CG 10 steps (fixed work)."""
    import torch

    from scico_amd import functional, linop, loss, phantom
    from scico_amd.linop.xray import XRayTransform3D
    from scico_amd.optimize import ADMM, LinearSubproblemSolver

    N, V, det = 256, 360, (368, 256)
    M = XRayTransform3D.matrices_from_euler_angles((N, N, N), det, "Y", phantom.angles(V))
    H = XRayTransform3D((N, N, N), M, det)
    vol = torch.as_tensor(phantom.ellipsoid_phantom((N, N, N), 30, seed=0), device="cuda")
    y = H(vol)
    y = y + 0.01 * float(y.max()) * torch.randn(y.shape, device="cuda",
                                                generator=torch.Generator("cuda").manual_seed(1))
    s = ADMM(f=loss.SquaredL2Loss(y=y, A=H), g_list=[2e-2 * functional.L1Norm()],
             C_list=[linop.FiniteDifference((N, N, N), append=0)], rho_list=[1.0],
             maxiter=steps + warmup, subproblem_solver=LinearSubproblemSolver(
                 cg_kwargs={"tol": 0.0, "maxiter": 10}))
    assert s._fused is not None
    dt, hist = run_timed(s, steps, warmup, True)
    kt = time_kernels(H, reps=5)
    inter = float(N) ** 3 * V
    return {"workload": "admm_tv3d_256cube_360views_det368x256_cg10_fixedwork",
            "steps": steps, "ms_per_iter": round(1000 * dt / steps, 3),
            "iters_per_s": round(steps / dt, 3),
            "seconds_50_iters": round(50 * dt / steps, 3) if steps == 50 else None,
            "forward_ms": round(kt["forward"] * 1e3, 3), "adjoint_ms": round(kt["adjoint"] * 1e3, 3),
            "interactions_per_s": {k: round(inter / v, 1) for k, v in kt.items()},
            "last_objective": float(hist[-1, 0])}


def extra_3d_general(reps=10):
    """The reference's own 3D projector comparison geometry
    (``examples/scripts/ct_projector_comparison_3d.py:31-58``): a
    129 x 130 x 131 volume, 3 views with Euler angles "XY" (74 deg, 0/60/120
    deg), detector 128 x 129 -- not slice-decomposable, so the general-mode
    gather kernels run."""
    from scico_amd.linop.xray import XRayTransform3D

    shape, det = (129, 130, 131), (128, 129)
    ang = np.stack(np.broadcast_arrays(90.0 - 16.0, np.linspace(0, 180, 3, endpoint=False)),
                   axis=-1)
    M = XRayTransform3D.matrices_from_euler_angles(shape, det, "XY", ang, degrees=True)
    H = XRayTransform3D(shape, M, det)
    assert H.info()["mode"] == "general"
    kt = time_kernels(H, reps=reps)
    inter = float(np.prod(shape)) * 3
    return {"workload": "xray3d_general_129x130x131_3views_XY74_det128x129",
            "forward_ms": round(kt["forward"] * 1e3, 4),
            "adjoint_ms": round(kt["adjoint"] * 1e3, 4),
            "interactions_per_s": {k: round(inter / v, 1) for k, v in kt.items()}}


def extra_batch(steps=50, warmup=1):
    """256 slices: N from {128, 192, 256, 384, 512}, 90..720 views,
    det_count = ceil(sqrt(2) N), ellipse phantoms (10) + 1 % noise;
    autotuned persistent projector; ADMM-TV (lam 2e-2, rho 1, CG 10 fixed
    work) on all slices per graph replay."""
    import torch

    from scico_amd import phantom
    from scico_amd.batch import ADMMTVBatch, XRayTransform2DBatch

    rng = np.random.default_rng(8)
    shapes, angles = [], []
    for _ in range(256):
        n = int(rng.choice([128, 192, 256, 384, 512]))
        shapes.append((n, n))
        angles.append(phantom.angles(int(rng.integers(90, 721))))
    B = XRayTransform2DBatch(shapes, angles)
    t0 = time.perf_counter()
    tuned = B.autotune(vpgs=(1, 2), grids=(0, 2, 4), orders=(0, 1, 2), reps=2, adj_groups=(3, 2))
    t_tune = time.perf_counter() - t0
    xs = [phantom.ellipse_phantom(sh[0], 10, seed=k) for k, sh in enumerate(shapes)]
    y = B.project(B.pack_images(xs))
    y = y + 0.01 * float(y.max()) * torch.randn(y.shape, device="cuda",
                                                generator=torch.Generator("cuda").manual_seed(1))
    bs = ADMMTVBatch(B, y, 2e-2, 1.0, cg_maxiter=10, cg_tol=0.0, hist_cap=steps + warmup + 4)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        bs.enqueue()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        bs.enqueue()
    for _ in range(warmup):
        graph.replay()
    bs.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        graph.replay()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    inter = float(sum(s[0] * s[1] * len(a) for s, a in zip(shapes, angles)))
    return {"workload": "batched_256_ragged_slices_admm_tv_cg10_fixedwork", "slices": 256,
            "interactions_per_pass": inter, "autotuned": {k: tuned[k] for k in
                                                          ("vpg", "grid", "order",
                                                           "adj_groups")},
            "autotune_seconds": round(t_tune, 2),
            "projector_pass_ms": round(tuned["ms"], 3),
            "projector_interactions_per_s": round(2 * inter / (tuned["ms"] * 1e-3), 1),
            "steps": steps, "ms_per_iter": round(1000 * dt / steps, 2),
            "slice_iters_per_s": round(256 * steps / dt, 1),
            "seconds_50_iters": round(dt, 3) if steps == 50 else None}


def extra_unrolled(K=20):
    """512^2, 360 views (det_count 725), 20-step unrolled APGM-TV with a
    per-pixel TV weight map, checkpoints every 5 steps: forward and
    forward + backward (gradients w.r.t. the sinogram and the weights)."""
    import torch

    from scico_amd import autograd as AG
    from scico_amd import phantom
    from scico_amd.linop import operator_norm
    from scico_amd.linop.xray import XRayTransform2D

    N, V = 512, 360
    H = XRayTransform2D((N, N), phantom.angles(V))
    xt = torch.as_tensor(phantom.ellipse_phantom(N, 20, seed=0), device="cuda")
    y0 = H(xt)
    L = 1.01 * float(operator_norm(H, maxiter=30)) ** 2

    def once(grad):
        y = y0.clone().requires_grad_(grad)
        w = torch.ones((N, N), device="cuda", requires_grad=grad)
        x = AG.unrolled_apgm_tv(H, y, w, L, 2e-2, K=K, checkpoint_every=5)
        if grad:
            ((x - xt) ** 2).sum().backward()

    for g in (False, True):
        once(g)
    out = {}
    for g in (False, True):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            once(g)
        torch.cuda.synchronize()
        out["fwd_bwd_ms" if g else "fwd_ms"] = round((time.perf_counter() - t0) / 3 * 1e3, 3)
    return {"workload": "unrolled_apgm_tv_512x512_360views_K20_ckpt5", **out}


def kernels_only():
    """Projector launches only (counter runs): 10 x 2D forward + adjoint at
    512^2 / 720 / 724, 10 x 64 MiB copy (calibration), 3 x 3D forward +
    adjoint at 256^3 / 360 / (368, 256)."""
    import torch

    from scico_amd import _vec, phantom
    from scico_amd.linop.xray import XRayTransform2D, XRayTransform3D

    H = XRayTransform2D((512, 512), phantom.angles(720), det_count=724)
    x = torch.rand((512, 512), device="cuda")
    y = torch.rand((720, 724), device="cuda")
    s, b = torch.empty_like(y), torch.empty_like(x)
    big = torch.rand(16 << 20, device="cuda")
    big2 = torch.empty_like(big)
    for _ in range(10):
        H.project(x, out=s)
    for _ in range(10):
        H.back_project(y, out=b)
    for _ in range(10):
        _vec.copy(big, out=big2)
    M = XRayTransform3D.matrices_from_euler_angles((256, 256, 256), (368, 256), "Y",
                                                    phantom.angles(360))
    H3 = XRayTransform3D((256, 256, 256), M, (368, 256))
    v = torch.rand((256, 256, 256), device="cuda")
    p = torch.rand((360, 368, 256), device="cuda")
    for _ in range(3):
        H3.project(v)
        H3.back_project(p)
    torch.cuda.synchronize()
    log("bench: kernel launches done")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=18)
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--kernels", action="store_true",
                    help="projector launches only (for counter runs)")
    ap.add_argument("--only", choices=["deconvolution", "xray3d", "xray3d_general", "batched",
                                       "unrolled_grad"],
                    help="run one extra workload only and print its figures (profiling runs)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(pass --gpus 1).", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.kernels:
        kernels_only()
        return
    if args.only:
        fn = {"deconvolution": extra_deconv, "xray3d": extra_3d,
              "xray3d_general": extra_3d_general, "batched": extra_batch,
              "unrolled_grad": extra_unrolled}[args.only]
        print(json.dumps({args.only: fn()}, default=float))
        return

    N, A, D = 512, 720, 724
    lam, rho, cg_maxiter = 2e-2, 1.0, 10
    log(f"bench: building 512^2/720/724 problem on {torch.cuda.get_device_name(0)}")
    ang, xgt, y, H = build_problem(N, A, D, 20, 0)
    use_graph = not args.no_graph
    # Headline: fixed work per iteration (CG tol 0: exactly 10 CG steps, i.e.
    # 11 forward + 11 back projections per ADMM iteration), from x0 = 0.
    solver = make_solver(H, y, N, lam, rho, cg_maxiter, args.steps + args.warmup, cg_tol=0.0)
    log(f"bench: timing {args.steps} fixed-work ADMM iterations ({'graph' if use_graph else 'eager'})")
    dt, hist = run_timed(solver, args.steps, args.warmup, use_graph)
    ips = args.steps / dt
    log(f"bench: {args.steps} iterations in {dt:.3f} s ({ips:.1f} it/s); last itstat "
        f"obj={hist[-1, 0]:.6e} prml={hist[-1, 1]:.3e} dual={hist[-1, 2]:.3e} "
        f"cg_it={int(hist[-1, 3])} cg_res={hist[-1, 4]:.3e}")
    # Configured run: CG tol 1e-4 (the reference default for this workload),
    # 100 iterations from x0 = 0; CG stops early once warm starts converge.
    solver2 = make_solver(H, y, N, lam, rho, cg_maxiter, 100 + args.warmup, cg_tol=1e-4)
    dt2, hist2 = run_timed(solver2, 100, args.warmup, use_graph)
    cfg = {"iters": 100, "seconds": round(dt2, 4), "iters_per_s": round(100 / dt2, 2),
           "mean_cg_iters": round(float(hist2[:, 3].mean()), 2),
           "final_objective": float(hist2[-1, 0])}
    log(f"bench: configured tol=1e-4 run: 100 iterations in {dt2:.3f} s, mean CG "
        f"{cfg['mean_cg_iters']}")
    del solver, solver2

    # roofline figures use launches timed one at a time (as rocprofv3 times
    # each dispatch); back-to-back launches overlap one launch's ramp-down
    # with the next one's ramp-up and are reported beside them
    kt = time_kernels(H, isolated=True)
    kb = time_kernels(H)
    bytes_per = 4.0 * (N * N + A * D)  # compulsory bytes: image + sinogram
    inter = float(N * N * A)  # (pixel, view) interactions per application
    dom = max(kt, key=kt.get)
    t_dom = kt[dom]
    achieved = bytes_per / t_dom / 1e9
    vpi = VALU_PER_INTERACTION[dom]
    roofline = {
        "kernel": "xr2d " + ("forward (image transpose + projection)" if dom == "forward"
                             else "adjoint"),
        "bound": "hbm",
        "achieved": round(achieved, 3),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 6),
        "traffic": round(TRAFFIC_BYTES[dom]),
        # traffic and instructions per interaction are counter readings of a
        # separate rocprofv3 --pmc run (constants above), not of this run
        "counters_from": COUNTERS_FROM,
        "algorithmic_bytes": bytes_per,
        "avg_launch_us": {k: round(v * 1e6, 2) for k, v in kt.items()},
        "back_to_back_us": {k: round(v * 1e6, 2) for k, v in kb.items()},
        "interactions_per_s": {k: round(inter / v, 1) for k, v in kt.items()},
        # the projector is bound by vector-ALU issue, not HBM: its
        # lane-instruction rate against the VALU peak
        "valu": None if vpi is None else {
            "instr_per_interaction": vpi,
            "achieved": round(inter * vpi / t_dom, 1),
            "peak": VALU_PEAK,
            "unit": "lane-instr/s",
            "frac": round(inter * vpi / t_dom / VALU_PEAK, 4)},
    }
    log(f"bench: forward {kt['forward']*1e6:.1f} us, adjoint {kt['adjoint']*1e6:.1f} us per launch "
        f"(back to back {kb['forward']*1e6:.1f} us, {kb['adjoint']*1e6:.1f} us)")

    extras = None
    if not args.no_extras:
        extras = {}
        for name, fn in (("deconvolution", extra_deconv), ("xray3d", extra_3d),
                         ("xray3d_general", extra_3d_general),
                         ("batched", extra_batch), ("unrolled_grad", extra_unrolled)):
            log(f"bench: extra {name}")
            t0 = time.perf_counter()
            extras[name] = fn()
            log(f"bench: extra {name} done in {time.perf_counter() - t0:.1f} s: {extras[name]}")

    cpu = None
    if not args.no_cpu and args.cpu_iters > 0:
        log(f"bench: CPU baseline, {args.cpu_iters} iterations")
        cpu = cpu_baseline(y, N, A, D, lam, rho, cg_maxiter, args.cpu_iters)
        log(f"bench: CPU {cpu['value']:.4f} it/s on {cpu['cores']} threads")

    out = {
        "metric": "ADMM-TV reconstruction iterations/s (512^2, 720 views, det 724, CG 10)",
        "value": round(ips, 3),
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * dt / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ellipse phantom, 1% Gaussian noise)",
        "config": {"workload": "admm_tv_512x512_720views_det724_cg10_fixedwork",
                   "image": [N, N], "views": A, "det_count": D, "lambda": lam, "rho": rho,
                   "cg_maxiter": cg_maxiter, "cg_tol": 0.0, "graph": use_graph,
                   "parallelism": "single-gpu"},
        "configured_run": cfg,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "gpu_vs_cpu": round(ips / cpu["value"], 1) if cpu else None,
        "extras": extras,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
