"""Write ``tests/golden/admm3d_slab.npz``: the oracle's 3D ADMM-TV at
configuration length on a slab of the 3D configuration, for the GPU test.
Run in the build container (CPU only, about an hour):
``python tests/golden/make_admm3d_fixture.py``.

Problem: the 3D configuration (256^3 ellipsoid phantom, 30 ellipsoids,
seed 0; 360 views about volume axis 1 on [0, pi); detector 368 x 256;
1 % noise, seed 1; ADMM with 2e-2 * anisotropic TV of
FiniteDifference(append=0), rho 1, CG 10 steps, tol 1e-4; 50 iterations)
restricted to its central 16 slices along the rotation axis: volume
256 x 16 x 256, detector 368 x 16.  Every slice is one of the full
configuration's slices, projected with the same in-plane geometry; only the
slab is thinner, because the oracle's 3D projector takes about a second per
projection at this size (a full-size run would take hours per variant).
The projector is the C restatement ``oracle.cpu.project_3d_sep`` /
``back_project_3d_sep``, bitwise equal to ``oracle.xray.project_3d`` /
``back_project_3d`` (``tests/test_oracle_cpu.py``).

Floors as in ``make_admm_fixtures.py``: per iteration the largest deviation
of three other valid float32 runs (float64 accumulation of the projection
sums; two runs with projector-output noise of the size of the oracle's own
rounding, seeds 1 and 2) from the plain run, residual columns scaled by
their running maximum.
"""

from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import cpu  # noqa: E402
from oracle import solvers as S  # noqa: E402
from oracle import xray as X  # noqa: E402

N, SLAB, V, DET = 256, 16, 360, (368, 16)
ITERS, LAM, RHO, CG_TOL, CG_MAXITER = 50, 2e-2, 1.0, 1e-4, 10
VARIANTS = ("f64acc", 1, 2)
POOL = (8, 2, 8)
COLS = ((1, "objective"), (2, "prml"), (3, "dual"), (5, "cg_res"))
f32 = np.float32


def problem():
    """(matrices float32, x_true, y) of the slab (the tests call this too)."""
    from scico_amd import phantom

    vol = phantom.ellipsoid_phantom((N, N, N), 30, seed=0)
    s0 = (N - SLAB) // 2
    xgt = np.ascontiguousarray(vol[:, s0:s0 + SLAB, :])
    M = X.matrices_from_euler_angles(xgt.shape, DET, "Y", phantom.angles(V)).astype(f32)
    y = phantom.noisy(cpu.project_3d_sep(xgt, M, DET), 0.01, seed=1)
    return M, xgt, y


def sha256(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def snr(x, ref):
    ref = np.asarray(ref, np.float64)
    return float(10 * np.log10(np.var(ref) / np.mean((np.asarray(x, np.float64) - ref) ** 2)))


def pool(x):
    a, b, c = x.shape
    p, q, r = POOL
    return x.astype(np.float64).reshape(a // p, p, b // q, q, c // r, r).mean(axis=(1, 3, 5)) \
        .astype(f32)


def scales(hist):
    """|objective|, and the running maximum of each residual column."""
    out = {"objective": np.abs(hist[:, 1])}
    for col, key in ((2, "prml"), (3, "dual"), (5, "cg_res")):
        out[key] = np.maximum.accumulate(np.abs(hist[:, col]))
    return {k: np.maximum(v, 1e-300) for k, v in out.items()}


def projectors(M, variant=None, sig=(0.0, 0.0)):
    shape = (N, SLAB, N)
    if variant == "f64acc":
        return (lambda v: cpu.project_3d_sep(v, M, DET, f64acc=True)), \
            (lambda v: cpu.back_project_3d_sep(v, M, shape, f64acc=True))
    A = lambda v: cpu.project_3d_sep(v, M, DET)
    AT = lambda v: cpu.back_project_3d_sep(v, M, shape)
    if variant is None:
        return A, AT
    rng = np.random.default_rng(variant)
    jit = lambda r, s: (r * (1.0 + s * rng.standard_normal(r.shape))).astype(f32)
    return (lambda v: jit(A(v), sig[0])), (lambda v: jit(AT(v), sig[1]))


def run(M, y, variant=None, sig=(0.0, 0.0)):
    A, AT = projectors(M, variant, sig)
    x, hist = S.admm_tv(y, A, AT, LAM, RHO, np.zeros((N, SLAB, N), f32), ITERS, cg_tol=CG_TOL,
                        cg_maxiter=CG_MAXITER, isotropic=False)
    return x, np.asarray(hist, dtype=np.float64)


def main():
    M, xgt, y = problem()
    rel = lambda a, b: float(np.linalg.norm((a - b).astype(np.float64)) / np.linalg.norm(b))
    shape = (N, SLAB, N)
    sig = (rel(cpu.project_3d_sep(xgt, M, DET), cpu.project_3d_sep(xgt, M, DET, f64acc=True)),
           rel(cpu.back_project_3d_sep(y, M, shape),
               cpu.back_project_3d_sep(y, M, shape, f64acc=True)))
    print(f"oracle rounding (rel-L2 vs float64 accumulation): forward {sig[0]:.2e}, adjoint "
          f"{sig[1]:.2e}", flush=True)
    t0 = time.time()
    xr, hr = run(M, y)
    print(f"clean run {time.time() - t0:.1f} s, CG its {hr[:, 4].astype(int).tolist()}",
          flush=True)
    sc = scales(hr)
    env = {k: np.zeros(ITERS) for _, k in COLS}
    fin = {"x_rel_l2": 0.0, "x_pool_rel_l2": 0.0, "snr_abs": 0.0, "cg_count_diffs": 0}
    for v in VARIANTS:
        t0 = time.time()
        xn, hn = run(M, y, variant=v, sig=sig)
        for col, k in COLS:
            env[k] = np.maximum(env[k], np.abs(hn[:, col] - hr[:, col]) / sc[k])
        d = {"x_rel_l2": rel(xn, xr), "x_pool_rel_l2": rel(pool(xn), pool(xr)),
             "snr_abs": abs(snr(xn, xgt) - snr(xr, xgt)),
             "cg_count_diffs": int(np.sum(hn[:, 4] != hr[:, 4]))}
        fin = {k: max(fin[k], d[k]) for k in fin}
        print(f"variant {v}: {time.time() - t0:.1f} s, {d}", flush=True)
    out = {"hist": hr, "x_pool": pool(xr)}
    for k, v in env.items():
        out[f"floor_{k}"] = v
    meta = {"volume": [N, SLAB, N], "views": V, "det": list(DET), "iters": ITERS, "lam": LAM,
            "rho": RHO, "cg_tol": CG_TOL, "cg_maxiter": CG_MAXITER, "slab_offset": (N - SLAB) // 2,
            "noise": "0.01 * max(A xgt) * N(0,1), numpy default_rng(1)",
            "jitter_rel_std": {"forward": sig[0], "adjoint": sig[1]},
            "variants": [str(v) for v in VARIANTS], "pool": list(POOL), "y_sha256": sha256(y),
            "snr_db": snr(xr, xgt), "floor_final": fin}
    out["meta"] = np.array(json.dumps(meta))
    path = os.path.join(HERE, "admm3d_slab.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes); final floors {fin}", flush=True)


if __name__ == "__main__":
    main()
