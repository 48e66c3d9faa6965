"""Write ``tests/golden/unrolled_grad_fd.npz``: float64 central differences of
the unrolled AcceleratedPGM-TV reconstruction loss, for the GPU gradient
test.  Run in the build container (CPU only):
``python tests/golden/make_grad_fixture.py``.

Problem (the differentiable configuration's protocol at the largest size
the float64 oracle runs in a few minutes): 64^2 ellipse phantom (8
ellipses, seed 0), 60 views on [0, pi), default detector (91 bins), 1 %
noise (seed 1), K = 20 steps of AcceleratedPGM with the weighted Haar TV
prox (lam = 0.05), step 1/L with L = 1.01 ||A||^2 (float64 power
iteration), loss l(y, w) = ||x_K(y, w) - x_true||^2.  Reference pattern:
``scico/flax/inverse.py:35-110`` (a differentiable unrolled
reconstruction).

For 64 random directions (dy ~ N(0, 1) scaled to the sinogram's range,
dw ~ N(0, 1); NumPy default_rng(5)) the fixture holds the central
difference of ``oracle.solvers.unrolled_apgm_tv(dtype=float64)`` at
eps = 1e-3 (the protocol's step), 5e-4 and 1e-5.  The TV prox is only
piecewise smooth, so a step that carries a shrinkage coefficient across
its threshold makes a central difference inexact: the eps 1e-3 / 5e-4
pair measures that error direction by direction, and eps 1e-5 (float64
rounding still far below it) rarely crosses a kink.  The test evaluates
the GPU gradient (float32, checkpoints every 5 steps) at the same float32
(y, w) and compares.
"""

from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import solvers as S  # noqa: E402
from oracle import xray as X  # noqa: E402

N, V, K, LAM = 64, 60, 20, 0.05
DIRECTIONS, DIR_SEED = 64, 5
EPS = (1e-3, 5e-4, 1e-5)


def problem():
    """(g, xgt, y float32, w0 float32, A64, AT64) of the test problem."""
    from scico_amd import phantom

    ang = phantom.angles(V)
    g = X.xr2d_defaults((N, N), ang)
    xgt = phantom.ellipse_phantom(N, 8, seed=0)
    y = phantom.noisy(X.project(xgt, g["x0"], g["dx"], g["y0"], g["ny"], g["angles"]), 0.01,
                      seed=1)
    w0 = np.random.default_rng(4).uniform(0.5, 1.5, (N, N)).astype(np.float32)
    A64 = lambda v: X.project_f64(v, g["x0"], g["dx"], g["y0"], g["ny"], g["angles"])
    AT64 = lambda v: X.back_project_f64(v, g["x0"], g["dx"], g["nx"], g["y0"], g["angles"])
    return g, xgt, y, w0, A64, AT64


def directions(y_shape, y_scale):
    rng = np.random.default_rng(DIR_SEED)
    out = []
    for _ in range(DIRECTIONS):
        dy = rng.standard_normal(y_shape) * y_scale
        dw = rng.standard_normal((N, N))
        out.append((dy, dw))
    return out


def lipschitz(A64, AT64, iters=200):
    v = np.random.default_rng(0).standard_normal((N, N))
    mu = 0.0
    for _ in range(iters):
        u = AT64(A64(v))
        mu = float(np.linalg.norm(u) / np.linalg.norm(v))
        v = u / np.linalg.norm(u)
    return 1.01 * mu


def main():
    g, xgt, y, w0, A64, AT64 = problem()
    t0 = time.time()
    L = lipschitz(A64, AT64)
    x64 = xgt.astype(np.float64)

    def loss(yv, wv):
        xk = S.unrolled_apgm_tv(yv, A64, AT64, L, LAM, wv, K, np.zeros((N, N)), dtype=np.float64)
        return float(np.sum((xk - x64) ** 2))

    yc, wc = y.astype(np.float64), w0.astype(np.float64)
    fd = np.zeros((len(EPS), DIRECTIONS))
    for i, (dy, dw) in enumerate(directions(y.shape, float(np.abs(y).max()))):
        for j, eps in enumerate(EPS):
            fd[j, i] = (loss(yc + eps * dy, wc + eps * dw) - loss(yc - eps * dy, wc - eps * dw)) \
                / (2 * eps)
    est = np.abs(fd[0] - fd[1]) / np.abs(fd[0])
    print(f"L {L:.6g}; {time.time() - t0:.1f} s; central-difference self-difference "
          f"median {np.median(est):.2e} max {est.max():.2e}")
    path = os.path.join(HERE, "unrolled_grad_fd.npz")
    np.savez_compressed(path, fd=fd, eps=np.array(EPS), L=np.array(L),
                        loss0=np.array(loss(yc, wc)))
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
