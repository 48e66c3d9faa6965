"""Seeded synthetic phantoms (host, NumPy) for tests and benchmarks.

The reference has no ellipse generator; this one is defined here:
coordinates are normalised to [-1, 1] at pixel centres; each ellipse has
centre ~ U[-0.7, 0.7]^d, semi-axes ~ U[0.05, 0.4]^d, rotation ~ U[0, pi)
(three Euler angles for ellipsoids, applied as Z-Y-Z) and intensity
~ U[0.1, 1].  Ellipses are rasterised additively in float64 by a
pixel-centre inside test and the result cast to float32.
"""

from __future__ import annotations

import numpy as np


def ellipse_phantom(N: int, n_ellipses: int = 10, seed: int = 0, M: int | None = None) -> np.ndarray:
    M = N if M is None else M
    rng = np.random.default_rng(seed)
    yy = -1.0 + (2.0 * np.arange(M) + 1.0) / M
    xx = -1.0 + (2.0 * np.arange(N) + 1.0) / N
    Y, Xg = np.meshgrid(yy, xx, indexing="ij")
    img = np.zeros((M, N), dtype=np.float64)
    for _ in range(n_ellipses):
        c = rng.uniform(-0.7, 0.7, size=2)
        ax = rng.uniform(0.05, 0.4, size=2)
        th = rng.uniform(0.0, np.pi)
        val = rng.uniform(0.1, 1.0)
        ct, st = np.cos(th), np.sin(th)
        dy, dxx = Y - c[0], Xg - c[1]
        u = ct * dxx + st * dy
        v = -st * dxx + ct * dy
        img[(u / ax[0]) ** 2 + (v / ax[1]) ** 2 <= 1.0] += val
    return img.astype(np.float32)


def ellipsoid_phantom(shape, n_ellipsoids: int = 30, seed: int = 0) -> np.ndarray:
    """3D analogue of :func:`ellipse_phantom`: centres U[-0.7, 0.7]^3,
    semi-axes U[0.05, 0.4]^3, rotation Rz(a) Ry(b) Rz(g) with each angle
    U[0, pi), intensities U[0.1, 1]; voxel centres on [-1, 1]^3.  Each
    ellipsoid is rasterised inside its bounding box only."""
    rng = np.random.default_rng(seed)
    grids = [-1.0 + (2.0 * np.arange(s) + 1.0) / s for s in shape]
    vol = np.zeros(shape, dtype=np.float64)
    for _ in range(n_ellipsoids):
        c = rng.uniform(-0.7, 0.7, size=3)
        ax = rng.uniform(0.05, 0.4, size=3)
        a, b, g = rng.uniform(0.0, np.pi, size=3)
        val = rng.uniform(0.1, 1.0)
        R = _rz(a) @ _ry(b) @ _rz(g)
        r = float(np.max(ax))
        sl = []
        for k, gk in enumerate(grids):
            lo = int(np.searchsorted(gk, c[k] - r, side="left"))
            hi = int(np.searchsorted(gk, c[k] + r, side="right"))
            sl.append(slice(lo, hi))
        if any(t.stop <= t.start for t in sl):
            continue
        Z = np.stack(np.meshgrid(grids[0][sl[0]], grids[1][sl[1]], grids[2][sl[2]],
                                 indexing="ij"), axis=-1)
        d = (Z - c) @ R
        inside = np.sum((d / ax) ** 2, axis=-1) <= 1.0
        vol[tuple(sl)][inside] += val
    return vol.astype(np.float32)


def _rz(t):
    c, s = np.cos(t), np.sin(t)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def _ry(t):
    c, s = np.cos(t), np.sin(t)
    return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])


def noisy(y: np.ndarray, sigma_frac: float = 0.01, seed: int = 1) -> np.ndarray:
    """``y + sigma * n`` with ``sigma = sigma_frac * max(y)``, n ~ N(0, 1)."""
    rng = np.random.default_rng(seed)
    sigma = sigma_frac * float(np.max(y))
    return (y + sigma * rng.standard_normal(y.shape)).astype(np.float32)


def angles(n: int) -> np.ndarray:
    """``n`` views uniform on [0, pi) (float64 -> float32)."""
    return np.linspace(0, np.pi, n, endpoint=False).astype(np.float32)
