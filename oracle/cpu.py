"""Python side of the C restatement (TEST INFRASTRUCTURE ONLY).

Projector calls into ``libxray_cpu.so`` (OpenMP) and an ADMM-TV iteration
built on them plus the NumPy operators of ``oracle.solvers``; timed by
bench.py as the CPU baseline.
"""

from __future__ import annotations

import numpy as np

from . import _cbuild
from . import solvers as S

_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _cbuild.load()
    return _lib


def use_library(path) -> None:
    """Load the C restatement from ``path`` (e.g. a -march=native build)
    for the following calls."""
    global _lib
    _lib = _cbuild.load(path)


def _geom_args(g):
    x0 = np.ascontiguousarray(np.asarray(g["x0"], dtype=np.float64))
    dx = np.ascontiguousarray(np.asarray(g["dx"], dtype=np.float64))
    ang = np.ascontiguousarray(np.asarray(g["angles"], dtype=np.float32))
    return x0, dx, ang


def weights(g):
    M, N = g["nx"]
    A = len(g["angles"])
    x0, dx, ang = _geom_args(g)
    inds = np.empty((A, M, N), np.int32)
    w = np.empty((A, M, N), np.float32)
    lib().cpu_xr2d_weights(M, N, A, x0.ctypes.data, dx.ctypes.data, float(g["y0"]),
                           ang.ctypes.data, inds.ctypes.data, w.ctypes.data)
    return inds, w


def project(im, g):
    im = np.ascontiguousarray(im, dtype=np.float32)
    M, N = im.shape
    A, D = len(g["angles"]), g["ny"]
    x0, dx, ang = _geom_args(g)
    y = np.empty((A, D), np.float32)
    lib().cpu_xr2d_forward(M, N, A, D, x0.ctypes.data, dx.ctypes.data, float(g["y0"]),
                           ang.ctypes.data, im.ctypes.data, y.ctypes.data)
    return y


def back_project(y, g):
    y = np.ascontiguousarray(y, dtype=np.float32)
    M, N = g["nx"]
    A, D = y.shape
    x0, dx, ang = _geom_args(g)
    out = np.empty((M, N), np.float32)
    lib().cpu_xr2d_adjoint(M, N, A, D, x0.ctypes.data, dx.ctypes.data, float(g["y0"]),
                           ang.ctypes.data, y.ctypes.data, out.ctypes.data)
    return out


def project_f64acc(im, g):
    """:func:`project` with float64 accumulation of each bin."""
    im = np.ascontiguousarray(im, dtype=np.float32)
    M, N = im.shape
    A, D = len(g["angles"]), g["ny"]
    x0, dx, ang = _geom_args(g)
    y = np.empty((A, D), np.float32)
    lib().cpu_xr2d_forward_f64acc(M, N, A, D, x0.ctypes.data, dx.ctypes.data, float(g["y0"]),
                                  ang.ctypes.data, im.ctypes.data, y.ctypes.data)
    return y


def back_project_f64acc(y, g):
    """:func:`back_project` with float64 accumulation of each pixel."""
    y = np.ascontiguousarray(y, dtype=np.float32)
    M, N = g["nx"]
    A, D = y.shape
    x0, dx, ang = _geom_args(g)
    out = np.empty((M, N), np.float32)
    lib().cpu_xr2d_adjoint_f64acc(M, N, A, D, x0.ctypes.data, dx.ctypes.data, float(g["y0"]),
                                  ang.ctypes.data, y.ctypes.data, out.ctypes.data)
    return out


def admm_tv(y, g, lam, rho, x0, maxiter, cg_maxiter=10, cg_tol=1e-4):
    """ADMM-TV (oracle.solvers.admm_tv) with the C/OpenMP projector."""
    return S.admm_tv(y, lambda v: project(v, g), lambda v: back_project(v, g), lam, rho, x0,
                     maxiter, cg_tol=cg_tol, cg_maxiter=cg_maxiter)


def xr3d_separable(matrices):
    """(s, ri) if the view matrices (V, 2, 4) rotate about volume axis s:
    detector row ri has coefficient 0 on axis s in every view and row
    1 - ri coefficient 0 on the two other axes; else None."""
    m = np.asarray(matrices, dtype=np.float32)
    for s in range(3):
        others = [a for a in range(3) if a != s]
        for ri in range(2):
            rs = 1 - ri
            if np.all(m[:, ri, s] == 0) and np.all(m[:, rs, others] == 0):
                return s, ri
    return None


def _sep_call(fn, data, matrices, vol_shape, det_shape, f64acc):
    sep = xr3d_separable(matrices)
    if sep is None:
        raise ValueError("view matrices are not separable (no common rotation axis)")
    m = np.ascontiguousarray(np.asarray(matrices, dtype=np.float32))
    X, Y, Z = vol_shape
    D0, D1 = det_shape
    return fn(X, Y, Z, len(m), D0, D1, m.ctypes.data, sep[0], sep[1], int(f64acc), data)


def project_3d_sep(vol, matrices, det_shape, f64acc=False):
    """``oracle.xray.project_3d`` for a separable geometry (C/OpenMP)."""
    vol = np.ascontiguousarray(vol, dtype=np.float32)
    out = np.empty((len(matrices), *det_shape), np.float32)
    rc = _sep_call(lambda *a: lib().cpu_xr3d_sep_forward(*a, out.ctypes.data), vol.ctypes.data,
                   matrices, vol.shape, det_shape, f64acc)
    if rc:
        raise MemoryError("cpu_xr3d_sep_forward failed")
    return out


def back_project_3d_sep(proj, matrices, input_shape, f64acc=False):
    """``oracle.xray.back_project_3d`` for a separable geometry (C/OpenMP)."""
    proj = np.ascontiguousarray(proj, dtype=np.float32)
    out = np.empty(tuple(input_shape), np.float32)
    rc = _sep_call(lambda *a: lib().cpu_xr3d_sep_adjoint(*a, out.ctypes.data), proj.ctypes.data,
                   matrices, input_shape, proj.shape[1:], f64acc)
    if rc:
        raise MemoryError("cpu_xr3d_sep_adjoint failed")
    return out
