"""Oracle for the parallel-beam X-ray projectors (2D and per-slice 3D).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates, in NumPy, the arithmetic of

* ``scico/linop/xray/_xray2d.py:86-121``  (constructor defaults),
* ``scico/linop/xray/_xray2d.py:309-351`` (``_calc_weights``),
* ``scico/linop/xray/_xray2d.py:223-265`` (``_project``, scatter-add),
* ``scico/linop/xray/_xray2d.py:267-306`` (``_back_project``, gather),
* ``scico/linop/xray/_xray2d.py:158-221`` (``fbp``, ``_ramp_filter``),
* ``scico/linop/xray/_xray3d.py:206-266`` (3D ``_calc_weights``),
* ``scico/linop/xray/_xray3d.py:110-204`` (3D project / back project),
* ``scico/linop/xray/_xray3d.py:268-327`` (``matrices_from_euler_angles``).

The fp32 functions follow the reference's operation order with every
operation rounded to float32 (JAX default, ``jax_enable_x64`` off) and no
FMA contraction; trigonometry is the correctly rounded float32 of a
float64 evaluation.  The ``*_f64`` twins evaluate the same model in
float64 and serve as the FP64 reference for error reporting.

How each function is pinned (tests/test_oracle_*.py):

* ``view_params`` / ``calc_weights``: brute-force float64 interval
  overlap on tiny images (an independent formulation of the pixel
  boxcar), invariants (weights in (0, 1], the two bin weights sum to 1),
  the closed form at angle 0 with unit pixels (bins = image rows, the
  reference docstring ``_xray2d.py:66-70``) and angle pi/2 (columns).
* ``project`` / ``back_project``: dense float64 matrix assembled from the
  brute-force overlap; adjointness ``<Ax, y> = <x, A^T y>`` (the
  reference's ``test_xray_2d.py:52-85``); row/column-sum closed forms;
  mass conservation for a constant image.
* 3D: the reference's known-answer arrays ``test_xray_3d.py:29-69`` and
  the dot test ``test_xray_3d.py:9-26``.
* ``fbp``: the reference's PSNR > 28 dB pin ``test_xray_2d.py:88-102``.
"""

from __future__ import annotations

import numpy as np

f32 = np.float32


# ---------------------------------------------------------------- 2D --


def xr2d_defaults(input_shape, angles, x0=None, dx=None, y0=None, det_count=None):
    """Constructor defaults, ``_xray2d.py:86-121``.

    Returns float32 scalars as they enter the jitted kernels (x64 off).
    """
    nx = tuple(int(s) for s in input_shape)
    if dx is None:
        dx = 2 * (np.sqrt(2) / 2,)
    if np.ndim(dx) == 0:
        dx = 2 * (dx,)
    if x0 is None:
        x0 = -(np.array(nx) * np.asarray(dx, dtype=np.float64)) / 2
    if det_count is None:
        det_count = int(np.ceil(np.linalg.norm(input_shape)))
    if y0 is None:
        y0 = -det_count / 2
    return dict(
        nx=nx,
        x0=np.asarray(x0, dtype=np.float64).astype(f32),
        dx=np.asarray(dx, dtype=np.float64).astype(f32),
        y0=f32(y0),
        ny=int(det_count),
        angles=np.asarray(angles).astype(f32),
    )


def view_params(x0, dx, angles, y0):
    """Per-view geometry, ``_xray2d.py:326-347``, float32.

    Returns ``(Pxmin, P0, P1, width)`` arrays of shape (A,).
    """
    th = np.asarray(angles, dtype=f32)
    u0 = np.cos(th.astype(np.float64)).astype(f32)
    u1 = np.sin(th.astype(np.float64)).astype(f32)
    x00, x01 = f32(x0[0]), f32(x0[1])
    dx0, dx1 = f32(dx[0]), f32(dx[1])
    Px0 = (x00 * u0 + x01 * u1) - f32(y0)  # :327
    P0 = dx0 * u0  # :328
    P1 = dx1 * u1
    c = np.stack([Px0, Px0 + P0, Px0 + P1, (Px0 + P0) + P1])  # :329
    Pxmin = c.min(axis=0)
    diag1 = np.abs(P0 + P1)  # :342
    diag2 = np.abs(P0 - P1)  # :343
    w = np.maximum(diag1, diag2)  # :344
    f = np.minimum(diag1, diag2)  # :345
    width = (w + f) / f32(2)  # :347
    return Pxmin.astype(f32), P0.astype(f32), P1.astype(f32), width.astype(f32)


def calc_weights(x0, dx, nx, angles, y0, views=None):
    """``_calc_weights`` (``_xray2d.py:309-351``) for the selected views.

    Returns ``inds`` int32 (A', M, N) and ``weights`` float32 (A', M, N).
    """
    Pxmin, P0, P1, width = view_params(x0, dx, angles, y0)
    if views is not None:
        Pxmin, P0, P1, width = Pxmin[views], P0[views], P1[views], width[views]
    i = np.arange(nx[0], dtype=np.int32).astype(f32).reshape(1, -1, 1)
    j = np.arange(nx[1], dtype=np.int32).astype(f32).reshape(1, 1, -1)
    e = lambda v: v.reshape(-1, 1, 1)
    Px = (e(Pxmin) + e(P0) * i) + e(P1) * j  # :331-335
    inds = np.floor(Px).astype(np.int32)  # :338
    dist = f32(1) - (Px - inds.astype(f32))  # :348
    weights = np.minimum(dist, e(width)) / e(width)  # :349
    return inds, weights.astype(f32)


def _view_chunks(A, n_pix, budget=1 << 24):
    step = max(1, budget // max(1, n_pix))
    for a0 in range(0, A, step):
        yield np.arange(a0, min(A, a0 + step))


def project(im, x0, dx, y0, ny, angles):
    """Forward projection, ``_xray2d.py:248-265``, float32 scatter-add.

    The scatter visits updates in row-major (view, row, col) order, first
    for bin ``ind`` and then for bin ``ind + 1`` (XLA CPU scatter order).
    """
    im = np.asarray(im, dtype=f32)
    nx = im.shape
    A = len(angles)
    y = np.zeros((A, ny), dtype=f32)
    for vs in _view_chunks(A, im.size):
        inds, w = calc_weights(x0, dx, nx, angles, y0, views=vs)
        a = np.broadcast_to(vs.reshape(-1, 1, 1), inds.shape)
        ok = (inds >= 0) & (inds < ny)  # :255
        yc = y[vs[0] : vs[-1] + 1]
        np.add.at(yc, (a[ok] - vs[0], inds[ok]), (im * w)[ok])  # :256-259
        ok1 = (inds + 1 >= 0) & (inds + 1 < ny)  # :262
        np.add.at(yc, (a[ok1] - vs[0], inds[ok1] + 1), (im * (f32(1) - w))[ok1])  # :263
    return y


def back_project(y, x0, dx, nx, y0, angles):
    """Back projection (exact adjoint), ``_xray2d.py:292-306``, float32.

    Gather-then-sum over views for bin ``ind`` and, separately, ``ind+1``.
    Out-of-range bins contribute zero (the reference gathers a clamped
    element and multiplies by a zero weight).
    """
    y = np.asarray(y, dtype=f32)
    ny = y.shape[1]
    A = len(angles)
    s0 = np.zeros(nx, dtype=f32)
    s1 = np.zeros(nx, dtype=f32)
    for vs in _view_chunks(A, nx[0] * nx[1]):
        inds, w = calc_weights(x0, dx, nx, angles, y0, views=vs)
        a = vs.reshape(-1, 1, 1)
        w0 = np.where((inds >= 0) & (inds < ny), w, f32(0))  # :295
        g0 = y[a, np.clip(inds, 0, ny - 1)]
        s0 += np.sum(g0 * w0, axis=0, dtype=f32)  # :299
        w1 = np.where((inds + 1 >= 0) & (inds + 1 < ny), f32(1) - w, f32(0))  # :301
        g1 = y[a, np.clip(inds + 1, 0, ny - 1)]
        s1 += np.sum(g1 * w1, axis=0, dtype=f32)  # :302-304
    return (s0 + s1).astype(f32)


# float64 twins: same model, every quantity in float64.


def view_params_f64(x0, dx, angles, y0):
    th = np.asarray(angles, dtype=np.float64)
    u0, u1 = np.cos(th), np.sin(th)
    x0 = np.asarray(x0, dtype=np.float64)
    dx = np.asarray(dx, dtype=np.float64)
    Px0 = x0[0] * u0 + x0[1] * u1 - float(y0)
    P0, P1 = dx[0] * u0, dx[1] * u1
    Pxmin = np.min(np.stack([Px0, Px0 + P0, Px0 + P1, Px0 + P0 + P1]), axis=0)
    d1, d2 = np.abs(P0 + P1), np.abs(P0 - P1)
    width = (np.maximum(d1, d2) + np.minimum(d1, d2)) / 2
    return Pxmin, P0, P1, width


def _weights_f64(x0, dx, nx, angles, y0, views):
    Pxmin, P0, P1, width = (v[views] for v in view_params_f64(x0, dx, angles, y0))
    i = np.arange(nx[0], dtype=np.float64).reshape(1, -1, 1)
    j = np.arange(nx[1], dtype=np.float64).reshape(1, 1, -1)
    e = lambda v: v.reshape(-1, 1, 1)
    Px = e(Pxmin) + e(P0) * i + e(P1) * j
    inds = np.floor(Px).astype(np.int64)
    w = np.minimum(1.0 - (Px - inds), e(width)) / e(width)
    return inds, w


def project_f64(im, x0, dx, y0, ny, angles):
    im = np.asarray(im, dtype=np.float64)
    A = len(angles)
    y = np.zeros((A, ny), dtype=np.float64)
    for vs in _view_chunks(A, im.size):
        inds, w = _weights_f64(x0, dx, im.shape, angles, y0, vs)
        a = np.broadcast_to((vs - vs[0]).reshape(-1, 1, 1), inds.shape)
        yc = y[vs[0] : vs[-1] + 1]
        for ind, wt in ((inds, w), (inds + 1, 1.0 - w)):
            ok = (ind >= 0) & (ind < ny)
            flat = a[ok] * ny + ind[ok]
            yc += np.bincount(flat, weights=(im * wt)[ok], minlength=yc.size).reshape(yc.shape)
    return y


def back_project_f64(y, x0, dx, nx, y0, angles):
    y = np.asarray(y, dtype=np.float64)
    ny = y.shape[1]
    out = np.zeros(nx, dtype=np.float64)
    for vs in _view_chunks(len(angles), nx[0] * nx[1]):
        inds, w = _weights_f64(x0, dx, nx, angles, y0, vs)
        a = vs.reshape(-1, 1, 1)
        for ind, wt in ((inds, w), (inds + 1, 1.0 - w)):
            ok = (ind >= 0) & (ind < ny)
            out += np.sum(np.where(ok, y[a, np.clip(ind, 0, ny - 1)] * wt, 0.0), axis=0)
    return out


def dense_matrix_overlap(x0, dx, nx, angles, y0, ny):
    """Brute-force float64 system matrix from interval overlaps.

    Independent of ``calc_weights``' floor/min arithmetic: pixel (i, j)
    in view a is a boxcar of height 1/width on ``[Px, Px + width)`` with
    ``Px`` its lowest projected corner (the float32 value of
    ``_xray2d.py:331-335``); bin b covers ``[b, b + 1)``.  The entry is
    the boxcar's integral over the bin.  Valid while ``width <= 1``.
    Returns an ``(A*ny, M*N)`` float64 array.
    """
    Pxmin, P0, P1, width = view_params(x0, dx, angles, y0)
    A = len(angles)
    H = np.zeros((A * ny, nx[0] * nx[1]), dtype=np.float64)
    for a in range(A):
        for i in range(nx[0]):
            for j in range(nx[1]):
                Px = float((Pxmin[a] + P0[a] * f32(i)) + P1[a] * f32(j))
                wd = float(width[a])
                for b in range(int(np.floor(Px)) - 1, int(np.floor(Px + wd)) + 2):
                    if 0 <= b < ny:
                        ov = min(b + 1.0, Px + wd) - max(float(b), Px)
                        if ov > 0:
                            H[a * ny + b, i * nx[1] + j] += ov / wd
    return H


def ramp_filter(x, tau):
    """``_xray2d.py:199-221`` (Kak & Slaney eq. 61), float64."""
    x = np.asarray(x)
    out = np.zeros(x.shape, dtype=np.float64)
    out[x == 0] = 1.0 / (4.0 * tau**2)
    odd = (x % 2) != 0
    out[odd] = -1.0 / (x[odd].astype(np.float64) ** 2 * np.pi**2 * tau**2)
    return out


def fbp(y, x0, dx, nx, y0, angles):
    """Filtered back projection, ``_xray2d.py:158-197`` (float32 result)."""
    y = np.asarray(y, dtype=f32)
    A, N = y.shape
    nvec = np.arange(N) - (N - 1) // 2
    h = ramp_filter(nvec, 1.0).astype(f32).reshape(1, -1)
    mask = back_project(np.ones_like(y), x0, dx, nx, y0, angles) >= f32(A * (1.0 - 1e-5))
    L = 2 * N - 1
    hf = np.fft.fft(h.astype(np.complex64), n=L, axis=1)
    yf = np.fft.fft(y.astype(np.complex64), n=L, axis=1)
    hy = np.fft.ifft(hf * yf, n=L, axis=1)[:, (N - 1) // 2 : -(N - 1) // 2].real.astype(f32)  # :192-194
    scale = f32(np.pi * float(dx[0]) * float(dx[1]) / A)
    return (scale * mask * back_project(hy, x0, dx, nx, y0, angles)).astype(f32)


# ---------------------------------------------------------------- 3D --


def matrices_from_euler_angles(input_shape, output_shape, seq, angles, degrees=False,
                               voxel_spacing=None, det_spacing=None):
    """``_xray3d.py:268-327`` (float64; cast to float32 by the operator)."""
    from scipy.spatial.transform import Rotation

    if voxel_spacing is None:
        voxel_spacing = np.ones(3)
    if det_spacing is None:
        det_spacing = np.ones(2)
    m = Rotation.from_euler(seq, angles, degrees=degrees).as_matrix()[:, :2, :]
    m = m * np.asarray(voxel_spacing, dtype=np.float64).reshape(1, 1, 3)
    m = (1 / np.asarray(det_spacing, dtype=np.float64)).reshape(1, 2, 1) * m
    t = -np.einsum("vmn,n->vm", m, np.array(input_shape) / 2) + np.array(output_shape) / 2
    return np.concatenate((m, t[..., None]), axis=2)


def calc_weights_3d(input_shape, matrix, det_shape, slice_offset=0):
    """``_xray3d.py:206-266`` for one view, float32."""
    m = np.asarray(matrix, dtype=f32)
    g = np.mgrid[: input_shape[0], : input_shape[1], : input_shape[2]].astype(np.int32)
    x = g.astype(f32) + f32(0.5)  # :211
    x[0] = x[0] + f32(slice_offset)  # :212
    Px = np.stack(
        [((m[r, 0] * x[0] + m[r, 1] * x[1]) + m[r, 2] * x[2]) + m[r, 3] for r in range(2)]
    )  # :214-219
    w = f32(0.5)
    left = Px - w / f32(2)  # :223
    t = np.minimum(np.ceil(left) - left, w)  # :224
    ul = np.floor(left).astype(np.int32)  # :225
    s = f32(4.0)  # 1 / w**2
    wul = (t[0] * t[1]) * s  # :227-230
    wur = ((w - t[0]) * t[1]) * s
    wll = (t[0] * (w - t[1])) * s
    wlr = ((w - t[0]) * (w - t[1])) * s

    def inb(d0, d1):
        return ((ul[0] + d0 >= 0) & (ul[0] + d0 < det_shape[0])
                & (ul[1] + d1 >= 0) & (ul[1] + d1 < det_shape[1]))

    z = f32(0)
    wul = np.where(inb(0, 0), wul, z)  # :233-264
    wur = np.where(inb(1, 0), wur, z)
    wll = np.where(inb(0, 1), wll, z)
    wlr = np.where(inb(1, 1), wlr, z)
    return ul, wul, wur, wll, wlr


def project_3d(vol, matrices, det_shape):
    """``_xray3d.py:110-159``: per view, four scatter-adds (mode="drop")."""
    vol = np.asarray(vol, dtype=f32)
    out = np.zeros((len(matrices), *det_shape), dtype=f32)
    for v, mat in enumerate(np.asarray(matrices, dtype=f32)):
        ul, wul, wur, wll, wlr = calc_weights_3d(vol.shape, mat, det_shape)
        p = out[v]
        for (d0, d1), wt in (((0, 0), wul), ((1, 0), wur), ((0, 1), wll), ((1, 1), wlr)):
            i0, i1 = ul[0] + d0, ul[1] + d1
            ok = (i0 >= 0) & (i0 < det_shape[0]) & (i1 >= 0) & (i1 < det_shape[1])
            np.add.at(p, (i0[ok], i1[ok]), (wt * vol)[ok])
    return out


def back_project_3d(proj, matrices, input_shape):
    """``_xray3d.py:161-204``: scan over views, four gathers each."""
    proj = np.asarray(proj, dtype=f32)
    det_shape = proj.shape[1:]
    out = np.zeros(input_shape, dtype=f32)
    for v, mat in enumerate(np.asarray(matrices, dtype=f32)):
        ul, wul, wur, wll, wlr = calc_weights_3d(input_shape, mat, det_shape)
        y = proj[v]
        for (d0, d1), wt in (((0, 0), wul), ((1, 0), wur), ((0, 1), wll), ((1, 1), wlr)):
            i0 = np.clip(ul[0] + d0, 0, det_shape[0] - 1)
            i1 = np.clip(ul[1] + d1, 0, det_shape[1] - 1)
            out = out + y[i0, i1] * wt
    return out
