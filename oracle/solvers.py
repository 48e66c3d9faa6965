"""Oracle for the operators, functionals and solvers around the projector.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates in NumPy (float32 unless noted, reference operation order):

* finite differences  ``scico/linop/_diff.py:236-272`` (append=0,
  circular) and their adjoint (``scico/linop/_stack.py:71-72``);
* L1 / L21 norms and proxes ``scico/functional/_norm.py:65-116, 208-296``
  with ``no_nan_divide`` ``scico/numpy/util.py:302-315``;
* the Haar-based TV prox ``scico/functional/_tvnorm.py:117-236, 355-538``;
* ``SquaredL2Loss`` ``scico/loss.py:154-262``;
* ``cg`` ``scico/solver.py:326-405``;
* ADMM with ``LinearSubproblemSolver`` ``scico/optimize/_admm.py:96-378``,
  ``scico/optimize/_admmaux.py:160-269`` and its iteration statistics;
* ``AcceleratedPGM`` ``scico/optimize/_pgm.py:60-252`` (constant step);
* ``ProximalADMM`` ``scico/optimize/_padmm.py:171-361`` and the identity
  ``SquaredL2Loss`` prox ``scico/loss.py:220-226``;
* ``CircularConvolve`` ``scico/linop/_circconv.py:79-209``.

Pins (tests/test_oracle_solvers.py): the reference's explicit FD outputs
(``test_diff.py:11-63``), the L21 closed form (``test_norm.py:10-28``),
the Haar identities ``W^T W = 2I / 4I`` (``test_tvnorm.py:13-23``), prox
optimality against brute-force minimisation on tiny inputs, CG against a
direct solve (``test_solver.py:28-114``), the ADMM / APGM fixed-point
conditions and the quadratic closed forms (``test_admm.py:111-258``,
``test_pgm.py:21-120``), proximal ADMM converging to the closed-form
minimiser of a quadratic problem (``test_padmm.py:180-200``), CircularConvolve against a direct circular sum
(``test_circconv.py:35-56``) and ``numpy.fft`` in float64.
"""

from __future__ import annotations

import numpy as np

f32 = np.float32


# ------------------------------------------------------------- operators --


def fd(x, circular=False):
    """``FiniteDifference(shape, append=0 | circular=True)`` over all axes."""
    x = np.asarray(x, dtype=f32)
    out = []
    for ax in range(x.ndim):
        sl = [slice(None)] * x.ndim
        sl[ax] = slice(0, 1) if circular else slice(-1, None)
        out.append(np.diff(x, axis=ax, append=x[tuple(sl)]))
    return np.stack(out).astype(f32)


def fd_adj(g, circular=False):
    """Adjoint of :func:`fd`: sum over axes of the per-axis transposes."""
    g = np.asarray(g, dtype=f32)
    total = None
    for ax in range(g.shape[0]):
        ga = g[ax]
        if circular:
            t = np.roll(ga, 1, axis=ax) - ga
        else:
            n = ga.shape[ax]
            t = np.zeros_like(ga)
            take = lambda a, i: np.take(a, i, axis=ax)
            idx = [slice(None)] * ga.ndim

            def put(i, v):
                idx[ax] = i
                t[tuple(idx)] = v

            put(0, -take(ga, 0))
            if n > 2:
                idx[ax] = slice(1, n - 1)
                lo = [slice(None)] * ga.ndim
                lo[ax] = slice(0, n - 2)
                hi = [slice(None)] * ga.ndim
                hi[ax] = slice(1, n - 1)
                t[tuple(idx)] = ga[tuple(lo)] - ga[tuple(hi)]
            if n > 1:
                put(n - 1, take(ga, n - 2))
        total = t if total is None else (total + t).astype(f32)
    return total.astype(f32)


# ------------------------------------------------------------ functionals --


def no_nan_divide(x, y):
    return np.where(y != 0, np.divide(x, np.where(y != 0, y, f32(1))), f32(0)).astype(f32)


def l21_norm(x):
    """``L21Norm()(x)`` with the l2 norm over axis 0 (``_norm.py:246-250``)."""
    x = np.asarray(x, dtype=np.float64)
    return float(np.sum(np.abs(np.sqrt(np.sum(x**2, axis=0)))))


def l21_prox(v, lam):
    """``L21Norm._prox`` over axis 0 (``_norm.py:254-263``), float32."""
    v = np.asarray(v, dtype=f32)
    lam = f32(lam)
    length = np.sqrt(np.sum(np.abs(v) ** 2, axis=0, keepdims=True, dtype=f32)).astype(f32)
    direction = no_nan_divide(v, length)
    nl = length - lam
    nl = f32(0.5) * (nl + np.abs(nl))
    return (nl * direction).astype(f32)


def l1_norm(x):
    return float(np.sum(np.abs(np.asarray(x, dtype=np.float64))))


def l1_prox(v, lam):
    """``L1Norm.prox`` (``_norm.py:110-116``), float32."""
    v = np.asarray(v, dtype=f32)
    tmp = np.abs(v) - f32(lam)
    tmp = f32(0.5) * (tmp + np.abs(tmp))
    return (np.sign(v) * tmp).astype(f32)


def haar(x, axes=None):
    """Single-level shift-invariant Haar transform ``W`` (``_tvnorm.py:486-538``).

    Returns ``(ndim, 2, *shape)``: low-pass ``s*(x + roll(x,-1))``,
    high-pass ``s*(roll(x,-1) - x)`` per axis, ``s = 1/sqrt(2)``.
    """
    x = np.asarray(x, dtype=f32)
    s = f32(1.0 / np.sqrt(2.0))
    axes = range(x.ndim) if axes is None else axes
    out = []
    for ax in axes:
        r = np.roll(x, -1, axis=ax)
        out.append(np.stack([s * (x + r), s * (r - x)]))
    return np.stack(out).astype(f32)


def haar_adj(w, axes=None):
    """Adjoint of :func:`haar`."""
    w = np.asarray(w, dtype=f32)
    s = f32(1.0 / np.sqrt(2.0))
    axes = range(w.shape[0]) if axes is None else axes
    total = None
    for k, ax in enumerate(axes):
        lo, hi = w[k, 0], w[k, 1]
        t = s * (lo + np.roll(lo, 1, axis=ax)) + s * (np.roll(hi, 1, axis=ax) - hi)
        total = t if total is None else total + t
    return total.astype(f32)


def tv_prox(v, lam, isotropic=True, circular=False):
    """``TVNorm.prox`` (``_tvnorm.py:183-236``), all axes, float32.

    Non-circular: edge-pad +1 on each axis, Haar, shrink the high-pass
    restricted to the unpadded region with threshold ``sqrt(2)*K*lam``
    (``K = 2*ndim``), adjoint Haar, crop, times ``1/K``.
    """
    v = np.asarray(v, dtype=f32)
    nd = v.ndim
    K = 2 * nd
    thr = f32(np.sqrt(2.0)) * f32(K) * f32(lam)
    if circular:
        w = haar(v)
        hp = w[:, 1]
        w[:, 1] = l21_prox(hp, thr) if isotropic else l1_prox(hp, thr)
        return (f32(1.0 / K) * haar_adj(w)).astype(f32)
    vp = np.pad(v, [(0, 1)] * nd, mode="edge")
    w = haar(vp)
    inner = (slice(None), 1) + tuple(slice(0, -1) for _ in range(nd))
    hp = w[inner]
    w[inner] = l21_prox(hp, thr) if isotropic else l1_prox(hp, thr)
    u = haar_adj(w)[tuple(slice(0, -1) for _ in range(nd))]
    return (f32(1.0 / K) * u).astype(f32)


def tv_prox_weighted(v, lam, weight=None, isotropic=True, circular=False, dtype=f32):
    """:func:`tv_prox` with a per-pixel threshold map ``thr * weight`` on the
    high-pass coefficients at each unpadded location (``weight=None`` is the
    reference prox, ``_tvnorm.py:183-236``).  ``dtype=np.float64`` gives the
    float64 twin used for finite-difference Jacobians.
    """
    dt = np.dtype(dtype)
    v = np.asarray(v, dtype=dt)
    nd = v.ndim
    K = 2 * nd
    s = dt.type(1.0 / np.sqrt(2.0))
    if dt == np.float32:
        thr = f32(np.sqrt(2.0)) * f32(K) * f32(lam)
    else:
        thr = np.sqrt(2.0) * K * float(lam)
    tmap = thr if weight is None else (dt.type(thr) * np.asarray(weight, dtype=dt)).astype(dt)

    def W(x):
        out = []
        for ax in range(nd):
            r = np.roll(x, -1, axis=ax)
            out.append(np.stack([s * (x + r), s * (r - x)]))
        return np.stack(out).astype(dt)

    def WT(w):
        total = None
        for ax in range(nd):
            lo, hi = w[ax, 0], w[ax, 1]
            t = s * (lo + np.roll(lo, 1, axis=ax)) + s * (np.roll(hi, 1, axis=ax) - hi)
            total = t if total is None else total + t
        return total.astype(dt)

    def shrink(hp):
        if isotropic:
            ln = np.sqrt(np.sum(hp * hp, axis=0))
            nl = dt.type(0.5) * ((ln - tmap) + np.abs(ln - tmap))
            with np.errstate(invalid="ignore", divide="ignore"):
                d = np.where(ln != 0, hp / np.where(ln != 0, ln, 1), 0)
            return (nl * d).astype(dt)
        t = np.abs(hp) - tmap
        return (np.sign(hp) * (dt.type(0.5) * (t + np.abs(t)))).astype(dt)

    if circular:
        w = W(v)
        w[:, 1] = shrink(w[:, 1])
        return (dt.type(1.0 / K) * WT(w)).astype(dt)
    vp = np.pad(v, [(0, 1)] * nd, mode="edge")
    w = W(vp)
    inner = (slice(None), 1) + tuple(slice(0, -1) for _ in range(nd))
    w[inner] = shrink(w[inner])
    u = WT(w)[tuple(slice(0, -1) for _ in range(nd))]
    return (dt.type(1.0 / K) * u).astype(dt)


def unrolled_apgm_tv(y, A, AT, L, lam, weight, K, x0, dtype=f32):
    """``K`` AcceleratedPGM steps for ``0.5 ||A x - y||^2 + lam TV_w(x)`` with
    the weighted Haar TV prox (:func:`tv_prox_weighted`), constant step
    ``1/L`` (``_pgm.py:221-240``); the momentum sequence is the float32
    one of ``AcceleratedPGM``.  Returns ``x_K``."""
    dt = np.dtype(dtype)
    x = np.asarray(x0, dtype=dt)
    v = x
    t = 1.0
    gam = 1.0 / L
    for _ in range(K):
        x_old = x
        u = (v - dt.type(gam) * AT((A(v) - y).astype(dt))).astype(dt)
        x = tv_prox_weighted(u, gam * lam, weight, dtype=dt)
        t_old = t
        t = float(f32(0.5) * (f32(1) + np.sqrt(f32(1) + f32(4) * f32(t_old) ** 2)))
        c = float(f32(f32(t_old) - f32(1)) / f32(t))
        v = (x + dt.type(c) * (x - x_old)).astype(dt)
    return x


def tv_norm(x, isotropic=True, circular=False):
    """``TVNorm.__call__`` = norm(FD(x)) (``_tvnorm.py:89-115``)."""
    g = fd(x, circular=circular)
    return l21_norm(g) if isotropic else l1_norm(g)


# -------------------------------------------------------------------- cg --


def cg(A, b, x0, tol=1e-5, atol=0.0, maxiter=1000):
    """``scico.solver.cg`` without preconditioner (``solver.py:378-403``).

    Reductions are taken in float64 and stored as float32 scalars.
    """
    x = np.asarray(x0, dtype=f32)
    b = np.asarray(b, dtype=f32)
    Ax = A(x)
    bn = f32(np.sqrt(np.sum(b.astype(np.float64) ** 2)))
    r = (b - Ax).astype(f32)
    p = r
    num = f32(np.sum(r.astype(np.float64) ** 2))
    ii = 0
    tol2 = f32(max(tol * bn, atol)) ** 2
    while ii < maxiter and num > tol2:
        Ap = A(p)
        alpha = f32(num / f32(np.sum(p.astype(np.float64) * Ap)))
        x = (x + alpha * p).astype(f32)
        r = (r - alpha * Ap).astype(f32)
        num_old = num
        num = f32(np.sum(r.astype(np.float64) ** 2))
        beta = f32(num / num_old)
        p = (r + beta * p).astype(f32)
        ii += 1
    return x, {"num_iter": ii, "rel_res": float(np.sqrt(num) / bn)}


# ------------------------------------------------------------------ ADMM --


def admm_tv(y, A, AT, lam, rho, x0, maxiter, cg_tol=1e-4, cg_maxiter=100, circular=False,
            isotropic=True, z0=None, u0=None, state=False):
    """ADMM for ``0.5*||y - A x||^2 + lam*||FD x||_{2,1 or 1}``.

    Mirrors ``ADMM(f=SquaredL2Loss(y, A), g_list=[lam*L21Norm()],
    C_list=[FiniteDifference(append=0)], rho_list=[rho],
    subproblem_solver=LinearSubproblemSolver(cg_kwargs))``:
    ``_admm.py:297-332`` (init), ``:364-378`` (step),
    ``_admmaux.py:222-269`` (x-step), ``_admm.py:241-295`` (itstat).
    Returns ``(x, history)`` with history rows
    ``(Iter, Objective, Prml Rsdl, Dual Rsdl, CG It, CG Res)``.
    """
    C = lambda v: fd(v, circular)
    CT = lambda g: fd_adj(g, circular)
    prox = l21_prox if isotropic else l1_prox
    gnorm = l21_norm if isotropic else l1_norm
    rho32 = f32(rho)
    x = np.asarray(x0, dtype=f32)
    z = C(x) if z0 is None else np.asarray(z0, dtype=f32)
    u = np.zeros_like(z) if u0 is None else np.asarray(u0, dtype=f32)
    lhs = lambda v: (rho32 * CT(C(v)) + AT(A(v))).astype(f32)
    hist = []
    for it in range(maxiter):
        rhs = (AT(y) + rho32 * CT((z - u).astype(f32))).astype(f32)
        x, info = cg(lhs, rhs, x, tol=cg_tol, maxiter=cg_maxiter)
        z_old = z
        Cx = C(x)
        z = prox((Cx + u).astype(f32), f32((1.0 / rho) * lam))
        u = ((u + Cx) - z).astype(f32)
        r = (y - A(x)).astype(np.float64)
        obj = 0.5 * float(np.sum(r * r)) + lam * gnorm(z)
        prml = float(np.sqrt(rho * np.sum((C(x).astype(np.float64) - z) ** 2)))
        dual = float(np.linalg.norm((rho32 * CT((z - z_old).astype(f32))).astype(np.float64)))
        hist.append((it, obj, prml, dual, info["num_iter"], info["rel_res"]))
    if state:
        return x, hist, z, u
    return x, hist


# ------------------------------------------------------------------ APGM --


def apgm(grad_f, prox_g, L0, x0, maxiter, objective=None):
    """``AcceleratedPGM`` with constant step ``L0`` (``_pgm.py:221-240``).

    Returns ``(x, residuals)``, or ``(x, residuals, objectives)`` with
    ``objective(x)`` evaluated after every iteration, as the reference's
    iteration statistics do (``_common.py:306-312``, ``_pgm.py:151-156``).
    """
    x = np.asarray(x0, dtype=f32)
    v = x
    t = f32(1.0)
    L = f32(L0)
    res, obj = [], []
    for _ in range(maxiter):
        x_old = x
        x = prox_g((v - f32(1.0 / L) * grad_f(v)).astype(f32), f32(1.0 / L))
        res.append(float(np.linalg.norm((x - v).astype(np.float64))))
        if objective is not None:
            obj.append(float(objective(x)))
        # the momentum sequence is float32 (jax_enable_x64 off): t_old ** 2,
        # 4 *, 1 +, sqrt, 1 +, 0.5 * and (t_old - 1) / t each round
        t_old = t
        t = f32(0.5) * (f32(1) + np.sqrt(f32(1) + f32(4) * (t_old * t_old)))
        c = (t_old - f32(1)) / t
        v = (x + c * (x - x_old)).astype(f32)
    return (x, res) if objective is None else (x, res, obj)


# -------------------------------------------------------- proximal ADMM --


def _bl(op, *xs):
    """Apply ``op`` blockwise to lists of arrays (a plain array is one block)."""
    if isinstance(xs[0], list):
        return [op(*t) for t in zip(*xs)]
    return op(*xs)


def padmm(f_prox, g_prox, A, AT, x0, z0, u0, rho, mu, nu, maxiter, B=None, BT=None, c=None,
          dtype=f32):
    """``ProximalADMM`` iterations (``scico/optimize/_padmm.py:348-361``).

    ``z`` and ``u`` may be lists of arrays (block arrays); ``A`` returns and
    ``AT`` takes such lists.  ``B = None`` means ``B = -I``; ``c = None``
    means zero.  Returns ``(x, z, u)``.
    """
    cast = lambda a: np.asarray(a, dtype=dtype)
    add = lambda a, b: _bl(lambda p, q: (p + q).astype(dtype), a, b)
    sub = lambda a, b: _bl(lambda p, q: (p - q).astype(dtype), a, b)
    scl = lambda s, a: _bl(lambda p: (dtype(s) * p).astype(dtype), a)
    Bz = (lambda z: scl(-1.0, z)) if B is None else B
    BH = (lambda w: scl(-1.0, w)) if BT is None else BT

    def constraint(Ax, z):
        r = add(Ax, Bz(z))
        return r if c is None else sub(r, c)

    x = cast(x0)
    z = _bl(cast, z0)
    u = _bl(cast, u0)
    u_old = u
    for _ in range(maxiter):
        proxarg = sub(x, scl(1.0 / mu, AT(sub(scl(2.0, u), u_old))))
        x = f_prox(proxarg, 1.0 / (rho * mu))
        Ax = A(x)
        proxarg = sub(z, scl(1.0 / nu, BH(add(constraint(Ax, z), u))))
        z = g_prox(proxarg, 1.0 / (rho * nu))
        u_old = u
        u = add(u, constraint(Ax, z))
    return x, z, u


def sql2_identity_prox(v, lam, y, scale=0.5, dtype=f32):
    """``SquaredL2Loss(y, A=Identity).prox`` (``loss.py:220-226``):
    ``(c y + v) / (c + 1)`` with ``c = 2 scale lam``."""
    c = dtype(2.0 * scale * lam)
    return ((c * np.asarray(y, dtype) + np.asarray(v, dtype)) / (c + dtype(1))).astype(dtype)


# ---------------------------------------------------------------- phantoms --


def circular_phantom(shape, radius_list, val_list):
    """``scico.examples.create_circular_phantom`` (``examples.py:384-436``):
    concentric discs about the array centre ``(dim - 1) / 2``, later
    entries overwriting earlier ones where ``dist < r`` (float64)."""
    center = [(d - 1) / 2 for d in shape]
    mesh = np.meshgrid(*[np.arange(d) for d in shape], sparse=True, indexing="ij")
    dist = np.sqrt(sum((mesh[i] - center[i]) ** 2 for i in range(len(shape))))
    img = np.zeros(shape)
    for r, v in zip(radius_list, val_list):
        img[dist < r] = v
    return img


def tv_denoise_problem(ndim):
    """``(x_gt, y)`` of the reference's TV denoising tests: ``test_tvnorm.py:61-72``
    (ndim 2: 32^2 circular phantom plus a ramp) and ``:26-34`` (ndim 1:
    128-sample clipped sine), noise sigma 0.02.  The reference draws the
    noise from ``jax.random`` (absent here); this draws it from NumPy's
    generator, seed 0."""
    if ndim == 2:
        N = 32
        x_gt = circular_phantom((N, N), [0.6 * N, 0.4 * N, 0.2 * N, 0.1 * N],
                                [0.25, 1, 0, 0.5]).astype(f32)
        gr, gc = np.ogrid[0:N, 0:N]
        x_gt += ((gr + gc) / (4 * N)).astype(f32)
    else:
        g = np.linspace(0, 2 * np.pi, 128, dtype=f32)
        x_gt = np.sin(2 * g)
        x_gt[x_gt > 0.5] = 0.5
        x_gt[x_gt < -0.5] = -0.5
    noise = np.random.default_rng(0).standard_normal(x_gt.shape).astype(f32)
    return x_gt, (x_gt + f32(0.02) * noise).astype(f32)


def tv_denoise_pair(y, lam, isotropic, circular, admm_its, L0, apgm_its, rho=1e1,
                    thr_factor=1.0):
    """The two solutions the reference's TV tests compare
    (``test_tvnorm.py:74-120``): ADMM (x0 = y, CG tol 1e-4 / maxiter 25)
    with ``lam * L21 | L1`` of ``FiniteDifference``, and AcceleratedPGM
    (x0 = y, constant ``L0``) with the Haar TV prox.  ``thr_factor`` scales
    the prox threshold (1 for the reference)."""
    I = lambda v: v
    x_admm, hist = admm_tv(y, I, I, lam, rho, y, admm_its, cg_tol=1e-4, cg_maxiter=25,
                           circular=circular, isotropic=isotropic)
    grad = lambda v: (v - y).astype(f32)
    prox = lambda v, s: tv_prox(v, thr_factor * lam * s, isotropic=isotropic, circular=circular)
    x_apgm, _ = apgm(grad, prox, L0, y, apgm_its)
    return x_admm, x_apgm, hist


# ------------------------------------------------------- CircularConvolve --


def dft_center_shift(shape, center):
    """``CircularConvolve._dft_center_shift`` (``_circconv.py:155-186``), complex64."""
    offset = -np.asarray(center, dtype=np.float64).reshape(-1)
    factors = []
    for k, s in zip(offset, shape):
        n = np.arange(s)
        f = np.select(
            [n < s / 2, n == s / 2, n > s / 2],
            [np.exp(-1j * k * 2 * np.pi * n / s), np.cos(k * np.pi) + 0j,
             np.exp(1j * k * 2 * np.pi * (s - n) / s)],
        ).astype(np.complex64)
        factors.append(f)
    shift = factors[0].reshape(-1, *([1] * (len(shape) - 1)))
    for d, f in enumerate(factors[1:], start=1):
        shift = shift * f.reshape([-1 if i == d else 1 for i in range(len(shape))])
    return shift.astype(np.complex64)


def circconv_hdft(h, shape, center=None):
    """``h_dft = fftn(h, s=shape)`` (complex64) times the centring shift."""
    hd = np.fft.fftn(np.asarray(h, dtype=np.complex64), s=shape).astype(np.complex64)
    if center is not None:
        hd = (hd * dft_center_shift(shape, center)).astype(np.complex64)
    return hd


def circconv(x, h_dft):
    """``CircularConvolve._eval`` (``_circconv.py:188-197``), complex64 FFTs."""
    xd = np.fft.fftn(np.asarray(x, dtype=np.complex64)).astype(np.complex64)
    return np.fft.ifftn((h_dft * xd).astype(np.complex64)).real.astype(f32)


def circconv_adj(y, h_dft):
    """``CircularConvolve._adj`` (``_circconv.py:199-209``)."""
    yd = np.fft.fftn(np.asarray(y, dtype=np.complex64)).astype(np.complex64)
    return np.fft.ifftn((np.conj(h_dft) * yd).astype(np.complex64)).real.astype(f32)


def circconv_f64(x, h, center=None):
    """Float64 FFT convolution (FP64 reference)."""
    shape = np.asarray(x).shape
    hd = np.fft.fftn(np.asarray(h, dtype=np.float64), s=shape)
    if center is not None:
        hd = hd * dft_center_shift(shape, center).astype(np.complex128)
    return np.fft.ifftn(hd * np.fft.fftn(np.asarray(x, dtype=np.float64))).real


def circconv_direct(x, h, center=None):
    """Direct circular sum ``y[n] = sum_m h[m] x[n - m + c]`` (float64)."""
    x = np.asarray(x, dtype=np.float64)
    h = np.asarray(h, dtype=np.float64)
    c = np.zeros(x.ndim, dtype=int) if center is None else np.asarray(center, dtype=int)
    y = np.zeros_like(x)
    for m in np.ndindex(*h.shape):
        y += h[m] * np.roll(x, shift=tuple(np.asarray(m) - c), axis=tuple(range(x.ndim)))
    return y
