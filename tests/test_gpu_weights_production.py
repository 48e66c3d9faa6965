"""Detector bins and weights of the production projector kernels, bit for bit.

``test_gpu_xray2d.py::test_weights_bit_exact`` pins the debug entry point
``scxr_xr2d_weights``; the forward (``scxr_xr2d_forward``) and adjoint
(``scxr_xr2d_adjoint``) kernels compute bins and weights with their own
code (hoisted line terms, bin clamps, the fraction path).  These tests read
the weights out of those kernels themselves (``_xray2d.py:331-349``):

* adjoint: a sinogram that is 1 on every third bin (offset o) of one view
  and 0 elsewhere back projects, at every pixel, to exactly w (its bin
  ``ind`` on the comb), ``1 - w`` (``ind + 1`` on the comb) or 0; the three
  offsets over every view check every (pixel, view) pair's bin modulo 3,
  w and 1 - w.
* forward: an image that is 1 on a few random pixels projects, in a bin
  that exactly one of those pixels' taps reaches, to exactly that tap's
  weight (w or 1 - w), and to exactly 0 in bins no tap reaches; random
  pixel sets are drawn until at least 97 % of all (pixel, view) pairs had
  both in-range taps isolated.

The expected values are the oracle's float32 (ind, w) (``oracle.cpu``,
bitwise the NumPy restatement); the bookkeeping runs in torch on the GPU.
"""

import numpy as np
import pytest
import torch

from oracle import cpu
from oracle import xray as X
from scico_amd.linop.xray import XRayTransform2D

pytestmark = pytest.mark.gpu

# (shape, views, kwargs, adjoint view stride, forward pixel fraction): the
# C1 geometry in full, the C2 geometry on every 37th view (adjoint) and an
# eighth of its pixels (forward), and small geometries whose projected
# coordinates go below 1 (so the adjoint takes its floor path, not the
# fraction path), whose bins fall partly outside the detector, with two or
# four views per forward workgroup, a wide detector (one chain per wave) or
# one pixel per adjoint thread.
GEOMS = [
    ((256, 256), 180, {"det_count": 363}, 1, 1.0),
    ((512, 512), 720, {"det_count": 724}, 37, 0.125),
    ((64, 48), 45, {"det_count": 90, "y0": -40.0}, 1, 1.0),
    ((33, 65), 31, {"det_count": 20}, 1, 1.0),
    ((40, 24), 1100, {}, 9, 1.0),
    ((20, 20), 2100, {"dx": 0.9}, 21, 1.0),
    ((30, 30), 7, {"det_count": 1100}, 1, 1.0),
    ((16, 16), 5, {"dx": 1.8, "det_count": 60}, 1, 1.0),
    ((24, 20), 13, {"x0": (-7.0, -15.0), "dx": (0.8, 1.25)}, 1, 1.0),
    # general-geometry kernels: pixels of 3 detector units (the adjoint tile
    # projects over more than 64 bins) and of 0.02 units (forward lane
    # stride beyond 64)
    ((40, 36), 20, {"dx": 3.0, "det_count": 180}, 1, 1.0),
    ((48, 40), 9, {"dx": 0.02, "det_count": 5}, 1, 1.0),
]


def _setup(shape, nang, kw):
    ang = np.linspace(0, np.pi, nang, endpoint=False)
    H = XRayTransform2D(shape, ang, **kw)
    g = X.xr2d_defaults(shape, ang, x0=kw.get("x0"), dx=kw.get("dx"), y0=kw.get("y0"),
                        det_count=kw.get("det_count"))
    inds, w = cpu.weights(g)
    ri, rw = X.calc_weights(g["x0"], g["dx"], g["nx"], g["angles"][:2], g["y0"])
    assert np.array_equal(inds[:2], ri) and np.array_equal(w[:2], rw)
    w1 = (np.float32(1) - w).astype(np.float32)  # the reference's (1 - weights)
    return H, g, inds, w, w1


def _min_px(g):
    """Smallest projected coordinate over all pixels and views (float64)."""
    vp = np.stack(X.view_params(g["x0"], g["dx"], g["angles"], g["y0"]), axis=1).astype(np.float64)
    M, N = g["nx"]
    ext = np.stack([vp[:, 1] * (M - 1), vp[:, 2] * (N - 1)], 1)
    return float(np.min(vp[:, 0] + np.minimum(ext[:, 0], 0) + np.minimum(ext[:, 1], 0)))


@pytest.mark.parametrize("shape,nang,kw,stride,frac", GEOMS)
def test_adjoint_kernel_weights_bit_exact(gpu, shape, nang, kw, stride, frac):
    H, g, inds, w, w1 = _setup(shape, nang, kw)
    A, D = H.output_shape
    views = np.arange(0, A, stride)
    print(f"adjoint weights {shape} A={nang} {kw}: {len(views)} views, min Px "
          f"{_min_px(g):.2f}", flush=True)
    ti = torch.as_tensor(inds[views], device=gpu).long()
    tw = torch.as_tensor(w[views], device=gpu)
    tw1 = torch.as_tensor(w1[views], device=gpu)
    y = torch.zeros((A, D), device=gpu)
    out = torch.empty((len(views), 3, *shape), device=gpu)
    for k, a in enumerate(views):
        for o in range(3):
            y.zero_()
            y[a, o::3] = 1.0
            H.back_project(y, out=out[k, o])
    bad = 0
    for o in range(3):
        in0 = (ti >= 0) & (ti < D) & (torch.remainder(ti, 3) == o)
        in1 = (ti + 1 >= 0) & (ti + 1 < D) & (torch.remainder(ti + 1, 3) == o)
        exp = torch.where(in0, tw, torch.where(in1, tw1, torch.zeros_like(tw)))
        got = out[:, o]
        bad += int((got.view(torch.int32) != exp.view(torch.int32)).sum())
    print(f"  {len(views) * shape[0] * shape[1]} (pixel, view) pairs x 3 offsets: "
          f"{bad} values differ", flush=True)
    assert bad == 0


def test_adjoint_weight_paths_covered():
    """The geometries above include plans whose adjoint takes the fraction
    path (every projected coordinate >= 1) and plans that take the floor
    path (some below 1)."""
    mins = []
    for shape, nang, kw, _, _ in GEOMS:
        ang = np.linspace(0, np.pi, nang, endpoint=False)
        g = X.xr2d_defaults(shape, ang, x0=kw.get("x0"), dx=kw.get("dx"), y0=kw.get("y0"),
                            det_count=kw.get("det_count"))
        mins.append(_min_px(g))
    assert min(mins) < 1.0 and max(mins) >= 1.0


@pytest.mark.parametrize("shape,nang,kw,stride,frac", GEOMS)
def test_forward_kernel_weights_bit_exact(gpu, shape, nang, kw, stride, frac):
    H, g, inds, w, w1 = _setup(shape, nang, kw)
    A, D = H.output_shape
    M, N = shape
    P = M * N
    rng = np.random.default_rng(0)
    # the pixels under test: all, or a random subset
    pool = np.sort(rng.permutation(P)[: max(1, int(round(frac * P)))])
    Q = len(pool)
    print(f"forward weights {shape} A={nang} {kw}: {Q} of {P} pixels", flush=True)
    ti = torch.as_tensor(inds.reshape(A, P)[:, pool], device=gpu).long()
    tw = torch.as_tensor(w.reshape(A, P)[:, pool], device=gpu)
    tw1 = torch.as_tensor(w1.reshape(A, P)[:, pool], device=gpu)
    del inds, w, w1
    tpool = torch.as_tensor(pool, device=gpu)
    # a few pixels per image: about one per 16 bins of the image's shadow
    span = int(ti.max() - ti.min()) + 2
    K = max(1, min(Q // 4, span // 16))
    B = 256  # images per batch
    covered = torch.zeros((A, Q), dtype=torch.bool, device=gpu)
    im = torch.zeros((B, M, N), device=gpu)
    ys = torch.empty((B, A, D), device=gpu)
    arA = torch.arange(A, device=gpu).view(1, A, 1)
    bad = stray = 0
    nimg = 0
    for rnd in range(12):
        perm = torch.as_tensor(rng.permutation(Q), device=gpu)
        nb = Q // K
        for s in range(0, nb, B):
            bsz = min(B, nb - s)
            pix = perm[s * K:(s + bsz) * K].view(bsz, K)
            im.zero_()
            im.view(B, P)[torch.arange(bsz, device=gpu).view(-1, 1), tpool[pix]] = 1.0
            for b in range(bsz):
                H.project(im[b], out=ys[b])
            nimg += bsz
            ik = ti[:, pix].permute(1, 0, 2)  # (bsz, A, K)
            taps = [(ik, tw[:, pix].permute(1, 0, 2)), (ik + 1, tw1[:, pix].permute(1, 0, 2))]
            cnt = torch.zeros((bsz, A, D), dtype=torch.int32, device=gpu)
            iso = torch.ones_like(ik, dtype=torch.bool)
            for tap, _ in taps:
                ok = (tap >= 0) & (tap < D)
                cnt.scatter_add_(2, tap.clamp(0, D - 1), ok.to(torch.int32))
            y = ys[:bsz]
            for tap, val in taps:
                ok = (tap >= 0) & (tap < D)
                tc = tap.clamp(0, D - 1)
                one = torch.gather(cnt, 2, tc) == 1
                got = torch.gather(y, 2, tc)
                check = ok & one
                bad += int(((got.view(torch.int32) != val.view(torch.int32)) & check).sum())
                iso &= one | ~ok
            # bins no tap reaches hold exactly 0
            stray += int(((cnt == 0) & (y.view(torch.int32) != 0)).sum())
            covered[arA.expand(bsz, A, K), pix.view(bsz, 1, K).expand(bsz, A, K)] |= iso
        cov = float(covered.float().mean())
        if cov >= 0.97:
            break
    print(f"  {nimg} images of {K} pixels: coverage {cov:.4f} of {A * Q} (pixel, view) pairs, "
          f"{bad} weights differ, {stray} stray nonzero bins", flush=True)
    assert bad == 0 and stray == 0
    assert cov >= 0.97
