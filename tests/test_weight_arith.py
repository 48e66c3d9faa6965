"""CPU checks of the arithmetic identities the 2D projector kernels use to
reproduce the reference's float32 weight `min(1 - (Px - ind), width) / width`
(``scico/linop/xray/_xray2d.py:349-351``) bit for bit.

* The quotient: RN32(RN64(d * RN64(1 / w))) == RN32(d / w) for float32
  d in (0, 1] and the per-view widths (the kernels' float64 product).
* The fraction path (taken when every Px >= 1): 1 - (Px - floor(Px)) is
  exact in float32, so the float64 multiply-add rw - f * rw rounds the
  same product once.  Below 1 the subtraction is not exact, which is why
  the host only takes that path for Px >= 1.
"""

import numpy as np

from oracle import xray as X

f32, f64 = np.float32, np.float64


def _widths():
    ang = np.linspace(0, np.pi, 4096, endpoint=False)
    out = []
    for dx in (np.sqrt(2) / 2, 1.0, 0.8125):
        g = X.xr2d_defaults((64, 64), ang, dx=dx)
        out.append(X.view_params(g["x0"], g["dx"], g["angles"], g["y0"])[3])
    return np.concatenate(out).astype(f32)


def test_float64_product_rounds_like_float32_division():
    rng = np.random.default_rng(7)
    w = _widths()
    rw = 1.0 / w.astype(f64)
    for _ in range(4):
        d = (1.0 - rng.random(w.size, dtype=f32)).astype(f32)  # (0, 1]
        ref = d / w  # IEEE float32 division, correctly rounded
        got = (d.astype(f64) * rw).astype(f32)
        np.testing.assert_array_equal(got, ref)
    # d of the fraction path near 1 and near one ulp
    d = np.array([1.0, np.nextafter(f32(1), f32(0)), 2.0 ** -23, 0.5], f32)
    for k in range(w.size):
        np.testing.assert_array_equal((d.astype(f64) * rw[k]).astype(f32), d / w[k])


def test_one_minus_fraction_is_exact_for_px_at_least_one():
    rng = np.random.default_rng(8)
    px = np.concatenate(
        [
            (1.0 + rng.random(200000) * 2000.0).astype(f32),
            np.float32(1.0) + rng.random(20000).astype(f32) * f32(2.0 ** -10),
            np.array([1.0, 1.5, 2.0 ** 23, 2.0 ** 23 + 1, 3e7], f32),
        ]
    )
    assert (px >= 1).all()
    f = (px - np.floor(px)).astype(f32)
    np.testing.assert_array_equal(f.astype(f64), px.astype(f64) - np.floor(px.astype(f64)))
    dd32 = (f32(1) - f).astype(f32)
    np.testing.assert_array_equal(dd32.astype(f64), 1.0 - f.astype(f64))
    # so rw - f rw (one rounding of the exact (1 - f) rw) equals RN64(dd32 * rw)


def test_one_minus_fraction_is_inexact_below_one():
    px = np.array([2.0 ** -30, 0.1, 1e-5], f32)
    f = (px - np.floor(px)).astype(f32)
    dd32 = (f32(1) - f).astype(f32)
    assert (dd32.astype(f64) != 1.0 - f.astype(f64)).any()
