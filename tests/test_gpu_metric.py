"""scico_amd.metric against the formulas of scico/metric.py:20-163 in numpy."""

import numpy as np
import pytest

from scico_amd import metric

pytestmark = pytest.mark.gpu


def test_metrics_vs_numpy(gpu):
    rng = np.random.default_rng(0)
    r = rng.standard_normal((64, 48)).astype(np.float32)
    c = (r + 0.1 * rng.standard_normal(r.shape)).astype(np.float32)
    d = (r + 0.3 * rng.standard_normal(r.shape)).astype(np.float32)
    r64, c64, d64 = (a.astype(np.float64) for a in (r, c, d))
    mse = np.mean((r64 - c64) ** 2)
    assert metric.mae(r, c) == pytest.approx(np.mean(np.abs(r64 - c64)), rel=1e-6)
    assert metric.mse(r, c) == pytest.approx(mse, rel=1e-6)
    assert metric.snr(r, c) == pytest.approx(10 * np.log10(np.var(r64) / mse), rel=1e-6)
    rng_ = r64.max() - r64.min()
    assert metric.psnr(r, c) == pytest.approx(10 * np.log10(rng_**2 / mse), rel=1e-6)
    assert metric.psnr(r, c, signal_range=255) == pytest.approx(10 * np.log10(255**2 / mse),
                                                                rel=1e-6)
    assert metric.isnr(r, d, c) == pytest.approx(
        10 * np.log10(np.mean((r64 - d64) ** 2) / mse), rel=1e-6)
    assert metric.bsnr(r, c) == pytest.approx(
        10 * np.log10(np.var(r64) / np.var(c64 - r64)), rel=1e-5)
    nrm = max(np.linalg.norm(c64), np.linalg.norm(r64))
    assert metric.rel_res(c, r) == pytest.approx(np.linalg.norm(r64 - c64) / nrm, rel=1e-6)
    assert metric.rel_res(np.zeros(4, np.float32), np.zeros(4, np.float32)) == 0.0
    assert metric.snr(r, r) == np.inf
