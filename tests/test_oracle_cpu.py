"""The C/OpenMP restatement agrees with the NumPy oracle (CPU)."""

import numpy as np
import pytest

from oracle import cpu
from oracle import xray as X
from tests.conftest import rel_l2


@pytest.mark.parametrize("shape,nang,kw", [((12, 13), 10, {}), ((40, 33), 25, {"det_count": 30}),
                                           ((64, 64), 90, {"dx": 0.5})])
def test_c_restatement_matches_numpy_oracle(shape, nang, kw):
    ang = np.linspace(0, np.pi, nang, endpoint=False)
    g = X.xr2d_defaults(shape, ang, **kw)
    ci, cw = cpu.weights(g)
    ri, rw = X.calc_weights(g["x0"], g["dx"], g["nx"], g["angles"], g["y0"])
    np.testing.assert_array_equal(ci, ri)
    np.testing.assert_array_equal(cw, rw)
    rng = np.random.default_rng(0)
    x = rng.random(shape).astype(np.float32)
    y = rng.standard_normal((nang, g["ny"])).astype(np.float32)
    yc = cpu.project(x, g)
    ry = X.project(x, g["x0"], g["dx"], g["y0"], g["ny"], g["angles"])
    np.testing.assert_array_equal(yc, ry)  # same scatter order
    assert rel_l2(cpu.back_project(y, g),
                  X.back_project(y, g["x0"], g["dx"], g["nx"], g["y0"], g["angles"])) < 1e-6
