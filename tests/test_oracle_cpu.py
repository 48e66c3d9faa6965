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


@pytest.mark.parametrize("seq,shape,det", [("Y", (10, 7, 12), (16, 7)), ("X", (6, 11, 9), (6, 15)),
                                           ("Y", (8, 5, 8), (12, 7))])
def test_c_separable_3d_matches_numpy_oracle(seq, shape, det):
    """The C restatement of the 3D projector for rotations about one volume
    axis equals the NumPy oracle (pinned to the reference's known-answer
    arrays) bit for bit, forward and adjoint; its float64-accumulation
    twin agrees to float32 rounding."""
    ang = np.linspace(0, np.pi, 9, endpoint=False)
    M = X.matrices_from_euler_angles(shape, det, seq, ang).astype(np.float32)
    sep = cpu.xr3d_separable(M)
    if sep is None:
        pytest.skip(f"rotation {seq} about a volume axis gives no separable matrices here")
    rng = np.random.default_rng(1)
    vol = rng.random(shape).astype(np.float32)
    y = rng.standard_normal((len(ang), *det)).astype(np.float32)
    np.testing.assert_array_equal(cpu.project_3d_sep(vol, M, det), X.project_3d(vol, M, det))
    np.testing.assert_array_equal(cpu.back_project_3d_sep(y, M, shape),
                                  X.back_project_3d(y, M, shape))
    assert rel_l2(cpu.project_3d_sep(vol, M, det, f64acc=True), X.project_3d(vol, M, det)) < 1e-6
    assert rel_l2(cpu.back_project_3d_sep(y, M, shape, f64acc=True),
                  X.back_project_3d(y, M, shape)) < 1e-6
