"""GPU parity of the 3D projector (slice and general modes) against the
oracle, through the C ABI.

Bar: ul indices and the four weights bit-exact; projections and back
projections within rel-L2 1e-5 of the float32 oracle; dot test <= 1e-5;
the reference's known-answer arrays (test_xray_3d.py:29-69).  The full
256^3 / 360-view case is checked slice by slice against the oracle on thin
volumes (slices are independent for a single rotation axis) plus the dot
test and per-view mass.
"""

import json
import os

import numpy as np
import pytest
import torch

from oracle import xray as X
from scico_amd import phantom
from scico_amd.linop import valid_adjoint
from scico_amd.linop.xray import XRayTransform3D
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _mats(shape, det, seq, ang, **kw):
    return XRayTransform3D.matrices_from_euler_angles(shape, det, seq, ang, **kw)


# (shape, det, seq, angles, kwargs, expected mode)
GEOMS = [
    ((16, 16, 16), (24, 16), "Y", np.linspace(0, np.pi, 12, endpoint=False), {}, "slice"),
    ((9, 7, 5), (14, 7), "Y", np.linspace(0, np.pi, 7, endpoint=False), {}, "slice"),
    ((8, 6, 10), (14, 10), "Y", np.linspace(0, np.pi, 5, endpoint=False), {}, "slice"),  # det != S
    ((6, 11, 13), (12, 18), "X", np.linspace(0, np.pi, 6, endpoint=False)[:, None], {}, "slice"),
    ((6, 11, 13), (11, 18), "X", np.linspace(0, np.pi, 6, endpoint=False)[:, None], {}, "general"),
    ((12, 10, 8), (16, 14), "Z", np.linspace(0, np.pi, 6, endpoint=False), {}, "general"),
    ((8, 6, 9), (13, 9), "Y", np.linspace(0, np.pi, 4, endpoint=False), {}, "general"),  # t_d = 1/4
    ((10, 9, 8), (16, 16), "XYZ",
     np.random.default_rng(0).uniform(0, np.pi, (5, 3)), {}, "general"),
    ((8, 8, 8), (12, 12), "Y", np.linspace(0, np.pi, 5, endpoint=False),
     {"voxel_spacing": [1.0, 1.0, 1.5]}, "slice"),
]


def _op(shape, det, seq, ang, kw, allow_slice=True):
    M = _mats(shape, det, seq, ang, **kw)
    H = XRayTransform3D(shape, M, det)
    H._allow_slice = allow_slice
    return H, H.matrices


@pytest.mark.parametrize("g", range(len(GEOMS)))
def test_mode_selection(gpu, g):
    shape, det, seq, ang, kw, mode = GEOMS[g]
    H, _ = _op(shape, det, seq, ang, kw)
    assert H.info()["mode"] == mode


@pytest.mark.parametrize("allow_slice", [True, False])
@pytest.mark.parametrize("g", range(len(GEOMS)))
def test_weights_bit_exact(gpu, g, allow_slice):
    shape, det, seq, ang, kw, _ = GEOMS[g]
    H, M = _op(shape, det, seq, ang, kw, allow_slice)
    for v in range(len(M)):
        ul, w = H.weights(v)
        rul, *rw = X.calc_weights_3d(shape, M[v], det)
        np.testing.assert_array_equal(ul.cpu().numpy(), rul)
        np.testing.assert_array_equal(w.cpu().numpy(), np.stack(rw))


@pytest.mark.parametrize("allow_slice", [True, False])
@pytest.mark.parametrize("g", range(len(GEOMS)))
def test_forward_adjoint_parity(gpu, g, allow_slice):
    shape, det, seq, ang, kw, _ = GEOMS[g]
    H, M = _op(shape, det, seq, ang, kw, allow_slice)
    print(f"xr3d {shape} {det} {seq} mode={H.info()['mode']}", flush=True)
    rng = np.random.default_rng(g)
    x = rng.random(shape).astype(np.float32)
    y = rng.standard_normal(H.output_shape).astype(np.float32)
    e_f = rel_l2(H(x).cpu().numpy(), X.project_3d(x, M, det))
    e_a = rel_l2(H.T(y).cpu().numpy(), X.back_project_3d(y, M, shape))
    print(f"  rel-L2 fwd {e_f:.2e} adj {e_a:.2e}", flush=True)
    assert e_f <= 1e-5 and e_a <= 1e-5
    assert valid_adjoint(H, H.T, eps=1e-5)


def test_reference_kat(gpu):
    """test_xray_3d.py:29-69 known-answer arrays, both modes."""
    kat = json.load(open(os.path.join(GOLD, "reference_kat.json")))["xray3d_scaling"]
    x = np.zeros(kat["input_shape"], np.float32)
    s = kat["ones_slice"]
    x[s[0][0] : s[0][1], s[1][0] : s[1][1], s[2][0] : s[2][1]] = 1
    for case in kat["cases"]:
        M = XRayTransform3D.matrices_from_euler_angles(
            x.shape, x.shape[:2], kat["seq"], kat["angles"], voxel_spacing=case["voxel_spacing"])
        for allow in (True, False):
            H = XRayTransform3D(x.shape, matrices=M, det_shape=x.shape[:2])
            H._allow_slice = allow
            np.testing.assert_allclose(H(x).cpu().numpy(), np.array(case["truth"]), atol=1e-6)


def test_matched_adjoint(gpu):
    """test_xray_3d.py:9-26."""
    N = 16
    dc = int(N * 1.05 / np.sqrt(2.0))
    M = XRayTransform3D.matrices_from_euler_angles(
        (N, N, N), (dc, dc), "X", np.linspace(0, np.pi, 3, endpoint=False)[:, None])
    H = XRayTransform3D((N, N, N), matrices=M, det_shape=(dc, dc))
    assert valid_adjoint(H, H.T, eps=1e-5)


def test_determinism_slice_mode(gpu):
    shape, det = (32, 24, 40), (48, 24)
    H, _ = _op(shape, det, "Y", np.linspace(0, np.pi, 30, endpoint=False), {})
    x = torch.rand(shape, device="cuda")
    y1, y2 = H(x), H(x)
    b1, b2 = H.T(y1), H.T(y1)
    assert torch.equal(y1, y2) and torch.equal(b1, b2)


def test_full_size_256_360(gpu):
    """256^3 volume, 360 views about axis 1, detector 368 x 256: slices
    {0, 77, 255} against the oracle on thin volumes, per-view mass, dot
    test."""
    N, V, det = 256, 360, (368, 256)
    ang = phantom.angles(V)
    M = XRayTransform3D.matrices_from_euler_angles((N, N, N), det, "Y", ang)
    H = XRayTransform3D((N, N, N), M, det)
    assert H.info()["mode"] == "slice" and H.info()["projection_in_place"]
    vol = phantom.ellipsoid_phantom((N, N, N), 30, seed=0)
    y = H(vol)
    yn = y.cpu().numpy()
    sub = np.arange(0, V, 18)
    Mt = X.matrices_from_euler_angles((N, 1, N), (det[0], 1), "Y", ang)
    rng = np.random.default_rng(3)
    yr = rng.standard_normal((V, *det)).astype(np.float32)
    bp = H.T(yr).cpu().numpy()
    for j in (0, 77, 255):
        thin = np.ascontiguousarray(vol[:, j : j + 1, :])
        ref = X.project_3d(thin, Mt[sub].astype(np.float32), (det[0], 1))[..., 0]
        e = rel_l2(yn[sub, :, j], ref)
        ref_bp = X.back_project_3d(np.ascontiguousarray(yr[:, :, j : j + 1]),
                                   Mt.astype(np.float32), (N, 1, N))[:, 0, :]
        eb = rel_l2(bp[:, j, :], ref_bp)
        print(f"  256^3/360 slice {j}: fwd rel-L2 {e:.2e}, adj rel-L2 {eb:.2e}", flush=True)
        assert e <= 1e-5 and eb <= 1e-5
    # every footprint is on the detector: each view carries the volume's mass
    sums = y.double().sum(dim=(1, 2)).cpu().numpy()
    np.testing.assert_allclose(sums, np.full(V, vol.astype(np.float64).sum()), rtol=1e-5)
    assert valid_adjoint(H, H.T, eps=1e-5)


STAGED_CASES = [g for g in range(len(GEOMS)) if GEOMS[g][5] == "slice"] + ["full", "slab"]


@pytest.mark.parametrize("case", STAGED_CASES)
def test_staged_adjoint_bitwise(gpu, case):
    """The slice adjoint that copies each view's detector rows for a 4 x 8
    pixel tile into LDS equals the one that reads them from global memory,
    bit for bit (same rows, same accumulation order); plans whose tile
    range can exceed the staging buffer keep the global-memory kernel."""
    if case == "full":
        shape, det, seq, ang, kw = (256, 256, 256), (368, 256), "Y", phantom.angles(360), {}
    elif case == "slab":
        shape, det, seq, ang, kw = (256, 16, 256), (368, 16), "Y", phantom.angles(360), {}
    else:
        shape, det, seq, ang, kw, _ = GEOMS[case]
    H, _ = _op(shape, det, seq, ang, kw)
    Hg, _ = _op(shape, det, seq, ang, kw)
    Hg._stage_adjoint = False
    assert H.info()["mode"] == "slice" and not Hg.info()["adjoint_staged"]
    y = torch.randn(H.output_shape, device="cuda",
                    generator=torch.Generator("cuda").manual_seed(11))
    staged = H.info()["adjoint_staged"]
    print(f"  staged adjoint {shape} {det} views {len(ang)}: staged={staged}", flush=True)
    if case in ("full", "slab"):
        assert staged
    assert torch.equal(H.T(y), Hg.T(y))
