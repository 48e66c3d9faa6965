"""Public wrappers refuse inputs and out= buffers of the wrong size, dtype,
device or layout before any pointer reaches the C ABI (the kernels trust
sizes), and a batched solver refuses a plan re-created after it was
built."""

import numpy as np
import pytest
import torch

from scico_amd import phantom
from scico_amd.batch import ADMMTVBatch, XRayTransform2DBatch
from scico_amd.linop import CircularConvolve, FiniteDifference
from scico_amd.linop.xray import XRayTransform2D, XRayTransform3D

pytestmark = pytest.mark.gpu


def _bad_buffers(shape, gpu):
    n = int(np.prod(shape))
    return [
        torch.empty(n + 1, dtype=torch.float32, device=gpu),          # one element too many
        torch.empty(max(n - 1, 1), dtype=torch.float32, device=gpu),  # one too few
        torch.empty(shape, dtype=torch.float64, device=gpu),          # wrong dtype
        torch.empty(shape, dtype=torch.float32),                      # host memory
        torch.empty((shape[-1], n // shape[-1]), dtype=torch.float32, device=gpu).t(),  # strided
    ]


def test_xray2d_refuses_bad_buffers(gpu):
    H = XRayTransform2D((16, 12), phantom.angles(5))
    x = torch.zeros(H.input_shape, device=gpu)
    y = torch.zeros(H.output_shape, device=gpu)
    for bad in _bad_buffers(H.output_shape, gpu):
        with pytest.raises(ValueError):
            H.project(x, out=bad)
    for bad in _bad_buffers(H.input_shape, gpu):
        with pytest.raises(ValueError):
            H.back_project(y, out=bad)
    with pytest.raises(ValueError):
        H.project(torch.zeros((16, 13), device=gpu))
    with pytest.raises(ValueError):
        H.back_project(torch.zeros((5, H.det_count + 1), device=gpu))
    with pytest.raises(ValueError):
        H.fbp(torch.zeros((4, H.det_count), device=gpu))
    # a conforming out= buffer of another shape with the same element count is fine
    out = torch.empty(int(np.prod(H.output_shape)), device=gpu)
    H.project(x, out=out)


def test_xray3d_fd_cconv_refuse_bad_buffers(gpu):
    shape = (6, 5, 4)
    mats = XRayTransform3D.matrices_from_euler_angles(shape, (7, 5), "Y", phantom.angles(3))
    H = XRayTransform3D(shape, mats, (7, 5))
    v = torch.zeros(shape, device=gpu)
    for bad in _bad_buffers(H.output_shape, gpu):
        with pytest.raises(ValueError):
            H.project(v, out=bad)
    with pytest.raises(ValueError):
        H.back_project(torch.zeros((3, 7, 6), device=gpu))
    C = FiniteDifference(shape, append=0)
    for bad in _bad_buffers(C.output_shape, gpu):
        with pytest.raises(ValueError):
            C._eval(v, out=bad)
    with pytest.raises(ValueError):
        C._adj(torch.zeros((2, *shape), device=gpu))
    h = np.ones((3, 3), np.float32) / 9
    K = CircularConvolve(h, (16, 16))
    img = torch.zeros((16, 16), device=gpu)
    for bad in _bad_buffers((16, 16), gpu):
        with pytest.raises(ValueError):
            K._eval(img, out=bad)
    with pytest.raises(ValueError):
        K._eval(torch.zeros((16, 15), device=gpu))


def test_batch_refuses_bad_buffers_and_stale_plans(gpu):
    shapes = [(24, 24), (16, 20)]
    angs = [phantom.angles(7), phantom.angles(5)]
    B = XRayTransform2DBatch(shapes, angs)
    xb = torch.zeros(B.n_img, device=gpu)
    yb = torch.zeros(B.n_sino, device=gpu)
    with pytest.raises(ValueError):
        B.project(torch.zeros(B.n_img + 1, device=gpu))
    with pytest.raises(ValueError):
        B.project(xb, out=torch.zeros(B.n_sino - 1, device=gpu))
    with pytest.raises(ValueError):
        B.project(xb, y=yb)  # residual partials need their workspace
    with pytest.raises(ValueError):
        B.back_project(yb, kind=1)  # the CG operator needs p
    with pytest.raises(ValueError):
        B.back_project(yb, kind=2, z=torch.zeros(B.n_img, device=gpu))
    with pytest.raises(ValueError):
        B.back_project(yb, dot_ws=torch.zeros(B.n_tiles - 1, dtype=torch.float64, device=gpu))
    solver = ADMMTVBatch(B, yb, 2e-2)
    solver.enqueue()
    B.vpg = 2 if B.vpg == 1 else 1
    B._create()
    with pytest.raises(RuntimeError):
        solver.enqueue()
    torch.cuda.synchronize()
