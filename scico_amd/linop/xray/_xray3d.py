"""3D parallel-beam X-ray transform on the HIP projector kernels.

Drop-in for ``scico.linop.xray.XRayTransform3D``
(``scico/linop/xray/_xray3d.py:27-327``): same constructor, attributes,
``project`` / ``back_project`` and ``matrices_from_euler_angles``.  The
projections run in ``scxr_xr3d_forward`` / ``scxr_xr3d_adjoint``.

Single-rotation-axis geometries (for example ``seq="Y"`` with the default
spacings) run in slice mode: every volume slice along the rotation axis
maps onto one detector column, and the kernels process all slices of a
(pixel, view) pair together.  Other geometries run in general mode.
"""

from __future__ import annotations

import ctypes
import weakref

import numpy as np

from ... import _device as dev
from ... import _pred
from ..._lib import call, load
from .._linop import LinearOperator


class XRayTransform3D(LinearOperator):
    r"""General-purpose 3D parallel-ray X-ray projector.

    View ``v`` is a float32 ``(2, 4)`` matrix ``[M | t]``: the centre of
    voxel ``(i, j, k)`` projects to ``M (i + 1/2, j + 1/2, k + 1/2) + t``
    in detector coordinates, and the voxel splats onto the 2 x 2 detector
    pixels under a half-pixel box around that point.  The back projection
    is the exact adjoint.
    """

    def __init__(self, input_shape, matrices, det_shape, batch_size: int = 8,
                 input_dtype=np.float32, input_device=None, output_device=None):
        # _xray3d.py:54-93
        self.input_shape = tuple(int(s) for s in input_shape)
        if len(self.input_shape) != 3:
            raise ValueError(f"input_shape must have three dimensions, got {self.input_shape}.")
        m = np.asarray(dev.to_numpy(matrices), dtype=np.float32)
        if m.ndim != 3 or m.shape[1:] != (2, 4):
            raise ValueError(f"matrices must have shape (num_views, 2, 4), got {m.shape}.")
        self.matrices = np.ascontiguousarray(m)
        self.det_shape = tuple(int(d) for d in det_shape)
        if len(self.det_shape) != 2:
            raise ValueError(f"det_shape must have two entries, got {self.det_shape}.")
        self.batch_size = batch_size
        self.output_shape = (len(m), *self.det_shape)
        self.input_device = input_device
        self.output_device = output_device
        if np.dtype(input_dtype) != np.float32:
            raise ValueError("only float32 volumes are supported")
        self._plan = None
        self._allow_slice = True
        self._stage_adjoint = True
        super().__init__(
            input_shape=self.input_shape,
            output_shape=self.output_shape,
            eval_fn=self.project,
            adj_fn=self.back_project,
            input_dtype=np.float32,
            output_dtype=np.float32,
        )

    @property
    def plan(self) -> ctypes.c_void_p:
        if self._plan is None:
            dev.require_gpu()
            lib = load()
            X, Y, Z = self.input_shape
            h = ctypes.c_void_p()
            call(
                "scxr_xr3d_plan_create", X, Y, Z, self.output_shape[0],
                self.matrices.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                self.det_shape[0], self.det_shape[1],
                int(self._allow_slice) | (0 if self._stage_adjoint else 2), ctypes.byref(h),
            )
            self._plan = h
            self._finalizer = weakref.finalize(self, lib.scxr_xr3d_plan_destroy, h)
        return self._plan

    def info(self) -> dict:
        """Execution mode chosen for this geometry."""
        out = np.zeros(9, dtype=np.int32)
        call("scxr_xr3d_info", self.plan, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        return {"mode": "slice" if out[0] == 1 else "general", "slice_axis": int(out[1]),
                "det_axis": int(out[2]), "delta": int(out[3]), "candidates_per_row": int(out[4]),
                "volume_in_place": bool(out[5]), "projection_in_place": bool(out[6]),
                "in_plane_bins": int(out[7]), "adjoint_staged": bool(out[8])}

    def project(self, im, out=None, pred=None):
        """X-ray projection (``_xray3d.py:95-100``)."""
        im = dev.input_array(im, self.input_shape, "im")
        out = dev.output_buffer(out, self.output_shape)
        call("scxr_xr3d_forward", self.plan, im.data_ptr(), out.data_ptr(),
             dev.ptr(_pred.resolve(pred)), dev.stream())
        return out

    def back_project(self, proj, out=None, pred=None):
        """Back projection (``_xray3d.py:102-108``)."""
        proj = dev.input_array(proj, self.output_shape, "proj")
        out = dev.output_buffer(out, self.input_shape)
        call("scxr_xr3d_adjoint", self.plan, proj.data_ptr(), out.data_ptr(),
             dev.ptr(_pred.resolve(pred)), dev.stream())
        return out

    def weights(self, view: int):
        """Debug: ``(ul (2, X, Y, Z) int32, w (4, X, Y, Z))`` of one view,
        as ``_calc_weights`` returns them (``_xray3d.py:206-266``)."""
        if not 0 <= int(view) < self.output_shape[0]:
            raise ValueError(f"view {view} out of range [0, {self.output_shape[0]})")
        ul = dev.empty((2, *self.input_shape), dtype=np.int32)
        w = dev.empty((4, *self.input_shape))
        call("scxr_xr3d_weights", self.plan, int(view), ul.data_ptr(), w.data_ptr(), dev.stream())
        return ul, w

    @staticmethod
    def matrices_from_euler_angles(input_shape, output_shape, seq: str, angles, degrees=False,
                                   voxel_spacing=None, det_spacing=None) -> np.ndarray:
        """Projection matrices from Euler angles (``_xray3d.py:268-327``).

        The rotation's first two rows, columns scaled by the voxel spacing
        and rows by the inverse detector spacing, and the translation that
        takes the volume centre to the detector centre; float64,
        ``(num_views, 2, 4)``.
        """
        from scipy.spatial.transform import Rotation

        vs = np.ones(3) if voxel_spacing is None else np.asarray(voxel_spacing, dtype=np.float64)
        ds = np.ones(2) if det_spacing is None else np.asarray(det_spacing, dtype=np.float64)
        rot = Rotation.from_euler(seq, angles, degrees=degrees).as_matrix()[:, :2, :]
        rows = rot * vs[None, None, :]
        rows = (1 / ds)[None, :, None] * rows
        centre = np.asarray(input_shape, dtype=np.float64) / 2
        t = np.asarray(output_shape, dtype=np.float64) / 2 - np.einsum("vmn,n->vm", rows, centre)
        return np.concatenate((rows, t[..., None]), axis=2)
