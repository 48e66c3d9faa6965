"""2D parallel-beam X-ray transform on the HIP projector kernels.

Drop-in for ``scico.linop.xray.XRayTransform2D``
(``scico/linop/xray/_xray2d.py:29-351``): same constructor, attributes,
warning and ``project`` / ``back_project`` / ``fbp`` methods.  The
projections run in ``scxr_xr2d_forward`` / ``scxr_xr2d_adjoint``.
"""

from __future__ import annotations

import ctypes
import weakref
from typing import Optional
from warnings import warn

import numpy as np

from ... import _device as dev
from ... import _pred
from ..._lib import call, load
from .._linop import LinearOperator


def is_scalar_equiv(s) -> bool:
    """``scico.numpy.util.is_scalar_equiv``: scalar or single-element array."""
    return np.isscalar(s) or (hasattr(s, "shape") and np.asarray(s).size == 1)


class XRayTransform2D(LinearOperator):
    r"""Parallel ray, single axis, 2D X-ray projector.

    Pixels project as boxcars of width ``width`` (per view) starting at
    the pixel's lowest projected corner; detector bins have unit width.
    Each pixel contributes to at most two bins while the projected pixel
    width is at most 1 (default pixel spacing :math:`\sqrt{2}/2`).
    """

    def __init__(
        self,
        input_shape,
        angles,
        x0=None,
        dx=None,
        y0: Optional[float] = None,
        det_count: Optional[int] = None,
        input_device=None,
        output_device=None,
    ):
        # _xray2d.py:86-121
        self.input_shape = tuple(int(s) for s in input_shape)
        if len(self.input_shape) != 2:
            raise ValueError(f"input_shape must have two dimensions, got {self.input_shape}.")
        self.angles = angles
        self.nx = tuple(self.input_shape)
        if dx is None:
            dx = 2 * (np.sqrt(2) / 2,)
        if is_scalar_equiv(dx):
            dx = 2 * (float(np.asarray(dx).reshape(-1)[0]),)
        self.dx = dx

        ang32 = np.asarray(dev.to_numpy(angles), dtype=np.float32).reshape(-1)
        self._angles32 = ang32
        dx32 = np.asarray([float(np.asarray(d)) for d in dx], dtype=np.float64).astype(np.float32)
        u0 = np.cos(ang32.astype(np.float64)).astype(np.float32)
        u1 = np.sin(ang32.astype(np.float64)).astype(np.float32)
        P0, P1 = dx32[0] * u0, dx32[1] * u1
        max_width = float(np.max(np.maximum(np.abs(P0 + P1), np.abs(P0 - P1))))
        if max_width > 1:
            warn(f"A projected pixel has width {max_width} > 1.0, which will reduce projector accuracy.")

        if x0 is None:
            x0 = -(np.array(self.nx) * np.asarray([float(np.asarray(d)) for d in self.dx])) / 2
        self.x0 = x0
        if det_count is None:
            det_count = int(np.ceil(np.linalg.norm(self.input_shape)))
        self.det_count = int(det_count)
        self.ny = self.det_count
        self.output_shape = (len(ang32), self.det_count)
        if y0 is None:
            y0 = -self.ny / 2
        self.y0 = y0
        self.dy = 1.0
        self.fbp_filter = None
        self.fbp_mask = None
        self.input_device = input_device
        self.output_device = output_device
        self._plan = None
        super().__init__(
            input_shape=self.input_shape,
            input_dtype=np.float32,
            output_shape=self.output_shape,
            output_dtype=np.float32,
            eval_fn=self.project,
            adj_fn=self.back_project,
        )

    # --- native plan (created on first use; holds the per-view tables)
    @property
    def plan(self) -> ctypes.c_void_p:
        if self._plan is None:
            dev.require_gpu()
            lib = load()
            x0 = np.asarray([float(np.asarray(v)) for v in np.asarray(self.x0).reshape(-1)],
                            dtype=np.float64)
            dx = np.asarray([float(np.asarray(v)) for v in self.dx], dtype=np.float64)
            ang = np.ascontiguousarray(self._angles32)
            h = ctypes.c_void_p()
            call(
                "scxr_xr2d_plan_create",
                self.nx[0], self.nx[1], len(ang), self.ny,
                x0.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                dx.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                float(self.y0),
                ang.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                ctypes.byref(h),
            )
            self._plan = h
            self._finalizer = weakref.finalize(self, lib.scxr_xr2d_plan_destroy, h)
        return self._plan

    def launch_info(self) -> dict:
        """Launch facts of the plan (``scxr_xr2d_info``): forward and adjoint
        workgroups (0 when that direction runs the general-geometry gather,
        which has no fused epilogues), views per forward workgroup, adjoint
        window width."""
        out = (ctypes.c_int32 * 4)()
        call("scxr_xr2d_info", self.plan, out)
        return {"forward_groups": int(out[0]), "adjoint_tiles": int(out[1]),
                "views_per_group": int(out[2]), "adjoint_window": int(out[3])}

    def set_forward_split(self, parts: int = 0) -> int:
        """Divide each forward group's image lines over ``parts`` workgroups
        (1..8; 0 = the automatic choice) and return the count in effect.
        The parts' partial sinograms are added in a fixed order; a split
        changes float32 rounding only.  With 1 the forward is bitwise equal
        to the batched projector's for the same slice.  Call it before
        capturing a graph that projects with this operator (a captured
        launch keeps its grid, and a larger split replaces the workspace)."""
        eff = ctypes.c_int32(0)
        call("scxr_xr2d_set_forward_split", self.plan, int(parts), ctypes.byref(eff))
        return int(eff.value)

    @property
    def tiled(self) -> bool:
        """Both directions run the tiled kernels (with fused epilogues)."""
        i = self.launch_info()
        return i["forward_groups"] > 0 and i["adjoint_tiles"] > 0

    def view_table(self) -> np.ndarray:
        """Per-view float32 (Pxmin, P0, P1, width) as computed by the plan."""
        out = np.empty((self.output_shape[0], 4), dtype=np.float32)
        call("scxr_xr2d_view_table", self.plan, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        return out

    def project(self, im, out=None, pred=None):
        """X-ray projection ``H @ im`` (``_xray2d.py:138-146``)."""
        im = dev.input_array(im, self.input_shape, "im")
        out = dev.output_buffer(out, self.output_shape)
        call("scxr_xr2d_forward", self.plan, im.data_ptr(), out.data_ptr(), dev.ptr(_pred.resolve(pred)),
             dev.stream())
        return out

    def back_project(self, y, out=None, pred=None):
        """Back projection ``H.T @ y`` (``_xray2d.py:148-156``)."""
        y = dev.input_array(y, self.output_shape, "y")
        out = dev.output_buffer(out, self.input_shape)
        call("scxr_xr2d_adjoint", self.plan, y.data_ptr(), out.data_ptr(), dev.ptr(_pred.resolve(pred)),
             dev.stream())
        return out

    def weights(self):
        """Debug: ``(inds, weights)`` of every (view, pixel), each (A, M, N)."""
        A = self.output_shape[0]
        inds = dev.empty((A, *self.nx), dtype=np.int32)
        w = dev.empty((A, *self.nx))
        call("scxr_xr2d_weights", self.plan, inds.data_ptr(), w.data_ptr(), dev.stream())
        return inds, w

    def fbp(self, y):
        """Filtered back projection (``_xray2d.py:158-197``)."""
        from ._fbp import fbp

        return fbp(self, y)
