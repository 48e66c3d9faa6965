"""ctypes binding of ``libscico_amd.so`` (the C ABI in ``include/scico_amd.h``).

The library is loaded lazily on first use.  There is no fallback: if the
library or a GPU is missing, every compute entry point raises
:class:`NativeUnavailable`.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int32, c_int64, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libscico_amd.so"

c_int32_p = POINTER(c_int32)
c_float_p = POINTER(c_float)
c_double_p = POINTER(c_double)

SCXR_OK = 0
_STATUS = {
    -1: "invalid argument",
    -2: "shape mismatch",
    -3: "unsupported geometry",
    -4: "HIP runtime error",
    -5: "device allocation failed",
    -6: "unsupported",
}


class NativeUnavailable(RuntimeError):
    """The HIP library (or a GPU) is not available."""


class ScxrError(RuntimeError):
    """A C ABI call returned a failure status."""

    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn}: {_STATUS.get(code, code)}: {msg}")
        self.code = code


# (name, restype, argtypes).  Every symbol declared in include/scico_amd.h
# appears here; tests check the two lists agree.
SIGNATURES = [
    ("scxr_version", c_int32, []),
    ("scxr_last_error", c_char_p, []),
    ("scxr_device_info", c_int32, [c_int32_p, c_int32_p, POINTER(c_int64), c_char_p, c_int32]),
    # 2D projector
    ("scxr_xr2d_plan_create", c_int32,
     [c_int32, c_int32, c_int32, c_int32, c_double_p, c_double_p, c_double, c_float_p,
      POINTER(c_void_p)]),
    ("scxr_xr2d_plan_destroy", c_int32, [c_void_p]),
    ("scxr_xr2d_view_table", c_int32, [c_void_p, c_float_p]),
    ("scxr_xr2d_forward", c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("scxr_xr2d_adjoint", c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("scxr_xr2d_weights", c_int32, [c_void_p, c_void_p, c_void_p, c_void_p]),
    ("scxr_xr2d_forward_t", c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("scxr_xr2d_info", c_int32, [c_void_p, c_int32_p]),
    ("scxr_xr2d_set_forward_split", c_int32, [c_void_p, c_int32, c_int32_p]),
    ("scxr_xr2d_forward_res", c_int32,
     [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("scxr_xr2d_adjoint_epi", c_int32,
     [c_void_p, c_void_p, c_void_p, c_int32, c_float, c_int32, c_void_p, c_void_p, c_void_p,
      c_void_p, c_void_p, c_void_p]),
    # ragged batches of 2D slices
    ("scxr_xr2d_batch_create", c_int32,
     [c_int32, c_int32_p, c_float_p, c_double_p, c_int32, POINTER(c_void_p)]),
    ("scxr_xr2d_batch_destroy", c_int32, [c_void_p]),
    ("scxr_xr2d_batch_info", c_int32, [c_void_p, POINTER(c_int64)]),
    ("scxr_xr2d_batch_offsets", c_int32, [c_void_p, POINTER(c_int64), POINTER(c_int64)]),
    ("scxr_xr2d_batch_configure", c_int32, [c_void_p, c_int32, c_int32, c_int32]),
    ("scxr_xr2d_batch_set_adjoint_groups", c_int32, [c_void_p, c_int32]),
    ("scxr_xr2d_batch_forward", c_int32,
     [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_void_p]),
    ("scxr_xr2d_batch_adjoint", c_int32,
     [c_void_p, c_void_p, c_void_p, c_int32, c_float, c_int32, c_void_p, c_void_p, c_void_p,
      c_void_p, c_void_p, c_int32, c_void_p]),
    ("scxr_batch_cg_start", c_int32,
     [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_double,
      c_double, c_void_p, c_void_p]),
    ("scxr_batch_cg_iter", c_int32,
     [c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
      c_void_p, c_void_p, c_void_p]),
    ("scxr_batch_admm_zu", c_int32,
     [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_int32, c_int32, c_void_p,
      c_void_p, c_void_p]),
    ("scxr_batch_admm_rhs", c_int32,
     [c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_int32, c_void_p, c_void_p]),
    ("scxr_batch_admm_dual", c_int32,
     [c_void_p, c_void_p, c_void_p, c_float, c_int32, c_void_p, c_void_p]),
    ("scxr_batch_admm_itstat", c_int32,
     [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_double, c_double,
      c_void_p, c_void_p, c_int32, c_void_p]),
    # 3D projector
    ("scxr_xr3d_plan_create", c_int32,
     [c_int32, c_int32, c_int32, c_int32, c_float_p, c_int32, c_int32, c_int32,
      POINTER(c_void_p)]),
    ("scxr_xr3d_plan_destroy", c_int32, [c_void_p]),
    ("scxr_xr3d_info", c_int32, [c_void_p, c_int32_p]),
    ("scxr_xr3d_forward", c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("scxr_xr3d_adjoint", c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("scxr_xr3d_weights", c_int32, [c_void_p, c_int32, c_void_p, c_void_p, c_void_p]),
    # finite differences, proxes
    ("scxr_fd_forward", c_int32,
     [c_int32, c_int32_p, c_void_p, c_void_p, c_int32, c_void_p, c_void_p]),
    ("scxr_fd_adjoint", c_int32,
     [c_int32, c_int32_p, c_void_p, c_void_p, c_int32, c_float, c_void_p, c_void_p]),
    ("scxr_fd_gram", c_int32,
     [c_int32, c_int32_p, c_float, c_void_p, c_void_p, c_int32, c_void_p, c_void_p]),
    ("scxr_l21_prox", c_int32, [c_int32, c_int64, c_void_p, c_float, c_void_p, c_void_p, c_void_p]),
    ("scxr_l1_prox", c_int32, [c_int64, c_void_p, c_float, c_void_p, c_void_p, c_void_p]),
    # CG
    ("scxr_cg_start", c_int32,
     [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_double,
      c_double, c_void_p, c_void_p]),
    ("scxr_cg_iter", c_int32,
     [c_int64, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
      c_void_p, c_int32, c_void_p, c_void_p]),
    ("scxr_cg_iter_t", c_int32,
     [c_int32, c_int32, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
      c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_void_p]),
    ("scxr_elem_blocks", c_int32, [c_int64]),
    # ADMM
    ("scxr_admm_rhs2d", c_int32,
     [c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_float, c_int32, c_void_p, c_void_p]),
    ("scxr_admm_zu2d", c_int32,
     [c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_int32, c_int32,
      c_void_p, c_void_p, c_void_p]),
    ("scxr_admm_dual2d", c_int32,
     [c_int32, c_int32, c_void_p, c_void_p, c_float, c_int32, c_void_p, c_void_p]),
    ("scxr_admm_rhs_nd", c_int32,
     [c_int32, c_int32_p, c_void_p, c_void_p, c_void_p, c_float, c_int32, c_void_p, c_void_p]),
    ("scxr_admm_lhs_nd", c_int32,
     [c_int32, c_int32_p, c_void_p, c_void_p, c_float, c_int32, c_void_p, c_void_p, c_void_p,
      c_void_p]),
    ("scxr_admm_zu_nd", c_int32,
     [c_int32, c_int32_p, c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_int32, c_int32,
      c_void_p, c_void_p, c_void_p]),
    ("scxr_admm_dual_nd", c_int32,
     [c_int32, c_int32_p, c_void_p, c_void_p, c_float, c_int32, c_void_p, c_void_p]),
    ("scxr_admm_itstat", c_int32,
     [c_void_p, c_int32, c_void_p, c_int32, c_void_p, c_int32, c_void_p, c_int32, c_void_p,
      c_double, c_double, c_void_p, c_void_p, c_int32, c_void_p]),
    # TV
    ("scxr_tv_prox", c_int32,
     [c_int32, c_int32_p, c_void_p, c_void_p, c_float, c_int32, c_int32, c_void_p, c_void_p]),
    ("scxr_tv_prox_w", c_int32,
     [c_int32, c_int32_p, c_void_p, c_void_p, c_void_p, c_float, c_int32, c_int32, c_void_p,
      c_void_p]),
    ("scxr_tv_prox_vjp_ws", c_int64, [c_int32, c_int32_p, c_int32]),
    ("scxr_tv_prox_vjp", c_int32,
     [c_int32, c_int32_p, c_void_p, c_void_p, c_void_p, c_float, c_int32, c_int32, c_void_p,
      c_void_p, c_void_p, c_void_p]),
    ("scxr_tv_norm", c_int32,
     [c_int32, c_int32_p, c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p]),
    # fused AcceleratedPGM deconvolution
    ("scxr_tv_prox_apgm_parts", c_int32, [c_int32, c_int32]),
    ("scxr_tv_prox_apgm", c_int32,
     [c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_float, c_int32, c_int32, c_float,
      c_void_p, c_void_p, c_void_p, c_void_p]),
    ("scxr_apgm_itstat", c_int32,
     [c_void_p, c_int32, c_double, c_void_p, c_int32, c_double, c_void_p, c_int32, c_void_p,
      c_void_p, c_int32, c_void_p]),
    ("scxr_cconv_apgm_grad", c_int32,
     [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_float,
      c_float, c_void_p, c_void_p, c_void_p]),
    ("scxr_cconv_apgm_grad_itstat", c_int32,
     [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_float,
      c_float, c_void_p, c_double, c_void_p, c_int32, c_double, c_void_p, c_int32, c_void_p,
      c_void_p, c_int32, c_void_p, c_void_p]),
    # CircularConvolve / FFT
    ("scxr_cconv_plan_create", c_int32, [c_int32, c_int32_p, c_int64, POINTER(c_void_p)]),
    ("scxr_cconv_plan_destroy", c_int32, [c_void_p]),
    ("scxr_cconv_set_spectrum", c_int32, [c_void_p, c_void_p, c_int32, c_void_p, c_void_p]),
    ("scxr_cconv_spectrum", c_int32, [c_void_p, c_void_p, c_void_p]),
    ("scxr_cconv_apply", c_int32, [c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_void_p]),
    ("scxr_fftn", c_int32, [c_void_p, c_void_p, c_int32, c_void_p, c_int32, c_void_p, c_void_p]),
    ("scxr_ramp_filter", c_int32,
     [c_int32, c_int32, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("scxr_vec_mul", c_int32, [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("scxr_vec_ge", c_int32, [c_int64, c_void_p, c_float, c_float, c_void_p, c_void_p, c_void_p]),
    ("scxr_vec_div_dev", c_int32, [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    # vector kernels
    ("scxr_vec_axpby", c_int32,
     [c_int64, c_float, c_void_p, c_float, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("scxr_vec_add", c_int32, [c_int64, c_void_p, c_float, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("scxr_vec_scale", c_int32, [c_int64, c_float, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("scxr_vec_fill", c_int32, [c_int64, c_float, c_void_p, c_void_p, c_void_p]),
    ("scxr_vec_copy", c_int32, [c_int64, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("scxr_vec_reduce", c_int32,
     [c_int32, c_int64, c_int32, c_void_p, c_void_p, c_double, c_int32, c_void_p, c_void_p,
      c_void_p, c_void_p]),
    ("scxr_vec_reduce_noalloc", c_int32,
     [c_int32, c_int64, c_int32, c_void_p, c_void_p, c_double, c_int32, c_void_p, c_void_p,
      c_void_p]),
]

_lib = None


def load(path: os.PathLike | str | None = None) -> ctypes.CDLL:
    """Load (once) and return the library; raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise NativeUnavailable(
            f"{p} not found: build it with `python -m scico_amd._build` "
            "(there is no CPU fallback)"
        )
    lib = ctypes.CDLL(str(p))
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(fn_name: str, rc: int) -> None:
    if rc != SCXR_OK:
        msg = _lib.scxr_last_error().decode() if _lib is not None else ""
        raise ScxrError(fn_name, rc, msg)


def call(fn_name: str, *args) -> None:
    """Call a status-returning entry point and raise on failure."""
    lib = load()
    check(fn_name, getattr(lib, fn_name)(*args))
