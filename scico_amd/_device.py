"""Device-array plumbing: torch tensors hold device memory and give the HIP
stream; every computation on them goes through the C ABI."""

from __future__ import annotations

import numpy as np
import torch

from ._lib import NativeUnavailable

DEVICE = "cuda"

_NP2TORCH = {
    np.dtype(np.float32): torch.float32,
    np.dtype(np.float64): torch.float64,
    np.dtype(np.int32): torch.int32,
    np.dtype(np.complex64): torch.complex64,
}
_TORCH2NP = {v: k for k, v in _NP2TORCH.items()}


def require_gpu() -> None:
    if not torch.cuda.is_available():
        raise NativeUnavailable("no HIP device visible: the scico_amd kernels need an MI355X")


def torch_dtype(dt) -> torch.dtype:
    if isinstance(dt, torch.dtype):
        return dt
    return _NP2TORCH[np.dtype(dt)]


def np_dtype(dt) -> np.dtype:
    if isinstance(dt, torch.dtype):
        return _TORCH2NP[dt]
    return np.dtype(dt)


def as_device(x, dtype=np.float32) -> torch.Tensor:
    """Contiguous device tensor of ``dtype`` (copies host arrays to the GPU).
    A block array becomes a block array of device tensors."""
    require_gpu()
    from .blockarray import BlockArray

    if isinstance(x, BlockArray):
        return BlockArray(as_device(b, dtype) for b in x)
    td = torch_dtype(dtype)
    if isinstance(x, torch.Tensor):
        if x.device.type != "cuda":
            x = x.to(DEVICE)
        if x.dtype != td:
            x = x.to(td)
        return x.contiguous()
    return torch.as_tensor(np.ascontiguousarray(x, dtype=np_dtype(dtype)), device=DEVICE)


def empty(shape, dtype=np.float32) -> torch.Tensor:
    require_gpu()
    return torch.empty(tuple(shape), dtype=torch_dtype(dtype), device=DEVICE)


def zeros(shape, dtype=np.float32) -> torch.Tensor:
    require_gpu()
    return torch.zeros(tuple(shape), dtype=torch_dtype(dtype), device=DEVICE)


# torch's raw current-stream query avoids building a Stream object per call
# (about 0.1 us instead of 3 us, which adds up over the dozens of launches
# of one solver iteration)
_RAW_STREAM = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def stream() -> int:
    """Raw hipStream_t of torch's current stream (captured by graphs)."""
    if _RAW_STREAM is not None:
        return _RAW_STREAM(torch.cuda.current_device())
    return torch.cuda.current_stream().cuda_stream


def input_array(x, shape, what: str = "input", dtype=np.float32) -> torch.Tensor:
    """``as_device(x, dtype)`` checked to have exactly ``shape`` before its
    pointer goes to the C ABI (which trusts sizes).  Raises ``ValueError``
    as scico does for non-conforming shapes (``_operator.py:219-223``)."""
    t = as_device(x, dtype)
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"Shapes do not conform: {what} array shape {tuple(t.shape)} != "
                         f"expected shape {tuple(shape)}")
    return t


def output_buffer(out, shape, what: str = "out", dtype=np.float32) -> torch.Tensor:
    """A new device array of ``shape``, or the caller's ``out`` once it is
    checked to be a contiguous device tensor of ``dtype`` holding exactly
    ``prod(shape)`` elements (a wrong buffer would be written out of
    bounds by the kernels, so it is refused with ``ValueError``)."""
    if out is None:
        return empty(shape, dtype)
    n = int(np.prod(shape, dtype=np.int64))
    if not isinstance(out, torch.Tensor) or out.device.type != "cuda":
        raise ValueError(f"{what} must be a device tensor")
    if out.dtype != torch_dtype(dtype):
        raise ValueError(f"{what} has dtype {out.dtype}, expected {torch_dtype(dtype)}")
    if not out.is_contiguous():
        raise ValueError(f"{what} must be contiguous")
    if out.numel() != n:
        raise ValueError(f"{what} holds {out.numel()} elements, expected {n} (shape {tuple(shape)})")
    return out


def check_buffer(t, n: int, what: str, dtype=np.float32, allow_none: bool = False,
                 at_least: bool = False):
    """Check a flat device operand of ``n`` elements (``at_least``: a
    workspace of at least ``n``) and return it; ``None`` passes when
    ``allow_none``."""
    if t is None:
        if allow_none:
            return None
        raise ValueError(f"{what} is required")
    if at_least and isinstance(t, torch.Tensor) and t.numel() >= n:
        return output_buffer(t, (t.numel(),), what, dtype)
    return output_buffer(t, (int(n),), what, dtype)


def ptr(t: torch.Tensor | None):
    return None if t is None else t.data_ptr()


def to_numpy(x) -> np.ndarray:
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    if isinstance(x, list):
        return [to_numpy(b) for b in x]
    return np.asarray(x)


def array_dtype(x) -> np.dtype:
    if isinstance(x, torch.Tensor):
        return np_dtype(x.dtype)
    if isinstance(x, list):
        return array_dtype(x[0])
    return np.asarray(x).dtype


def array_shape(x) -> tuple:
    return tuple(x.shape)
