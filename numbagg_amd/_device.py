"""Device-array interop for the host layer.

PyTorch-ROCm tensors are the only device-array type: NumPy (or array-like)
inputs are copied to the current GPU, computed on, and copied back to NumPy;
torch tensors already on the GPU are used in place and results stay on the GPU.
Kernels are enqueued on torch's current stream, and scratch space comes from
torch's caching allocator, so steady-state calls never hit hipMalloc.
"""

from __future__ import annotations

import ctypes
import threading

import numpy as np
import torch

from . import _lib
from ._lib import NBG_F32, NBG_F64, NBG_I32, NBG_I64

_NBG_CODE = {
    torch.float32: NBG_F32,
    torch.float64: NBG_F64,
    torch.int32: NBG_I32,
    torch.int64: NBG_I64,
}

_NP_TO_TORCH = {
    np.dtype(np.float16): torch.float16,
    np.dtype(np.float32): torch.float32,
    np.dtype(np.float64): torch.float64,
    np.dtype(np.int8): torch.int8,
    np.dtype(np.int16): torch.int16,
    np.dtype(np.int32): torch.int32,
    np.dtype(np.int64): torch.int64,
    np.dtype(np.uint8): torch.uint8,
    np.dtype(np.uint16): torch.uint16,
    np.dtype(np.uint32): torch.uint32,
    np.dtype(np.uint64): torch.uint64,
    np.dtype(np.bool_): torch.bool,
}

_TORCH_TO_NP = {v: k for k, v in _NP_TO_TORCH.items()}


def require_gpu() -> None:
    if not torch.cuda.is_available():
        raise RuntimeError(
            "numbagg_amd runs on a ROCm GPU and none is visible "
            "(torch.cuda.is_available() is False); there is no CPU fallback"
        )


def is_tensor(x) -> bool:
    return isinstance(x, torch.Tensor)


def np_dtype(x) -> np.dtype:
    """NumPy dtype of an ndarray, array-like or torch tensor."""
    if isinstance(x, torch.Tensor):
        return _TORCH_TO_NP[x.dtype]
    return np.asarray(x).dtype


def torch_dtype(dt: np.dtype) -> torch.dtype:
    try:
        return _NP_TO_TORCH[np.dtype(dt).newbyteorder("=")]
    except KeyError:
        raise TypeError(f"numbagg_amd: unsupported dtype {dt}") from None


def nbg_code(t: torch.Tensor) -> int:
    return _NBG_CODE[t.dtype]


def to_device(a, dtype: np.dtype | None = None) -> torch.Tensor:
    """A tensor on the current GPU holding ``a`` (cast to ``dtype`` if given)."""
    require_gpu()
    if isinstance(a, torch.Tensor):
        t = a if a.is_cuda else a.cuda()
    else:
        arr = np.asarray(a)
        if arr.dtype.byteorder not in ("=", "|"):
            arr = arr.astype(arr.dtype.newbyteorder("="))
        if arr.dtype not in _NP_TO_TORCH:
            if dtype is None:
                raise TypeError(f"numbagg_amd: unsupported dtype {arr.dtype}")
            arr = arr.astype(dtype)
        t = torch.from_numpy(np.ascontiguousarray(arr)).cuda()
    if dtype is not None:
        td = torch_dtype(dtype)
        if t.dtype != td:
            t = t.to(td)
    return t


def rows_view(t: torch.Tensor) -> tuple[torch.Tensor, int, int, int]:
    """View ``t`` of shape (..., n) as (rows, n) with a unit-stride core axis.

    Returns (tensor2d, rows, n, ld) where ld is the row stride in elements. Copies
    only when the outer dimensions cannot be flattened into a single stride.
    """
    n = t.shape[-1] if t.dim() else 1
    rows = int(np.prod(t.shape[:-1], dtype=np.int64)) if t.dim() > 1 else 1
    if t.dim() == 1:
        t2 = t.view(1, n) if t.stride(-1) == 1 else t.contiguous().view(1, n)
    else:
        t2 = t if t.dim() == 2 else None
        if t2 is None or t2.stride(-1) != 1 or (rows > 1 and t2.stride(0) < n):
            try:
                t2 = t.view(rows, n)
            except RuntimeError:
                t2 = None
        if t2 is None or t2.stride(-1) != 1 or (rows > 1 and t2.stride(0) < n):
            t2 = t.contiguous().view(rows, n)
    ld = int(t2.stride(0)) if rows > 1 else n
    return t2, rows, n, ld


def stream_handle() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def workspace(nbytes: int, device: torch.device) -> tuple[torch.Tensor | None, int]:
    if nbytes <= 0:
        return None, 0
    return torch.empty(int(nbytes), dtype=torch.uint8, device=device), int(nbytes)


def ptr(t: torch.Tensor | None) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr() if t is not None else 0)


_spin_lock = threading.Lock()
_spin_seen: dict[int, int] = {}


def check_spin_timeouts() -> None:
    """Raise if a bounded device-side wait expired since the last check.

    The look-back kernels (fill, exponential scans) wait on other workgroups
    with a bounded spin; an expiry means a broken cross-workgroup protocol and
    invalid outputs. The counter is per device and only grows (until
    ``nbg_reset_spin_timeouts``), so any increase since the last observation
    on this device is reported, whichever call caused it.
    """
    dev_id = torch.cuda.current_device()
    count = _lib.spin_timeouts()
    with _spin_lock:
        seen = _spin_seen.get(dev_id, 0)
        _spin_seen[dev_id] = count
    if count > seen:
        raise RuntimeError(
            f"numbagg_amd: {count - seen} bounded device-side wait(s) expired; "
            "the results of the calls since the last check are invalid"
        )


def to_output(t: torch.Tensor, as_tensor: bool):
    """Return ``t`` as a torch tensor (device inputs) or a NumPy array.

    The NumPy path synchronises anyway, so it also checks the spin-timeout
    counter; callers that keep tensors on the GPU poll
    ``numbagg_amd.spin_timeouts()`` themselves.
    """
    if as_tensor:
        return t
    res = t.cpu().numpy()
    check_spin_timeouts()
    return res
