"""Label factorisation on the GPU: the step before the grouped reductions.

numbagg's grouped functions take labels already factorised into 0..K-1
(numbagg/decorators.py:570-573), and its benchmarks produce them with
``pd.factorize(labels, sort=True)`` (numbagg/test/conftest.py:175-189).
``factorize`` does that on the device with the hash-and-sort kernels of
``csrc/factorize.hip``: same codes and uniques as ``pandas.factorize``
(NaN -> -1; uniques ascending with ``sort=True``, else in order of first
appearance).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _device as dev
from . import _lib


def factorize(values, sort: bool = True):
    """(codes, uniques) of a 1-D array of int32/int64/float32/float64 labels
    (other integer, bool and float16 inputs are widened first). NumPy in,
    NumPy out; torch tensors on the GPU stay there."""
    as_tensor = dev.is_tensor(values)
    dt = dev.np_dtype(values)
    if dt.kind not in "iufb":
        raise TypeError(f"factorize: unsupported dtype {dt}")
    if dt in (np.dtype(np.int32), np.dtype(np.int64), np.dtype(np.float32), np.dtype(np.float64)):
        work = dt
    elif dt.kind == "f":
        work = np.dtype(np.float32) if dt.itemsize < 4 else np.dtype(np.float64)
    elif dt.kind == "u" and dt.itemsize >= 8:
        raise TypeError("factorize: uint64 labels are not supported")
    else:
        work = np.dtype(np.int64) if (dt.kind == "u" and dt.itemsize == 4) or dt.itemsize > 4 else np.dtype(np.int32)
    t = dev.to_device(values, work)
    if t.dim() != 1:
        raise ValueError(f"factorize: 1-D labels expected, got shape {tuple(t.shape)}")
    t = t.contiguous()
    n = int(t.numel())
    code = dev.nbg_code(t)
    ws_bytes = _lib.query("nbg_factorize_workspace_bytes", n)
    ws, ws_bytes = dev.workspace(ws_bytes, t.device)
    total = torch.zeros(1, dtype=torch.int64, device=t.device)
    stream = dev.stream_handle()
    _lib.call("nbg_factorize_count", ctypes.c_int(code), dev.ptr(t), n, ctypes.c_int(int(sort)), dev.ptr(ws),
              ws_bytes, dev.ptr(total), stream)
    k = int(total.item())  # the uniques' count sizes the output
    codes = torch.empty(n, dtype=torch.int64, device=t.device)
    uniques = torch.empty(k, dtype=t.dtype, device=t.device)
    _lib.call("nbg_factorize_finish", ctypes.c_int(code), dev.ptr(t), n, k, ctypes.c_int(int(sort)),
              dev.ptr(codes), dev.ptr(uniques), dev.ptr(ws), ws_bytes, stream)
    if work != dt and not as_tensor:
        uniques = uniques.to(dev.torch_dtype(dt)) if dt.kind != "b" else uniques.to(torch.bool)
    return dev.to_output(codes, as_tensor), dev.to_output(uniques, as_tensor)


__all__ = ["factorize"]
