"""The allocation-free group_nansum entry point (static per-device scratch)
against the workspace entry point, through the C ABI.

Two calls in a row, the second on a smaller input (the static scratch keeps
the first call's contents, which the second must not see). Integer-valued
float data sums exactly in float64, so both entry points must agree bit for
bit there; on N(0, 1) data both must meet the parity bar of group_nansum
(within 1 float32 ULP of the float64 oracle, plus 1e-14 of the sum of |x|).
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sum_bar(got, v, lab, L):
    from oracle import exact

    want = exact.group("group_nansum", v, lab, L)
    sabs = exact.group("group_nansum", np.abs(v), lab, L)
    ulp = np.spacing(np.abs(want.astype(np.float32))).astype(np.float64)
    assert np.all(np.abs(got.astype(np.float64) - want) <= ulp + 1e-14 * sabs)


def _call_both(v, lab, L, strategy):
    import torch

    from numbagg_amd import _device as dev
    from numbagg_amd import _lib

    tv = torch.from_numpy(v).cuda()
    tl = torch.from_numpy(lab).cuda()
    rows, n = v.shape
    out_ws = torch.full((rows, L), 7.0, dtype=tv.dtype, device=tv.device)
    out_na = torch.full((rows, L), 7.0, dtype=tv.dtype, device=tv.device)
    nbytes = _lib.query("nbg_group_workspace_bytes", dev.nbg_code(tv), rows, n, L, strategy)
    ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=tv.device)
    stream = dev.stream_handle()
    _lib.call(
        "nbg_group_nansum", ctypes.c_int(dev.nbg_code(tv)), dev.ptr(tv), ctypes.c_int(dev.nbg_code(tl)),
        dev.ptr(tl), dev.ptr(out_ws), rows, n, n, L, ctypes.c_int(strategy), dev.ptr(ws), nbytes, stream,
    )
    _lib.call(
        "nbg_group_nansum_noalloc", ctypes.c_int(dev.nbg_code(tv)), dev.ptr(tv), ctypes.c_int(dev.nbg_code(tl)),
        dev.ptr(tl), dev.ptr(out_na), rows, n, n, L, ctypes.c_int(strategy), stream,
    )
    torch.cuda.synchronize()
    return out_ws.cpu().numpy(), out_na.cpu().numpy()


@pytest.mark.parametrize("strategy", [0, 1, 2])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_noalloc_matches_workspace_entry(gpu_lib, strategy, dtype):
    rng = np.random.default_rng(11 + strategy)
    for rows, n, L, integer in ((6, 300_001, 900, True), (3, 70_003, 250, True),
                                (6, 300_001, 900, False), (3, 70_003, 250, False)):
        if integer:
            v = rng.integers(-50, 50, size=(rows, n)).astype(dtype)
        else:
            v = rng.standard_normal((rows, n)).astype(dtype)
        v[rng.random(v.shape) < 0.05] = np.nan
        lab = rng.integers(-1, L, size=n).astype(np.int32)
        got_ws, got_na = _call_both(v, lab, L, strategy)
        if integer:
            np.testing.assert_array_equal(got_na.view(np.uint8), got_ws.view(np.uint8))
        if dtype == np.float32:
            _sum_bar(got_na, v, lab, L)
            _sum_bar(got_ws, v, lab, L)
        else:
            from oracle import exact

            sabs = exact.group("group_nansum", np.abs(v), lab, L)
            want = exact.group("group_nansum", v, lab, L)
            assert np.all(np.abs(got_na - want) <= 1e-13 * sabs + 1e-300)


def test_noalloc_rejects_oversized_inputs(gpu_lib):
    from numbagg_amd import _lib

    lib = _lib.load()
    # 1024 rows x 1024 labels need 32 MiB of scratch: rejected before any launch
    rc = lib.nbg_group_nansum_noalloc(
        _lib.NBG_F32, ctypes.c_void_p(256), _lib.NBG_I32, ctypes.c_void_p(256), ctypes.c_void_p(256),
        1024, 10, 10, 1024, 0, None,
    )
    assert rc == 1
    assert b"static scratch" in lib.nbg_last_error()
