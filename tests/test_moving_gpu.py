"""GPU parity of the moving-window kernels.

Bars (against the long-double exact oracle; NaN masks must be identical):
  float32 sum/mean/var/std: <= 1 float32 ULP of the exact value;
  float32 cov/corr: <= 2 ULP, or <= 1e-6 of the window's scale;
  float64 mean (the (100, 10000) window-20 config): <= 1e-12 relative;
  float64 sum: <= 1e-12 of sum|x| in the window; var/std/cov/corr: <= 1e-11 relative.
Against the C restatement of the reference the NaN masks are identical and the
GPU error never exceeds the reference's own error (both against the exact oracle).
"""

from __future__ import annotations

import numpy as np
import pytest
from numpy.testing import assert_allclose

from ._data import walk_array

pytestmark = pytest.mark.gpu

FUNCS1 = ["move_mean", "move_sum", "move_std", "move_var"]
FUNCS2 = ["move_cov", "move_corr"]


def ulp_err(got32, exact64):
    """|got - exact| in float32 ULPs of the exact value."""
    e = exact64.astype(np.float32)
    ulp = np.spacing(np.abs(e)).astype(np.float64)
    return np.abs(got32.astype(np.float64) - exact64) / ulp


def check_f32(func, got, want, scale=None):
    assert np.array_equal(np.isnan(got), np.isnan(want)), f"{func}: NaN mask differs"
    ok = ~np.isnan(want)
    if not ok.any():
        return
    err = ulp_err(got[ok], want[ok])
    if func in FUNCS2:
        absok = np.abs(got[ok] - want[ok]) <= 1e-6 * (scale[ok] if scale is not None else 1.0)
        assert np.all((err <= 2) | absok), f"{func}: max {err.max():.2f} ULP"
    else:
        assert err.max() <= 1.0, f"{func}: max {err.max():.3f} ULP"


@pytest.mark.parametrize("func", FUNCS1 + FUNCS2)
@pytest.mark.parametrize(
    "shape,window,min_count",
    [
        ((3, 500), 10, None),
        ((3, 500), 50, 3),
        ((4, 9000), 1000, 500),
        ((2, 20_000), 20, 20),
        ((2, 5000), 1, 0),
        ((1, 30_000), 12_000, 100),  # window longer than the LDS halo limit
        ((5, 4096), 4096, 1),
        ((3, 4097), 4097, 0),
    ],
)
def test_move_f32_against_exact(gpu_lib, func, shape, window, min_count):
    import numbagg_amd as nb
    from oracle import exact

    x = walk_array(*shape, seed=window + shape[1])
    y = walk_array(*shape, seed=window + shape[1] + 1) if func in FUNCS2 else None
    args = (x,) if y is None else (x, y)
    got = getattr(nb, func)(*args, window=window, min_count=min_count)
    assert got.dtype == np.float32 and got.shape == x.shape
    want = exact.move(func, x, window, min_count, y=y)
    check_f32(func, got, want, scale=np.ones_like(want))


@pytest.mark.parametrize("func", FUNCS1 + FUNCS2)
def test_move_f32_against_restated_reference(gpu_lib, func):
    """Same NaN mask as the reference restatement, and no larger error than the
    reference's own (float32-rounded products, running float64 sums), both
    measured against the exact oracle."""
    import numbagg_amd as nb
    from oracle import exact, restated

    x = walk_array(4, 6000, seed=5, offset=0.0)
    y = walk_array(4, 6000, seed=6, offset=0.0) if func in FUNCS2 else None
    args = (x,) if y is None else (x, y)
    got = getattr(nb, func)(*args, window=300, min_count=150)
    ref = restated.move(func, x, 300, 150, y=y)
    ex = exact.move(func, x, 300, 150, y=y)
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    ok = ~np.isnan(ex)
    err_gpu = np.abs(got[ok].astype(np.float64) - ex[ok])
    err_ref = np.abs(ref[ok].astype(np.float64) - ex[ok])
    half_ulp = 0.5 * np.spacing(np.abs(ex[ok]).astype(np.float32)).astype(np.float64)
    assert err_gpu.max() <= max(err_ref.max(), half_ulp.max())
    assert np.mean(err_gpu) <= np.mean(err_ref) + np.mean(half_ulp)


@pytest.mark.parametrize("func", FUNCS1 + FUNCS2)
@pytest.mark.parametrize("window,min_count", [(20, 20), (7, 2), (300, 0)])
def test_move_f64_against_exact(gpu_lib, func, window, min_count):
    import numbagg_amd as nb
    from oracle import exact

    rng = np.random.default_rng(window)
    x = rng.standard_normal((20, 10_000))
    x[rng.random(x.shape) < 0.1] = np.nan
    y = rng.standard_normal((20, 10_000)) if func in FUNCS2 else None
    args = (x,) if y is None else (x, y)
    got = getattr(nb, func)(*args, window=window, min_count=min_count)
    want = exact.move(func, x, window, min_count, y=y)
    assert np.array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    if func == "move_sum":
        scale = exact.move("move_sum", np.abs(x), window, 0)
        assert np.all(np.abs(got[ok] - want[ok]) <= 1e-12 * scale[ok] + 1e-300)
    elif func == "move_mean":
        assert_allclose(got[ok], want[ok], rtol=1e-12)
    else:
        assert_allclose(got[ok], want[ok], rtol=1e-11, atol=1e-13)


def test_move_mean_f64_config1(gpu_lib):
    """(100, 10000) float64, window=20, min_count=20, 10% NaN: within 1e-12
    relative of the exact window mean."""
    import numbagg_amd as nb
    from oracle import exact

    rng = np.random.default_rng(20240601)
    x = rng.standard_normal((100, 10_000))
    x[rng.random(x.shape) < 0.10] = np.nan
    got = nb.move_mean(x, window=20, min_count=20)
    want = exact.move("move_mean", x, 20, 20)
    assert np.array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    assert np.max(np.abs(got[ok] - want[ok]) / np.abs(want[ok])) <= 1e-12


@pytest.mark.parametrize("func", FUNCS1 + FUNCS2)
@pytest.mark.parametrize("window", [10, 50])
@pytest.mark.parametrize("min_count", [None, 0, 1, 3, "window"])
def test_move_reference_pandas_vectors(gpu_lib, golden, func, window, min_count):
    import numbagg_amd as nb

    a, a2 = golden["array_3x500"], golden["array2_3x500"]
    mc = window if min_count == "window" else min_count
    want = golden[f"move_{func[5:]}_w{window}_mc{'None' if min_count is None else mc}"]
    args = (a, a2) if func in FUNCS2 else (a,)
    assert_allclose(getattr(nb, func)(*args, window=window, min_count=mc), want)


def test_move_known_answers(gpu_lib):
    import numbagg_amd as nb

    rs = np.random.RandomState(0)
    arr = (rs.random(1000) * 1e13).astype(np.float32)
    assert_allclose(nb.move_mean(arr, window=1), arr)
    arr = np.tile((rs.rand(10) * 1e13).astype(np.float32), 100)
    assert nb.move_sum(arr, window=10)[-1] == np.sum(arr[:10])


def test_move_axes_and_layouts(gpu_lib):
    import torch

    import numbagg_amd as nb
    from oracle import exact

    rng = np.random.default_rng(9)
    x = rng.standard_normal((5, 6, 700)).astype(np.float32)
    x[rng.random(x.shape) < 0.2] = np.nan
    for axis in (0, 1, 2, -1, -2):
        got = nb.move_mean(x, window=3, min_count=1, axis=axis)
        xm = np.moveaxis(x, axis, -1)
        want = exact.move("move_mean", xm.reshape(-1, xm.shape[-1]), 3, 1).reshape(xm.shape)
        want = np.moveaxis(want, -1, axis)
        assert np.array_equal(np.isnan(got), np.isnan(want))
        assert_allclose(got, want, rtol=1e-6, atol=1e-7)
    # device tensors stay on the device; an unaligned row start takes the scalar path
    t = torch.from_numpy(x).cuda()[:, :, 3:]
    got = nb.move_sum(t, window=5)
    assert got.is_cuda
    want = exact.move("move_sum", x[:, :, 3:].reshape(-1, 697), 5).reshape(5, 6, 697)
    assert_allclose(got.cpu().numpy(), want, rtol=1e-6)
    # broadcasting of the two inputs of move_corr
    a = x[0]
    b = rng.standard_normal(700).astype(np.float32)
    got = nb.move_corr(a, b, window=40)
    want = exact.move("move_corr", a, 40, None, y=np.broadcast_to(b, a.shape))
    assert_allclose(got, want, rtol=1e-5, atol=1e-6)
    # integer input runs on the float64 loop (numpy's safe-cast loop choice)
    i = rng.integers(-50, 50, size=(3, 100)).astype(np.int32)
    got = nb.move_mean(i, window=4)
    assert got.dtype == np.float64
    assert_allclose(got, exact.move("move_mean", i.astype(np.float64), 4), rtol=1e-14)
    assert nb.move_mean(i.astype(np.int16), window=4).dtype == np.float32



@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("func", FUNCS1 + FUNCS2)
@pytest.mark.parametrize("window", [1, 37, 1000, 2047, 2048, 4081, 4096, 6000])
def test_move_span_carried_halo_is_bitwise_neutral(gpu_lib, monkeypatch, func, dtype, window):
    """A workgroup walking a span of tiles carries each halo forward in LDS
    (and optionally prefetches the next tile into registers); outputs must be
    bitwise identical to one tile per workgroup (every tile evaluates its
    window from the same staged inputs)."""
    import numbagg_amd as nb

    x = walk_array(3, 5 * 4096 + 123, seed=window, dtype=dtype)
    y = walk_array(3, 5 * 4096 + 123, seed=window + 1, dtype=dtype) if func in FUNCS2 else None
    args = (x,) if y is None else (x, y)
    outs = []
    for span, prefetch in (("1", "0"), ("2", "0"), ("4", "1"), ("16", "0"), ("16", "1")):
        monkeypatch.setenv("NBG_MOVE_SPAN", span)
        monkeypatch.setenv("NBG_MOVE_PREFETCH", prefetch)
        outs.append(getattr(nb, func)(*args, window=window, min_count=1))
    for o in outs[1:]:
        np.testing.assert_array_equal(o.view(np.uint8), outs[0].view(np.uint8))


@pytest.mark.parametrize("func", ["move_sum", "move_mean"])
@pytest.mark.parametrize("shape,window,min_count", [((3, 500), 10, None), ((4, 9000), 1000, 500), ((2, 20_000), 20, 20),
                                                    ((1, 30_000), 12_000, 100)])
def test_move_f32_twosum_accumulator(gpu_lib, monkeypatch, func, shape, window, min_count):
    """The FP32 TwoSum accumulator (NBG_F32_ACCUMULATOR=twosum) meets the same
    bar as the float64 default: <= 1 float32 ULP of the exact value."""
    import numbagg_amd as nb
    from oracle import exact

    monkeypatch.setenv("NBG_F32_ACCUMULATOR", "twosum")
    x = walk_array(*shape, seed=window + shape[1] + 5)
    got = getattr(nb, func)(x, window=window, min_count=min_count)
    check_f32(func, got, exact.move(func, x, window, min_count))


def test_out_argument(gpu_lib):
    """out= is filled and returned (numbagg/decorators.py:341, :414, :487)."""
    import numbagg_amd as nb

    rng = np.random.default_rng(0)
    x = rng.standard_normal((3, 200))
    want = nb.move_mean(x, window=5)
    dest = np.empty_like(x)
    assert nb.move_mean(x, window=5, out=dest) is dest
    np.testing.assert_array_equal(dest, want)
    d32 = np.empty(x.shape, dtype=np.float32)
    assert nb.move_exp_nanmean(x, alpha=0.3, out=d32) is d32
    np.testing.assert_array_equal(d32, nb.move_exp_nanmean(x, alpha=0.3).astype(np.float32))
    xt = np.ascontiguousarray(x.T)
    dm = np.empty((200, 3, 3))
    assert nb.move_covmatrix(xt, window=10, out=dm) is dm
    np.testing.assert_array_equal(dm, nb.move_covmatrix(xt, window=10))
    df = np.empty_like(x)
    xn = x.copy()
    xn[:, ::3] = np.nan
    assert nb.ffill(xn, out=df) is df
    np.testing.assert_array_equal(df, nb.ffill(xn))
    with pytest.raises(ValueError):
        nb.move_mean(x, window=5, out=np.empty((3, 100)))


@pytest.mark.parametrize("func", ["move_sum", "move_mean", "move_var", "move_std", "move_corr"])
def test_element_aligned_rows_match_contiguous(gpu_lib, func):
    """Rows that start one element past a 16-byte boundary (a sliced device
    tensor) take the 16-byte vector paths at element alignment (streaming
    prefetch loads, LDS-staged stores): bitwise the results of an aligned
    contiguous copy."""
    import torch
    import numbagg_amd as nb

    x = torch.from_numpy(walk_array(3, 40_001, seed=21, offset=0.0)).cuda()
    y = torch.from_numpy(walk_array(3, 40_001, seed=22, offset=0.0)).cuda()
    xv, yv = x[:, 1:], y[:, 1:]
    assert xv.data_ptr() % 16 == 4
    args_v = (xv, yv) if func == "move_corr" else (xv,)
    args_c = tuple(a.contiguous() for a in args_v)
    got = getattr(nb, func)(*args_v, window=1000, min_count=500)
    want = getattr(nb, func)(*args_c, window=1000, min_count=500)
    assert torch.equal(got.view(torch.int32), want.view(torch.int32))
