"""GPU parity of the exponentially weighted kernels.

Bars against the long-double exact oracle (NaN masks identical):
  float64 loop: count/mean/sum <= 1e-12 relative (+ 1e-12 of max|x|, for means
  that cancel to near zero); var/std/cov/corr <= 1e-9
  relative + 1e-9 absolute of the data scale (the reference formula
  E[x^2]-E[x]^2 loses digits for offset data);
  float32 loop: <= 2 float32 ULP (var/std/cov/corr: or 1e-6 absolute).
The reference's pandas vectors and known answers are reproduced.
"""

from __future__ import annotations

import numpy as np
import pytest
from numpy.testing import assert_allclose

from ._data import walk_array

pytestmark = pytest.mark.gpu

FUNCS1 = ["move_exp_nancount", "move_exp_nanmean", "move_exp_nansum", "move_exp_nanvar", "move_exp_nanstd"]
FUNCS2 = ["move_exp_nancov", "move_exp_nancorr"]
MOMENTS = {"move_exp_nanvar", "move_exp_nanstd", "move_exp_nancov", "move_exp_nancorr"}


def _args(func, x, y):
    return (x,) if func in FUNCS1 else (x, y)


@pytest.mark.parametrize("func", FUNCS1 + FUNCS2)
@pytest.mark.parametrize("shape,alpha", [((3, 500), 0.5), ((4, 70_000), 0.01), ((2, 9000), 0.001), ((1, 300_000), 0.05), ((33, 4097), 0.3)])
def test_exp_float64_loop_against_exact(gpu_lib, func, shape, alpha):
    """float32 data with a Python-float alpha runs on the float64 loop."""
    import numbagg_amd as nb
    from oracle import exact

    x = walk_array(*shape, seed=shape[1], offset=100.0)
    y = walk_array(*shape, seed=shape[1] + 7, offset=-50.0)
    got = getattr(nb, func)(*_args(func, x, y), alpha=alpha)
    assert got.dtype == np.float64
    want = exact.move_exp(func, x, alpha, y=y if func in FUNCS2 else None, loop_dtype=np.float64)
    assert np.array_equal(np.isnan(got), np.isnan(want)), func
    ok = ~np.isnan(want)
    scale = float(np.nanmax(np.abs(x)))
    if func in MOMENTS:
        assert_allclose(got[ok], want[ok], rtol=1e-9, atol=1e-9 * scale**2)
    else:
        # relative to the data's scale: a mean near zero is a cancellation
        assert_allclose(got[ok], want[ok], rtol=1e-12, atol=1e-12 * scale)


@pytest.mark.parametrize("two_pass", ["1", "0"])
@pytest.mark.parametrize("func", FUNCS1 + FUNCS2)
def test_exp_two_pass_against_exact(gpu_lib, monkeypatch, two_pass, func):
    """The two-pass path (tile aggregates, per-row scan, apply) and the
    single pass with look-back meet the same bars: rows of more than 256
    tiles (several chunks of the row scan), partial last tiles, NaN stretches
    longer than a tile, scalar and per-position alpha."""
    import numbagg_amd as nb
    from oracle import exact

    monkeypatch.setenv("NBG_EXP_TWO_PASS", two_pass)
    x = walk_array(2, 1_100_003, seed=11, offset=30.0)
    y = walk_array(2, 1_100_003, seed=12, offset=-20.0)
    x[0, 5000:20_000] = np.nan
    y[1, 700_000:720_000] = np.nan
    scale = float(np.nanmax(np.abs(x)))
    for alpha in (0.01, np.linspace(0.005, 0.05, x.shape[1])):
        got = getattr(nb, func)(*_args(func, x, y), alpha=alpha, min_weight=0.2)
        want = exact.move_exp(func, x, alpha, 0.2, y=y if func in FUNCS2 else None, loop_dtype=np.float64)
        assert np.array_equal(np.isnan(got), np.isnan(want)), func
        ok = ~np.isnan(want)
        if func in MOMENTS:
            assert_allclose(got[ok], want[ok], rtol=1e-9, atol=1e-9 * scale**2)
        else:
            assert_allclose(got[ok], want[ok], rtol=1e-12, atol=1e-12 * scale)
    # float32 loop (float32 alpha array), several rows of partial tiles
    x32 = walk_array(9, 300_001, seed=13, offset=0.0)
    y32 = walk_array(9, 300_001, seed=14, offset=0.0)
    al = np.full(300_001, 0.03, dtype=np.float32)
    got = getattr(nb, func)(*_args(func, x32, y32), alpha=al)
    want = exact.move_exp(func, x32, al, y=y32 if func in FUNCS2 else None, loop_dtype=np.float32)
    assert np.array_equal(np.isnan(got), np.isnan(want)), func
    ok = ~np.isnan(want)
    e = want[ok].astype(np.float32)
    ulps = np.abs(got[ok].astype(np.float64) - want[ok]) / np.spacing(np.abs(e)).astype(np.float64)
    if func in MOMENTS:
        assert np.all((ulps <= 2) | (np.abs(got[ok] - want[ok]) <= 1e-6 * np.maximum(1, np.abs(want[ok])))), ulps.max()
    else:
        assert ulps.max() <= 2, ulps.max()


@pytest.mark.parametrize("func", FUNCS1 + FUNCS2)
def test_exp_float32_loop_against_exact(gpu_lib, func):
    import numbagg_amd as nb
    from oracle import exact

    x = walk_array(6, 20_000, seed=3, offset=0.0)
    y = walk_array(6, 20_000, seed=4, offset=0.0)
    alpha = np.full(20_000, 0.02, dtype=np.float32)
    got = getattr(nb, func)(*_args(func, x, y), alpha=alpha, min_weight=0.5)
    assert got.dtype == np.float32
    want = exact.move_exp(func, x, alpha, 0.5, y=y if func in FUNCS2 else None, loop_dtype=np.float32)
    assert np.array_equal(np.isnan(got), np.isnan(want)), func
    ok = ~np.isnan(want)
    e = want[ok].astype(np.float32)
    ulps = np.abs(got[ok].astype(np.float64) - want[ok]) / np.spacing(np.abs(e)).astype(np.float64)
    if func in MOMENTS:
        assert np.all((ulps <= 2) | (np.abs(got[ok] - want[ok]) <= 1e-6 * np.maximum(1, np.abs(want[ok])))), ulps.max()
    else:
        assert ulps.max() <= 2, ulps.max()


@pytest.mark.parametrize("func", FUNCS1 + FUNCS2)
def test_exp_against_restated_reference(gpu_lib, func):
    import numbagg_amd as nb
    from oracle import restated

    rng = np.random.default_rng(11)
    x = rng.standard_normal((5, 12_345))
    x[rng.random(x.shape) < 0.05] = np.nan
    y = rng.standard_normal((5, 12_345))
    got = getattr(nb, func)(*_args(func, x, y), alpha=0.07, min_weight=0.3)
    ref = restated.move_exp(func, x, 0.07, 0.3, y=y if func in FUNCS2 else None)
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    assert_allclose(got, ref, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("func", FUNCS1 + FUNCS2)
@pytest.mark.parametrize("alpha", [0.5, 0.1])
def test_exp_reference_pandas_vectors(gpu_lib, golden, func, alpha):
    import numbagg_amd as nb

    a, a2 = golden["array_3x500"], golden["array2_3x500"]
    got = getattr(nb, func)(*_args(func, a, a2), alpha=alpha)
    assert_allclose(got, golden[f"{func}_a{alpha}"])
    # 1-D alpha along the axis, and the transposed layout (test_moving_exp.py:348-368)
    alphas = np.full(a.shape[-1], alpha)
    assert_allclose(getattr(nb, func)(*_args(func, a, a2), alpha=alphas), got)
    t_args = (a.T,) if func in FUNCS1 else (a.T, a2.T)
    assert_allclose(getattr(nb, func)(*t_args, alpha=alphas, axis=0).T, got)
    assert_allclose(getattr(nb, func)(*t_args, alpha=alpha, axis=0).T, got)


def test_exp_known_answers(gpu_lib):
    import numbagg_amd as nb

    assert_allclose(nb.move_exp_nanmean(np.array([10, 0, np.nan, 10]), alpha=0.5), [10.0, 3.3333333, 3.3333333, 8.1818182])
    assert_allclose(nb.move_exp_nanmean(np.array([10, 0, np.nan, 10]), alpha=0.25), [10.0, 4.2857143, 4.2857143, 7.1653543])
    assert_allclose(nb.move_exp_nancount(np.array([1, 0, np.nan, np.nan, 1, 0]), alpha=0.5), [1.0, 1.5, 0.75, 0.375, 1.1875, 1.59375])
    assert_allclose(nb.move_exp_nansum(np.array([10, 0, np.nan, 10]), alpha=0.25), [10.0, 7.5, 5.625, 14.21875])
    assert_allclose(
        nb.move_exp_nancorr(np.array([10, 0, 5, 10]), np.array([10, 0, 10, 5]), alpha=0.5),
        [np.nan, 1.0, 0.8485281, 0.2274294],
    )
    # count vs sum on ones, count before observations (test_moving_exp.py:123-157)
    for alpha in (0.1, 0.5, 0.9):
        arr = np.array([1, 1, np.nan, np.nan, 1, 1])
        assert_allclose(nb.move_exp_nancount(arr, alpha=alpha), nb.move_exp_nansum(arr, alpha=alpha))
        r = nb.move_exp_nancount(np.array([np.nan, np.nan, 1.0, np.nan, 1.0]), alpha=alpha)
        assert r[0] == 0.0 and r[1] == 0.0 and r[2] == 1.0
        assert np.all(np.isnan(nb.move_exp_nansum(np.full(4, np.nan), alpha=alpha)))


def test_exp_min_weight_and_nan_masks(gpu_lib):
    import numbagg_amd as nb

    from .test_oracle_moving_exp import MIN_WEIGHT_CASES, NAN_MASK_CASES

    array = np.ones(25)
    array[:5] = np.nan
    for func in ["move_exp_nancount", "move_exp_nanmean", "move_exp_nanstd", "move_exp_nansum", "move_exp_nanvar"]:
        for mw, alpha, expected in MIN_WEIGHT_CASES:
            out = getattr(nb, func)(array, min_weight=mw, alpha=alpha)
            assert np.sum(~np.isnan(out)[6:]) == expected, (func, mw, alpha)
    for func in ["move_exp_nanvar", "move_exp_nanstd", "move_exp_nancov", "move_exp_nancorr"]:
        for alpha in (0.1, 0.5, 0.9):
            for x, want, want_corr in NAN_MASK_CASES:
                x = np.array(x)
                exp_mask = want_corr if (func == "move_exp_nancorr" and want_corr is not None) else want
                args = (x,) if func in FUNCS1 else (x, x)
                assert list(np.isnan(getattr(nb, func)(*args, alpha=alpha))) == exp_mask, (func, x, alpha)


def test_exp_inf_and_big_values(gpu_lib):
    # numbagg/test/test_moving_exp.py:251-295
    import numbagg_amd as nb

    assert_allclose(nb.move_exp_nanmean(np.array([np.inf]), alpha=0.25), [np.inf])
    np.testing.assert_array_equal(nb.move_exp_nanmean(np.array([0, 0, np.inf], dtype=np.float16), alpha=1.0), [0, 0, np.inf])
    np.testing.assert_array_equal(nb.move_exp_nanmean(np.array([0, np.inf, np.inf], dtype=np.float16), alpha=1.0), [0, np.inf, np.nan])
    big = np.array([[0.00000000e000, 1.19846209e308], [1.19846209e308, 1.19846209e308]])
    np.testing.assert_array_equal(nb.move_exp_nanmean(big, alpha=1.0), big)
    np.testing.assert_array_equal(nb.move_exp_nanmean(np.array([[0, np.nan]]), alpha=1.0), [[0, np.nan]])
    np.testing.assert_array_equal(nb.move_exp_nanmean(np.array([[1, np.nan]]), alpha=1.0), [[1, np.nan]])


def test_exp_alpha_modes_and_dtypes(gpu_lib):
    import torch

    import numbagg_amd as nb
    from oracle import exact

    rng = np.random.default_rng(2)
    x = rng.standard_normal((4, 5, 3000)).astype(np.float32)
    x[rng.random(x.shape) < 0.1] = np.nan
    # N-D alpha laid out like the input
    al = rng.uniform(0.01, 0.5, size=x.shape)
    got = nb.move_exp_nanmean(x, alpha=al)
    want = exact.move_exp("move_exp_nanmean", x.reshape(-1, 3000), al.reshape(-1, 3000), loop_dtype=np.float64)
    assert_allclose(got.reshape(-1, 3000), want, rtol=1e-12, atol=1e-12 * np.nanmax(np.abs(x)))
    # axis=1 with an N-D alpha
    got = nb.move_exp_nanvar(x, alpha=al, axis=1)
    xm, am = np.moveaxis(x, 1, -1), np.moveaxis(al, 1, -1)
    want = exact.move_exp("move_exp_nanvar", xm.reshape(-1, 5), am.reshape(-1, 5), loop_dtype=np.float64).reshape(xm.shape)
    assert_allclose(got, np.moveaxis(want, -1, 1), rtol=1e-9, atol=1e-12)
    # axis tuple and empty tuple (test_moving_exp.py:371-391)
    arr = np.array([1.0, 2.0, 3.0, 4.0])
    assert_allclose(nb.move_exp_nansum(arr, alpha=0.5, axis=(0,)), nb.move_exp_nansum(arr, alpha=0.5, axis=0))
    assert_allclose(nb.move_exp_nansum(arr, alpha=np.full(4, 0.5), axis=(0,)), nb.move_exp_nansum(arr, alpha=0.5))
    assert nb.move_exp_nansum(arr, alpha=0.5, axis=()) is arr
    # float32 alpha keeps the float32 loop; device tensors stay on the device
    t = torch.from_numpy(x[0]).cuda()
    r = nb.move_exp_nanmean(t, alpha=np.float32(0.1))
    assert r.is_cuda and r.dtype == torch.float32
    # big-endian input (test_moving_exp.py:312-320)
    a = rng.random((3, 500))
    assert_allclose(nb.move_exp_nanmean(a.astype(">f8"), alpha=0.25), nb.move_exp_nanmean(a, alpha=0.25))
