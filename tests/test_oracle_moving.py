"""Pins the moving-window oracles (CPU).

* Both the C restatement and the long-double exact oracle reproduce pandas'
  rolling results on the reference's own fixture (the comparison the reference
  makes in numbagg/test/test_moving.py:29-53, rtol 1e-7).
* The exact oracle matches a brute-force per-window evaluation (the
  reference's `move_func`/`_mask` slow reference, test_moving.py:253-307) and
  the reference's known answers (test_moving.py:180-192).
* A planted off-by-one in min_count is caught.
"""

from __future__ import annotations

import warnings

import numpy as np
import pytest
from numpy.testing import assert_allclose

from oracle import exact, restated

FUNCS1 = ["move_mean", "move_sum", "move_std", "move_var"]
FUNCS2 = ["move_cov", "move_corr"]


def _tag(window, mc):
    return f"w{window}_mc{mc}"


@pytest.mark.parametrize("func", FUNCS1 + FUNCS2)
@pytest.mark.parametrize("window", [10, 50])
@pytest.mark.parametrize("min_count", [None, 0, 1, 3, "window"])
def test_oracles_match_pandas(golden, func, window, min_count):
    a, a2 = golden["array_3x500"], golden["array2_3x500"]
    mc = window if min_count == "window" else min_count
    want = golden[f"{func}_{_tag(window, 'None' if min_count is None else mc)}"]
    y = a2 if func in FUNCS2 else None
    assert_allclose(restated.move(func, a, window, mc, y=y), want)
    assert_allclose(exact.move(func, a, window, mc, y=y), want)


def _brute(func, a, window, min_count, b=None):
    """The reference's slow reference: nan-aware statistic over each window,
    masked where the window holds fewer than min_count values."""
    rows, n = a.shape
    out = np.full((rows, n), np.nan)
    mc = window if min_count is None else min_count
    if func == "move_mean" or func == "move_corr":
        mc = max(mc, 1)
    if func in ("move_std", "move_var", "move_cov"):
        mc = max(mc, 2)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        for r in range(rows):
            for i in range(n):
                lo = max(0, i - window + 1)
                x = a[r, lo : i + 1].astype(np.float64)
                ok = ~np.isnan(x)
                if b is not None:
                    y = b[r, lo : i + 1].astype(np.float64)
                    ok &= ~np.isnan(y)
                    y = y[ok]
                x = x[ok]
                if ok.sum() < mc:
                    continue
                if func == "move_sum":
                    out[r, i] = np.sum(x)
                elif func == "move_mean":
                    out[r, i] = np.mean(x)
                elif func == "move_var":
                    out[r, i] = np.var(x, ddof=1)
                elif func == "move_std":
                    out[r, i] = np.std(x, ddof=1)
                elif func == "move_cov":
                    out[r, i] = np.cov(x, y, ddof=1)[0, 1]
                else:
                    out[r, i] = np.corrcoef(x, y)[0, 1]
    return out


@pytest.mark.parametrize("func", FUNCS1 + FUNCS2)
@pytest.mark.parametrize("window,min_count", [(1, None), (3, 1), (7, None), (7, 0), (12, 5)])
def test_exact_oracle_matches_brute_force(func, window, min_count):
    rng = np.random.default_rng(2)
    a = rng.standard_normal((3, 40))
    a[rng.random(a.shape) < 0.25] = np.nan
    b = rng.standard_normal((3, 40)) if func in FUNCS2 else None
    if b is not None:
        b[rng.random(b.shape) < 0.1] = np.nan
    got = exact.move(func, a, window, min_count, y=b)
    want = _brute(func, a, window, min_count, b)
    assert_allclose(got, want, rtol=1e-12, atol=1e-14)


def test_known_answers():
    rs = np.random.RandomState(0)
    arr = (rs.random(1000) * 1e13).astype(np.float32)
    # window 1 returns the input (test_moving.py:180-183)
    assert_allclose(exact.move("move_mean", arr, 1), arr[None].astype(np.float64), rtol=1e-7)
    assert_allclose(restated.move("move_mean", arr, 1), arr[None])
    # a tiled array's last window sum equals the sum of one tile (test_moving.py:186-192)
    arr = np.tile((rs.rand(10) * 1e13).astype(np.float32), 100)
    assert restated.move("move_sum", arr, 10)[0, -1] == np.sum(arr[:10])
    assert np.float32(exact.move("move_sum", arr, 10)[0, -1]) == np.float32(np.sum(arr[:10].astype(np.float64)))


def test_planted_mistake_is_caught():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((2, 30))
    a[rng.random(a.shape) < 0.3] = np.nan
    wrong = exact.move("move_mean", a, 5, 4)  # min_count off by one
    want = _brute("move_mean", a, 5, 3)
    assert not np.array_equal(np.isnan(wrong), np.isnan(want))
