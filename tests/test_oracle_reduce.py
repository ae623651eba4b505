"""Pins the plain-reduction oracles (CPU).

* the C restatement of numbagg/funcs.py:23-228 against NumPy's nan-functions
  on the reference's random test arrays (numbagg/test/util.py:90-109) and
  its edge cases (numbagg/test/test_funcs.py:261-327): exact for counts,
  flags, extremes and indices, to NumPy's precision for sums, means and
  variances;
* the exact oracle (math.fsum) against the float64 restatement;
* planted mistakes (a population variance, the last index of a tied
  maximum) are caught.
"""

from __future__ import annotations

import numpy as np
import pytest
from numpy.testing import assert_allclose, assert_array_equal

from oracle import exact, restated


def _rows(seed, dtype, rows=6, n=500):
    rs = np.random.RandomState(seed)
    a = np.arange(rows * n, dtype=dtype)
    if np.issubdtype(dtype, np.floating):
        a[rs.rand(*a.shape) < 0.2] = np.inf
        a[rs.rand(*a.shape) < 0.2] = np.nan
        idx = rs.rand(*a.shape) < 0.2
        a[idx] *= -1
    rs.shuffle(a)
    return a.reshape(rows, n)


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32, np.int64])
def test_restatement_matches_numpy(dtype):
    x = _rows(1, dtype)
    with np.errstate(invalid="ignore"):
        assert_array_equal(restated.reduce("nancount", x), np.sum(~np.isnan(x.astype(np.float64)), axis=-1))
        assert_array_equal(restated.reduce("allnan", x), np.isnan(x.astype(np.float64)).all(axis=-1))
        assert_array_equal(restated.reduce("anynan", x), np.isnan(x.astype(np.float64)).any(axis=-1))
        mx = restated.reduce("nanmax", x)
        assert_array_equal(mx, np.nanmax(x, axis=-1))
        assert_array_equal(restated.reduce("nanmin", x), np.nanmin(x, axis=-1))
        if np.issubdtype(dtype, np.floating):
            xf = np.where(np.isinf(x), 0, x)  # np.nanargmax's tie rule matches without inf
            assert_array_equal(restated.reduce("nanargmax", xf), np.nanargmax(xf, axis=-1))
            assert_array_equal(restated.reduce("nanargmin", xf), np.nanargmin(xf, axis=-1))
            xs = np.where(np.isinf(x), 1.0, x)
            assert_allclose(restated.reduce("nansum", xs), np.nansum(xs, axis=-1), rtol=1e-6)
            assert_allclose(restated.reduce("nanmean", xs), np.nanmean(xs, axis=-1), rtol=1e-6)
            for ddof in (0, 1, 3):
                assert_allclose(restated.reduce("nanvar", xs, ddof=ddof), np.nanvar(xs, axis=-1, ddof=ddof), rtol=1e-5)
                assert_allclose(restated.reduce("nanstd", xs, ddof=ddof), np.nanstd(xs, axis=-1, ddof=ddof), rtol=1e-5)
        else:
            assert_array_equal(restated.reduce("nansum", x), np.sum(x, axis=-1))
            assert_array_equal(restated.reduce("nanargmax", x), np.argmax(x, axis=-1))


def test_edge_cases():
    # numbagg/test/test_funcs.py:261-327 and the ties/NaN rows of util.py:40-66
    assert restated.reduce("allnan", np.array([[np.nan, np.nan, np.nan]]))[0]
    assert not restated.reduce("allnan", np.array([[np.nan, 1.0, np.nan]]))[0]
    assert restated.reduce("anynan", np.array([[1.0, np.nan, 3.0]]))[0]
    assert not restated.reduce("anynan", np.array([[1.0]]))[0]
    assert restated.reduce("nanargmax", np.array([[0.0, 0.0, 0.0]]))[0] == 0
    assert restated.reduce("nanargmin", np.array([[np.nan, 2.0, 3.0]]))[0] == 1
    assert restated.reduce("nanargmax", np.array([[np.nan, np.nan]]))[0] == -1
    assert np.isnan(restated.reduce("nanmax", np.array([[np.nan, np.nan]]))[0])
    assert restated.reduce("nanmax", np.array([[np.inf, -np.inf]]))[0] == np.inf
    # max/min replace on >= / <=: the last of equal extremes (signed zeros)
    assert np.signbit(restated.reduce("nanmax", np.array([[0.0, -0.0]]))[0])
    assert not np.signbit(restated.reduce("nanmin", np.array([[-0.0, 0.0]]))[0])
    # int32 sums wrap
    assert restated.reduce("nansum", np.array([[2**31 - 1, 1]], dtype=np.int32))[0] == -(2**31)
    assert np.isnan(restated.reduce("nanvar", np.array([[1.0, np.nan]]))[0])  # count <= ddof


def test_exact_oracle_and_planted_mistakes():
    rng = np.random.default_rng(3)
    x = (rng.standard_normal((4, 3000)) + 1e4).astype(np.float64)
    x[rng.random(x.shape) < 0.1] = np.nan
    for f in ("nansum", "nanmean", "nanvar", "nanstd"):
        assert_allclose(exact.reduce(f, x), restated.reduce(f, x), rtol=1e-10)
    # a population variance (ddof 0 where 1 is asked) misses the bar
    pop = exact.reduce("nanvar", x, ddof=0)
    assert np.all(np.abs(pop / exact.reduce("nanvar", x) - 1) > 1e-4)
    # "last index wins" for a tied maximum gives another index
    t = np.array([[1.0, 5.0, 5.0, 2.0]])
    assert restated.reduce("nanargmax", t)[0] == 1
