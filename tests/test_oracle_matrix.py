"""Pins the matrix oracles (numbagg/funcs.py:338-532, numbagg/moving_matrix.py).

* the exact oracle (``oracle.exact.matrix``) and the C restatement of the
  reference loops (``oracle.restated.matrix``) against pandas' pairwise
  DataFrame statistics on the reference's fixture inputs (golden vectors
  written by tests/golden/make_golden.py: conftest.py:108-156, :204-211);
* the known answers of numbagg/test/test_matrix_functions.py on both oracles;
* closed forms: NumPy's corrcoef / cov without NaNs, a pair's entry equal to
  the two-series move_cov / move_corr / move_exp_nancov / nancorr oracles of
  the same columns, and brute-force pairwise-complete statistics.
"""

from __future__ import annotations

import pathlib

import numpy as np
import pytest
from numpy.testing import assert_allclose

import oracle

GOLD = np.load(pathlib.Path(__file__).resolve().parent / "golden" / "pandas_fixtures.npz")


def _same_nan(a, b):
    return np.array_equal(np.isnan(a), np.isnan(b))


@pytest.mark.parametrize("tag", ["3x500", "20x1000"])
@pytest.mark.parametrize("func", ["nancorrmatrix", "nancovmatrix"])
def test_static_against_pandas(func, tag):
    x = GOLD[f"matrix_array_{tag}"]
    want = GOLD[f"{func}_{tag}"]
    ex = oracle.exact.matrix(func, x)[0]
    assert _same_nan(ex, want)
    assert_allclose(ex, want, rtol=1e-12, atol=1e-15)
    rs = oracle.restated.matrix(func, x)[0]
    assert _same_nan(rs, want)
    assert_allclose(rs, want, rtol=1e-10, atol=1e-13)


@pytest.mark.parametrize("window,mc", [(20, None), (20, 3), (50, 1)])
@pytest.mark.parametrize("func", ["move_corrmatrix", "move_covmatrix"])
def test_moving_against_pandas(func, window, mc):
    x = np.ascontiguousarray(GOLD["matrix_array_3x500"].T)
    want = GOLD[f"{func}_w{window}_mc{mc}"]
    ex = oracle.exact.matrix(func, x, window=window, min_count=mc)[0]
    assert _same_nan(ex, want)
    assert_allclose(ex, want, rtol=1e-10, atol=1e-13)
    rs = oracle.restated.matrix(func, x, window=window, min_count=mc)[0]
    assert _same_nan(rs, want)
    assert_allclose(rs, want, rtol=1e-8, atol=1e-11)


@pytest.mark.parametrize("alpha", [0.5, 0.1])
@pytest.mark.parametrize("func", ["move_exp_nancorrmatrix", "move_exp_nancovmatrix"])
def test_exp_against_pandas(func, alpha):
    x = np.ascontiguousarray(GOLD["matrix_array_3x500"].T)
    want = GOLD[f"{func}_a{alpha}"]
    ex = oracle.exact.matrix(func, x, alpha=alpha)[0]
    assert _same_nan(ex, want)
    assert_allclose(ex, want, rtol=1e-10, atol=1e-13)
    rs = oracle.restated.matrix(func, x, alpha=alpha)[0]
    assert _same_nan(rs, want)
    assert_allclose(rs, want, rtol=1e-8, atol=1e-11)


@pytest.mark.parametrize("which", ["exact", "restated"])
def test_known_answers(which):
    """numbagg/test/test_matrix_functions.py:30-106, :202-220, :266-285,
    :344-434, :443-470."""
    o = getattr(oracle, which)
    data = np.array([[1, 2, 3, 4], [2, 4, 6, 8]], dtype=np.float64)
    assert_allclose(o.matrix("nancorrmatrix", data)[0], np.ones((2, 2)))
    cov = o.matrix("nancovmatrix", data)[0]
    assert_allclose(cov, cov.T)
    assert np.all(np.diag(cov) >= 0)
    const = np.array([[1, 1, 1, 1], [2, 2, 2, 2], [1, 2, 3, 4]], dtype=np.float64)
    c = o.matrix("nancorrmatrix", const)[0]
    assert np.isnan(c[0, 1]) and np.isnan(c[0, 2]) and c[2, 2] == 1.0
    v = o.matrix("nancovmatrix", const)[0]
    assert v[0, 0] == 0.0 and v[1, 1] == 0.0 and v[0, 1] == 0.0
    nans = np.array([[1, 2, np.nan, 4], [2, 4, 6, np.nan], [np.nan, 1, 2, 3]], dtype=np.float64)
    assert_allclose(np.diag(o.matrix("nancorrmatrix", nans)[0]), [1.0, 1.0, 1.0])

    rolling = np.array([[1, 2], [2, 4], [3, 6], [4, 8], [5, 10], [6, 12]], dtype=np.float64)
    r = o.matrix("move_corrmatrix", rolling, window=3, min_count=2)[0]
    for t in range(1, 6):
        assert_allclose(r[t], np.ones((2, 2)), rtol=1e-10)
    zero = np.array([[1, 2], [1, 2], [1, 2], [2, 3], [3, 4], [4, 5]], dtype=np.float64)
    assert np.isnan(o.matrix("move_corrmatrix", zero, window=3, min_count=2)[0][2, 0, 1])
    zc = o.matrix("move_covmatrix", zero, window=3, min_count=2)[0]
    assert zc[2, 0, 0] == 0.0 and zc[2, 1, 1] == 0.0 and zc[2, 0, 1] == 0.0
    mism = np.array([[1.0, 2.0], [2.0, 4.0], [3.0, np.nan], [np.nan, 8.0], [5.0, 10.0], [6.0, 12.0]])
    assert_allclose(o.matrix("move_corrmatrix", mism, window=6, min_count=2)[0][5, 0, 1], 1.0, rtol=1e-10)
    assert o.matrix("move_covmatrix", mism, window=6, min_count=2)[0][5, 0, 1] > 0

    ew = np.array([[1, 2], [2, 4], [3, 6], [4, 8]], dtype=np.float64)
    assert_allclose(o.matrix("move_exp_nancorrmatrix", ew, alpha=0.5)[0][-1], np.ones((2, 2)), rtol=1e-10)
    hi = o.matrix("move_exp_nancorrmatrix", ew, alpha=0.1, min_weight=0.8)[0]
    lo = o.matrix("move_exp_nancorrmatrix", ew, alpha=0.1, min_weight=0.1)[0]
    assert np.isnan(hi[0]).sum() >= np.isnan(lo[0]).sum()


def test_static_closed_forms():
    rng = np.random.default_rng(42)
    data = rng.standard_normal((5, 100))
    assert_allclose(oracle.exact.matrix("nancorrmatrix", data)[0], np.corrcoef(data), rtol=1e-12)
    assert_allclose(oracle.exact.matrix("nancovmatrix", data)[0], np.cov(data), rtol=1e-12)
    # brute force pairwise-complete statistics with NaNs and a batch axis
    x = rng.standard_normal((2, 4, 60))
    x[rng.random(x.shape) < 0.25] = np.nan
    cov = oracle.exact.matrix("nancovmatrix", x)
    corr = oracle.exact.matrix("nancorrmatrix", x)
    for b in range(2):
        for i in range(4):
            for j in range(4):
                ok = ~np.isnan(x[b, i]) & ~np.isnan(x[b, j])
                xi, xj = x[b, i, ok], x[b, j, ok]
                assert_allclose(cov[b, i, j], np.cov(xi, xj)[0, 1], rtol=1e-12)
                assert_allclose(corr[b, i, j], np.corrcoef(xi, xj)[0, 1], rtol=1e-12)


def test_pair_entries_equal_two_series_oracles():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((80, 3)) + 5
    x[rng.random(x.shape) < 0.15] = np.nan
    for func, one in (("move_covmatrix", "move_cov"), ("move_corrmatrix", "move_corr")):
        m = oracle.exact.matrix(func, x, window=9, min_count=4)[0]
        for i, j in ((0, 1), (2, 0), (1, 1)):
            mc = 4 if func == "move_covmatrix" else max(4, 2)
            want = oracle.exact.move(one, x[:, i][None], 9, mc, y=x[:, j][None])[0]
            assert_allclose(m[:, i, j], want, rtol=1e-14, equal_nan=True)
    for func, one in (("move_exp_nancovmatrix", "move_exp_nancov"), ("move_exp_nancorrmatrix", "move_exp_nancorr")):
        m = oracle.exact.matrix(func, x, alpha=0.3)[0]
        for i, j in ((0, 1), (2, 0)):
            want = oracle.exact.move_exp(one, x[:, i][None], 0.3, y=x[:, j][None])[0]
            assert_allclose(m[:, i, j], want, rtol=1e-14, equal_nan=True)


def test_restated_float32_accumulates_in_float32():
    """The reference's accumulators are the data dtype: on float32 data with a
    large level its rolling covariance drifts far from the exact value, which
    is why the GPU kernels are checked against the exact oracle."""
    rng = np.random.default_rng(0)
    x = (np.cumsum(rng.standard_normal((3000, 2)), axis=0) + 1e4).astype(np.float32)
    rs = oracle.restated.matrix("move_covmatrix", x, window=100, min_count=100)[0].astype(np.float64)
    ex = oracle.exact.matrix("move_covmatrix", x, window=100, min_count=100)[0]
    err = np.nanmax(np.abs(rs - ex) / np.abs(ex))
    assert err > 1e-4
