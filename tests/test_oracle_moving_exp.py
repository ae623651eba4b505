"""Pins the exponentially weighted oracles (CPU).

* C restatement and long-double exact oracle vs pandas ewm on the reference's
  fixture (the comparison in numbagg/test/test_moving_exp.py:19-31);
* the reference's known answers (test_moving_exp.py:111-120, :239-248,
  :323-345), min_weight counts (:34-79) and NaN masks (:160-236);
* a planted mistake (sum_weight_2 decayed by d instead of d**2) is caught.
"""

from __future__ import annotations

import numpy as np
import pytest
from numpy.testing import assert_allclose

from oracle import exact, restated

FUNCS1 = ["move_exp_nancount", "move_exp_nanmean", "move_exp_nansum", "move_exp_nanvar", "move_exp_nanstd"]
FUNCS2 = ["move_exp_nancov", "move_exp_nancorr"]


@pytest.mark.parametrize("func", FUNCS1 + FUNCS2)
@pytest.mark.parametrize("alpha", [0.5, 0.1])
def test_oracles_match_pandas(golden, func, alpha):
    a, a2 = golden["array_3x500"], golden["array2_3x500"]
    want = golden[f"{func}_a{alpha}"]
    y = a2 if func in FUNCS2 else None
    assert_allclose(restated.move_exp(func, a, alpha, y=y), want)
    assert_allclose(exact.move_exp(func, a, alpha, y=y), want)


KNOWN = [
    ("move_exp_nanmean", [10, 0, np.nan, 10], 0.5, [10.0, 3.3333333, 3.3333333, 8.1818182]),
    ("move_exp_nanmean", [10, 0, np.nan, 10], 0.25, [10.0, 4.2857143, 4.2857143, 7.1653543]),
    ("move_exp_nancount", [1, 0, np.nan, np.nan, 1, 0], 0.5, [1.0, 1.5, 0.75, 0.375, 1.1875, 1.59375]),
    ("move_exp_nancount", [1, 0, np.nan, np.nan, 1, 0], 0.25, [1.0, 1.75, 1.3125, 0.984375, 1.7382812, 2.3037109]),
    ("move_exp_nansum", [10, 0, np.nan, 10], 0.5, [10.0, 5.0, 2.5, 11.25]),
    ("move_exp_nansum", [10, 0, np.nan, 10], 0.25, [10.0, 7.5, 5.625, 14.21875]),
]


@pytest.mark.parametrize("func,x,alpha,want", KNOWN)
def test_known_answers(func, x, alpha, want):
    x = np.array(x, dtype=np.float64)
    assert_allclose(restated.move_exp(func, x, alpha)[0], want)
    assert_allclose(exact.move_exp(func, x, alpha)[0], want)


@pytest.mark.parametrize("alpha,want", [(0.5, [np.nan, 1.0, 0.8485281, 0.2274294]), (0.25, [np.nan, 1.0, 0.85, 0.4789468])])
def test_known_nancorr(alpha, want):
    a1 = np.array([10, 0, 5, 10.0])
    a2 = np.array([10, 0, 10, 5.0])
    assert_allclose(restated.move_exp("move_exp_nancorr", a1, alpha, y=a2)[0], want)
    assert_allclose(exact.move_exp("move_exp_nancorr", a1, alpha, y=a2)[0], want)


MIN_WEIGHT_CASES = [(0.0, 0.2, 19), (0.0, 0.8, 19), (0.5, 0.2, 17), (0.5, 0.8, 19), (0.9, 0.2, 10), (0.9, 0.8, 19), (1.0, 0.2, 0), (1.0, 0.8, 0)]


@pytest.mark.parametrize("func", ["move_exp_nancount", "move_exp_nanmean", "move_exp_nanstd", "move_exp_nansum", "move_exp_nanvar"])
def test_min_weight_counts(func):
    array = np.ones(25)
    array[:5] = np.nan
    for mw, alpha, expected in MIN_WEIGHT_CASES:
        for out in (restated.move_exp(func, array, alpha, mw)[0], exact.move_exp(func, array, alpha, mw)[0]):
            assert np.sum(~np.isnan(out)[6:]) == expected, (func, mw, alpha)


NAN_MASK_CASES = [
    ([np.nan, np.nan], [True, True], None),
    ([5.0, np.nan], [True, True], None),
    ([1.0, np.nan, 2.0], [True, True, False], None),
    ([1.0, np.nan, 1.0], [True, True, False], [True, True, True]),
    ([1.0, np.nan], [True, True], None),
    ([0.1, np.nan], [True, True], None),
    ([0.75, np.nan], [True, True], None),
    ([0.5, np.nan], [True, True], None),
    ([0.9, np.nan], [True, True], None),
    ([0.95, np.nan, 1.0], [True, True, False], None),
    ([0.59288027, np.nan, 0.4758262, 0.70877039], [True, True, False, False], None),
]


@pytest.mark.parametrize("func", ["move_exp_nanvar", "move_exp_nanstd", "move_exp_nancov", "move_exp_nancorr"])
@pytest.mark.parametrize("alpha", [0.1, 0.5, 0.9])
def test_nan_masks(func, alpha):
    # numbagg/test/test_moving_exp.py:160-236
    for x, want, want_corr in NAN_MASK_CASES:
        x = np.array(x)
        exp_mask = want_corr if (func == "move_exp_nancorr" and want_corr is not None) else want
        y = x if func in FUNCS2 else None
        assert list(np.isnan(restated.move_exp(func, x, alpha, y=y)[0])) == exp_mask, (func, x)
        assert list(np.isnan(exact.move_exp(func, x, alpha, y=y)[0])) == exp_mask, (func, x)


def test_min_weight_closed_form():
    # numbagg/test/test_moving_exp.py:82-108
    rs = np.random.RandomState(0)
    for n in (10, 200):
        array = rs.rand(n)
        array = np.where(array > 0.1, array, np.nan)
        for alpha in (0.1, 0.5, 0.9):
            w = np.array([(1 - alpha) ** (i - 1) for i in range(n, 0, -1)]) * alpha
            w = np.where(np.isnan(array), np.nan, w)
            weight = np.nansum(w)
            for oracle in (restated, exact):
                assert np.isnan(oracle.move_exp("move_exp_nanmean", array, alpha, weight + 0.01)[0, -1])
                assert not np.isnan(oracle.move_exp("move_exp_nanmean", array, alpha, weight - 0.01)[0, -1])


def test_planted_mistake_is_caught(golden):
    # sum_weight_2 decayed by d instead of d**2 would change the bias correction.
    a = golden["array_3x500"]
    x = a[0]
    alpha = 0.5
    s1 = s2 = sw = sw2 = 0.0
    out = []
    for v in x:
        d = 1 - alpha
        s1 *= d
        s2 *= d
        sw *= d
        sw2 *= d  # wrong: should be d**2
        if not np.isnan(v):
            s1 += v
            s2 += v * v
            sw += 1
            sw2 += 1
        vb = s2 / sw - (s1 / sw) ** 2 if sw else np.nan
        bias = 1 - sw2 / sw**2 if sw else np.nan
        out.append(vb / bias if bias > 0 else np.nan)
    with np.errstate(all="ignore"):
        assert not np.allclose(np.array(out), golden["move_exp_nanvar_a0.5"][0], equal_nan=True)
