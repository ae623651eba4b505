"""Pins the grouped oracles (CPU).

* C restatement and float64 closed forms vs pandas groupby on the reference's
  fixture (numbagg/test/test_grouped.py:137-166, incl. ddof 0/1/5);
* the reference's known answers (test_grouped.py:229-246, :359-394, :510-590);
* bit-exact integer/index outputs agree between the two oracles;
* a planted mistake (last-index tie rule for argmax) is caught.
"""

from __future__ import annotations

import numpy as np
import pytest
from numpy.testing import assert_allclose, assert_array_equal

from oracle import exact, restated

FUNCS = ["group_nansum", "group_nanmean", "group_nancount", "group_nanfirst", "group_nanlast", "group_nanargmax", "group_nanargmin"]


@pytest.mark.parametrize("func", FUNCS)
def test_oracles_match_pandas(golden, func):
    g = golden["group_array_1x500"]
    codes = golden["group_codes"]
    want = golden[func]
    assert_allclose(restated.group(func, g, codes, 12), want)
    assert_allclose(exact.group(func, g, codes, 12), want)


@pytest.mark.parametrize("func", ["group_nanvar", "group_nanstd"])
@pytest.mark.parametrize("ddof", [0, 1, 5])
def test_var_oracles_match_pandas(golden, func, ddof):
    g = golden["group_array_1x500"]
    codes = golden["group_codes"]
    want = golden[f"{func}_ddof{ddof}"]
    assert_allclose(restated.group(func, g, codes, 12, ddof=ddof), want)
    assert_allclose(exact.group(func, g, codes, 12, ddof=ddof), want)


def test_known_answers():
    # all-NaN labels: sum 0, argmin NaN (test_grouped.py:229-246)
    v = np.array([[np.nan, np.nan, np.nan]])
    lab = np.array([0, 1, 2])
    assert_array_equal(restated.group("group_nansum", v, lab, 3), [[0, 0, 0]])
    assert np.all(np.isnan(restated.group("group_nanargmin", v, lab, 3)))
    # integer values (test_grouped.py:359-394)
    iv = np.array([[1, 2, 3, 4, 5]], dtype=np.int32)
    assert_array_equal(restated.group("group_nancount", iv, np.array([0, 0, 0, 1, 1]), 2), [[3, 2]])
    fv = np.array([[1, 2, 3, 4, 5, 6]], dtype=np.float64)
    assert_allclose(restated.group("group_nanmean", fv, np.array([0, 0, 0, 1, 1, 2]), 3), [[2.0, 4.5, 6.0]])
    # argmax/argmin with flattened 2-D labels (test_grouped.py:543-590)
    values = np.array([[1.0, 5.0, 3.0, 2.0, 4.0, 6.0]])
    labels = np.array([0, 0, 1, 1, 2, 2])
    assert_array_equal(restated.group("group_nanargmax", values, labels, 3), [[1.0, 2.0, 5.0]])
    assert_array_equal(restated.group("group_nanargmin", values, labels, 3), [[0.0, 3.0, 4.0]])
    assert_array_equal(exact.group("group_nanargmax", values, labels, 3), [[1.0, 2.0, 5.0]])
    # argmax with (lat, lon) labels over a (3, 4) core (test_grouped.py:510-540)
    v3 = np.arange(12.0).reshape(1, 12)
    l3 = np.array([[0, 0, 1, 1], [0, 1, 1, 2], [2, 2, 2, 2]]).ravel()
    assert_array_equal(restated.group("group_nanargmax", v3, l3, 3), [[4.0, 6.0, 11.0]])
    assert_array_equal(restated.group("group_nanargmin", v3, l3, 3), [[0.0, 2.0, 7.0]])


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32, np.int64])
def test_integer_outputs_agree(dtype):
    rng = np.random.default_rng(4)
    v = (rng.standard_normal((3, 2000)) * 10).astype(dtype)
    if np.issubdtype(dtype, np.floating):
        v[rng.random(v.shape) < 0.2] = np.nan
        v[0, :50] = 1.5  # ties: the first index wins
    lab = rng.integers(-1, 40, size=2000)
    for func in ("group_nancount", "group_nanargmax", "group_nanargmin", "group_nanfirst", "group_nanlast"):
        if np.issubdtype(dtype, np.integer) and func in ("group_nanargmax", "group_nanargmin", "group_nanfirst", "group_nanlast"):
            # exact.group works in float64 and encodes "empty" as NaN
            got = restated.group(func, v, lab, 40).astype(np.float64)
            want = exact.group(func, v.astype(np.float64), lab, 40)
            empty = np.isnan(want)
            assert_array_equal(got[~empty], want[~empty])
            continue
        assert_array_equal(restated.group(func, v, lab, 40), exact.group(func, v, lab, 40).astype(v.dtype))


def test_planted_mistake_is_caught():
    values = np.array([[2.0, 1.0, 2.0, 0.5]])
    labels = np.array([0, 0, 0, 0])
    right = restated.group("group_nanargmax", values, labels, 1)
    # "last index wins" would give 2
    assert right[0, 0] == 0.0
    assert exact.group("group_nanargmax", values, labels, 1)[0, 0] == 0.0


@pytest.mark.parametrize("func", ["group_nanvar", "group_nanstd"])
def test_negative_ddof_empty_label_is_nan(func):
    # count 0, ddof -1: denom 1 > 0, but (0 - 0**2/0) / 1 is NaN (grouped.py:166-174)
    v = np.array([[1.0, 2.0, 4.0, np.nan]])
    lab = np.array([0, 0, 2, 1])
    for oracle_group in (restated.group, exact.group):
        got = oracle_group(func, v, lab, 3, ddof=-1)
        assert np.isnan(got[0, 1])
        want0 = 0.5 / 3 if func == "group_nanvar" else np.sqrt(0.5 / 3)
        assert_allclose(got[0, 0], want0)
        assert got[0, 2] == 0.0
