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


NEW_FUNCS = ["group_nanmin", "group_nanmax", "group_nanprod", "group_nansum_of_squares", "group_nanany", "group_nanall"]


@pytest.mark.parametrize("func", NEW_FUNCS)
def test_rest_of_family_oracles_match_pandas(golden, func):
    # conftest.py:444-487: pandas groupby min/max/prod/any/all and (x**2).sum()
    g = golden["group_array_1x500"]
    codes = golden["group_codes"]
    want = golden[func]
    assert_allclose(restated.group(func, g, codes, 12), want)
    assert_allclose(exact.group(func, g, codes, 12), want)


@pytest.mark.parametrize("dt", ["float64", "int32", "bool"])
def test_series_comparisons(golden, dt):
    # test_grouped.py:116-135: float64 / int32 / bool values, labels with -1
    vals = golden[f"series_{dt}_values"]
    labels = golden[f"series_{dt}_labels"]
    kdt = np.int32 if dt == "bool" else vals.dtype  # bool input runs the int32 loop
    v = vals.astype(kdt)[None, :]
    for func in NEW_FUNCS + ["group_nansum", "group_nancount", "group_nanfirst", "group_nanlast"]:
        if dt == "int32" and func == "group_nanprod":
            continue  # "group_nanprod result too large" (test_grouped.py:127-128)
        want = golden[f"series_{dt}_{func}"]
        got = restated.group(func, v, labels, 5)[0]
        assert_allclose(got.astype(np.float64), want, err_msg=func)
        if dt == "float64":
            assert_allclose(exact.group(func, v, labels, 5)[0], want, err_msg=func)


@pytest.mark.parametrize("func", NEW_FUNCS)
def test_rest_of_family_known_answers(func):
    # test_grouped.py:229-262: all-NaN label, no valid labels, single NaN label
    expect = {
        "group_nanmin": ([1.0, np.nan], [np.nan], [1.0, np.nan], [np.nan] * 3),
        "group_nanmax": ([2.0, np.nan], [np.nan], [2.0, np.nan], [np.nan] * 3),
        "group_nanprod": ([2.0, 1.0], [1.0], [2.0, 1.0], [1.0] * 3),
        "group_nansum_of_squares": ([5.0, 0.0], [0.0], [5.0, 0.0], [0.0] * 3),
        "group_nanany": ([1.0, 0.0], [0.0], [1.0, 0.0], [0.0] * 3),
        "group_nanall": ([1.0, 1.0], [1.0], [1.0, 1.0], [1.0] * 3),
    }[func]
    cases = [
        (np.array([1.0, 2.0, np.nan, np.nan]), np.array([0, 0, 1, 1]), 2),
        (np.array([1.0, 2.0, 3.0, 4.0]), np.array([-1, -1, -1, -1]), 1),
        (np.array([1.0, 2.0, np.nan]), np.array([0, 0, 1]), 2),
        (np.array([np.nan, np.nan, np.nan]), np.array([0, 1, 2]), 3),
    ]
    for (v, lab, L), want in zip(cases, expect):
        assert_allclose(restated.group(func, v[None, :], lab, L)[0], want)
        assert_allclose(exact.group(func, v[None, :], lab, L)[0], want)


def test_min_max_keep_the_first_signed_zero():
    # `value < min` is strict: among equal extremes (0.0 == -0.0) the first stays
    v = np.array([[0.0, -0.0, 1.0, -0.0, 0.0, 2.0]])
    lab = np.array([0, 0, 0, 1, 1, 1])
    for oracle_group in (restated.group, exact.group):
        mn = oracle_group("group_nanmin", v, lab, 2)[0]
        assert np.signbit(mn).tolist() == [False, True]
    # planted mistake: "last equal extreme wins" gives the other signs
    assert np.signbit(v[0, [1, 4]]).tolist() == [True, False]


def test_integer_products_and_squares_wrap():
    v = np.array([[2**20, 2**20, 3, 2**16, 2**16]], dtype=np.int32)
    lab = np.array([0, 0, 0, 1, 1])
    prod = restated.group("group_nanprod", v, lab, 2)[0]
    assert prod.tolist() == [np.int32(np.int64(2**40 * 3) % 2**32), 0]
    sq = restated.group("group_nansum_of_squares", v, lab, 2)[0]
    assert sq.tolist() == [np.int32((2 * 2**40 + 9) % 2**32), np.int32((2 * 2**32) % 2**32)]
