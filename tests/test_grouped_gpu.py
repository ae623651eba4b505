"""GPU parity of the grouped reductions.

count, argmax, argmin, first and last are bit-exact against the oracles (C
restatement and closed form). Float sums/means are compared to the float64
oracle: float32 outputs within 1 float32 ULP of the float64 result (float64
accumulation, one final rounding), float64 outputs within 1e-13 relative of
the sum of |x|; var/std within 1e-12 relative (float64 two-pass: means, then
squared deviations).
Both group-by strategies (LDS-privatised and global atomics) are covered.
"""

from __future__ import annotations

import numpy as np
import pytest
from numpy.testing import assert_allclose, assert_array_equal

pytestmark = pytest.mark.gpu

EXACT_FUNCS = ["group_nancount", "group_nanargmax", "group_nanargmin", "group_nanfirst", "group_nanlast"]
FLOAT_FUNCS = ["group_nansum", "group_nanmean", "group_nanvar", "group_nanstd"]


def _data(rows, n, L, dtype, seed, zipf=False, p_nan=0.05):
    rng = np.random.default_rng(seed)
    if np.issubdtype(dtype, np.floating):
        v = rng.standard_normal((rows, n)).astype(dtype)
        v[rng.random(v.shape) < p_nan] = np.nan
    else:
        v = rng.integers(-1000, 1000, size=(rows, n)).astype(dtype)
    if zipf:
        p = np.arange(1, L + 1, dtype=np.float64) ** -1.1
        lab = rng.choice(L, size=n, p=p / p.sum()).astype(np.int32)
    else:
        lab = rng.integers(-1, L, size=n).astype(np.int32)
    return v, lab


@pytest.fixture(params=["lds", "global"])
def strategy(request, monkeypatch):
    monkeypatch.setenv("NBG_GROUP_STRATEGY", request.param)
    return request.param


def _call(func, *args, **kw):
    """Run a grouped function; a forced LDS strategy whose per-label
    accumulators exceed the LDS budget must fail with the documented error."""
    import os

    import numbagg_amd as nb

    try:
        return getattr(nb, func)(*args, **kw)
    except RuntimeError as e:
        if os.environ.get("NBG_GROUP_STRATEGY") == "lds" and "do not fit the LDS strategy" in str(e):
            pytest.skip("label count exceeds the LDS budget (forced strategy rejected as documented)")
        raise


@pytest.mark.parametrize("func", EXACT_FUNCS)
@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32, np.int64])
@pytest.mark.parametrize("shape,L,zipf", [((3, 1000), 12, False), ((4, 300_000), 1000, True), ((2, 70_001), 5000, False)])
def test_group_exact_outputs(gpu_lib, strategy, func, dtype, shape, L, zipf):
    import numbagg_amd as nb
    from oracle import restated

    v, lab = _data(*shape, L, dtype, seed=shape[1] + L, zipf=zipf)
    if np.issubdtype(dtype, np.floating):
        v[:, 10:40] = 0.25  # ties
        v[0, 100] = -0.0
        v[0, 101] = 0.0
    got = _call(func, v, lab, num_labels=L, axis=-1)
    want = restated.group(func, v, lab.astype(np.int64), L)
    assert got.dtype == v.dtype
    assert_array_equal(got.view(np.uint8), want.view(np.uint8))


@pytest.mark.parametrize("func", FLOAT_FUNCS)
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape,L,zipf", [((3, 1000), 12, False), ((4, 300_000), 1000, True), ((2, 70_001), 9000, False)])
def test_group_float_outputs(gpu_lib, strategy, func, dtype, shape, L, zipf):
    import numbagg_amd as nb
    from oracle import exact

    v, lab = _data(*shape, L, dtype, seed=shape[1] + 3 * L, zipf=zipf)
    got = _call(func, v, lab, num_labels=L, axis=-1)
    _check_float(func, dtype, v, lab, L, got)


def _check_float(func, dtype, v, lab, L, got):
    from oracle import exact

    want = exact.group(func, v, lab, L)
    assert np.array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    if func in ("group_nansum", "group_nanmean") and dtype == np.float32:
        e = want[ok].astype(np.float32)
        ulp = np.spacing(np.abs(e)).astype(np.float64)
        err = np.abs(got[ok].astype(np.float64) - want[ok])
        # float64 accumulation of float32 values is within 1e-15 of sum|x|; after
        # rounding once that is at most 1 ULP off the rounded float64 oracle.
        sabs = exact.group("group_nansum", np.abs(v), lab, L)[ok]
        cnt = exact.group("group_nancount", v, lab, L)[ok]
        scale = sabs / (cnt if func == "group_nanmean" else 1.0)
        assert np.all(err <= ulp + 1e-14 * scale)
    elif func in ("group_nansum", "group_nanmean"):
        sabs = exact.group("group_nansum", np.abs(v), lab, L)[ok]
        cnt = exact.group("group_nancount", v, lab, L)[ok]
        scale = sabs / (cnt if func == "group_nanmean" else 1.0)
        assert np.all(np.abs(got[ok] - want[ok]) <= 1e-13 * scale + 1e-300)
    else:
        rtol = 2e-7 if dtype == np.float32 else 1e-12
        assert_allclose(got[ok], want[ok], rtol=rtol)


@pytest.mark.parametrize("rows,unroll,pipe", [("4", "4", "0"), ("1", "2", "0"), ("2", "2", "0"), ("2", "4", "1"),
                                              ("4", "2", "1"), ("1", "4", "0")])
@pytest.mark.parametrize("func", FLOAT_FUNCS)
def test_group_sum_pass_shapes(gpu_lib, monkeypatch, rows, unroll, pipe, func):
    """Every shape of the sum pass (rows per workgroup, vectors per thread,
    pipelined loads) meets the same bars; 9 rows leave partial row groups."""
    import numbagg_amd as nb

    monkeypatch.setenv("NBG_GROUP_ROWS", rows)
    monkeypatch.setenv("NBG_GROUP_UNROLL", unroll)
    monkeypatch.setenv("NBG_GROUP_PIPE", pipe)
    for dtype in (np.float32, np.float64):
        v, lab = _data(9, 300_000, 1000, dtype, seed=int(rows) * 10 + int(unroll), zipf=True)
        got = getattr(nb, func)(v, lab, num_labels=1000, axis=-1)
        _check_float(func, dtype, v, lab, 1000, got)


@pytest.mark.parametrize("per_cu", ["0", "1", "3", "8"])
def test_group_workgroups_per_cu(gpu_lib, monkeypatch, per_cu):
    """Grids of any number of workgroups per CU (longer or shorter chunks per
    workgroup) give the same bits for the exact functions and meet the same
    bars for the float ones."""
    import numbagg_amd as nb
    from oracle import restated

    monkeypatch.setenv("NBG_GROUP_BLOCKS_PER_CU", per_cu)
    v, lab = _data(7, 500_000, 700, np.float32, seed=int(per_cu) + 40, zipf=True)
    for func in EXACT_FUNCS:
        got = getattr(nb, func)(v, lab, num_labels=700, axis=-1)
        want = restated.group(func, v, lab.astype(np.int64), 700)
        assert_array_equal(got.view(np.uint8), want.view(np.uint8))
    for func in FLOAT_FUNCS:
        _check_float(func, np.float32, v, lab, 700, getattr(nb, func)(v, lab, num_labels=700, axis=-1))


@pytest.mark.parametrize("hot", ["1", "0"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_group_hot_labels(gpu_lib, monkeypatch, hot, dtype):
    """Labels 0 and 1 accumulated in registers (or not) meet the same bars:
    skewed and uniform labels, a single label (no label 1), 5 rows (a partial
    row group), out-of-range labels."""
    import numbagg_amd as nb

    monkeypatch.setenv("NBG_GROUP_HOT", hot)
    for L, zipf in ((1000, True), (1000, False), (1, False), (2, True)):
        v, lab = _data(5, 200_003, L, dtype, seed=L + (7 if zipf else 0), zipf=zipf)
        if L == 1:
            lab = np.where(lab > 0, 0, lab).astype(np.int32)
        for func in FLOAT_FUNCS:
            got = getattr(nb, func)(v, lab, num_labels=L, axis=-1)
            _check_float(func, dtype, v, lab, L, got)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_group_var_offset_data(gpu_lib, dtype):
    """Data far from zero (mean 1e4, unit spread): the per-label means come
    from float64 sums, so the squared deviations keep the variance to the
    same bars as for centred data."""
    import numbagg_amd as nb

    rng = np.random.default_rng(7)
    v = (1e4 + rng.standard_normal((3, 200_000))).astype(dtype)
    v[rng.random(v.shape) < 0.05] = np.nan
    lab = rng.integers(-1, 300, 200_000).astype(np.int32)
    for func in ("group_nanvar", "group_nanstd"):
        got = getattr(nb, func)(v, lab, num_labels=300, axis=-1)
        _check_float(func, dtype, v, lab, 300, got)


def test_group_ddof(gpu_lib):
    import numbagg_amd as nb
    from oracle import exact

    v, lab = _data(2, 5000, 30, np.float64, seed=1)
    for ddof in (0, 1, 5):
        for func in ("group_nanvar", "group_nanstd"):
            got = getattr(nb, func)(v, lab, num_labels=30, axis=-1, ddof=ddof)
            assert_allclose(got, exact.group(func, v, lab, 30, ddof=ddof), rtol=1e-12)


def test_group_reference_pandas_vectors(gpu_lib, golden):
    import numbagg_amd as nb

    g = golden["group_array_1x500"]
    codes = golden["group_codes"]
    for func in ["group_nansum", "group_nanmean", "group_nancount", "group_nanfirst", "group_nanlast", "group_nanargmax", "group_nanargmin"]:
        assert_allclose(getattr(nb, func)(g, codes, num_labels=12, axis=-1), golden[func])
    for ddof in (0, 1, 5):
        assert_allclose(nb.group_nanvar(g, codes, num_labels=12, axis=-1, ddof=ddof), golden[f"group_nanvar_ddof{ddof}"])
        assert_allclose(nb.group_nanstd(g, codes, num_labels=12, axis=-1, ddof=ddof), golden[f"group_nanstd_ddof{ddof}"])


def test_group_known_answers_and_axes(gpu_lib):
    import numbagg_amd as nb

    values = np.array([[1.0, 5.0, 3.0], [2.0, 4.0, 6.0]])
    labels = np.array([[0, 0, 1], [1, 2, 2]])
    assert_allclose(nb.group_nanargmax(values, labels), [1.0, 2.0, 5.0])
    assert_allclose(nb.group_nanargmin(values, labels), [0.0, 3.0, 4.0])
    v3 = np.arange(24.0).reshape(2, 3, 4)
    l3 = np.array([[0, 0, 1, 1], [0, 1, 1, 2], [2, 2, 2, 2]])
    r = nb.group_nanargmax(v3, l3, axis=(1, 2))
    assert r.shape == (2, 3)
    assert_allclose(r[0], [4.0, 6.0, 11.0])
    assert_allclose(nb.group_nanargmin(v3, l3, axis=(1, 2))[0], [0.0, 2.0, 7.0])
    # empty groups
    v = np.array([np.nan, np.nan, np.nan])
    g = np.array([0, 1, 2])
    np.testing.assert_equal(nb.group_nansum(v, g), [0, 0, 0])
    assert np.all(np.isnan(nb.group_nanargmin(v, g)))
    assert np.all(np.isnan(nb.group_nanmean(v, g)))
    # no valid labels
    r = nb.group_nansum(np.array([1.0, 2.0, 3.0, 4.0]), np.array([-1, -1, -1, -1]), num_labels=1)
    assert_allclose(r, [0.0])
    # integer values, int mean and bool mean (test_grouped.py:359-394, :606-612)
    assert_allclose(nb.group_nancount(np.array([1, 2, 3, 4, 5], dtype=np.int32), np.array([0, 0, 0, 1, 1], dtype=np.int32)), [3, 2])
    assert_allclose(nb.group_nanmean(np.array([1, 2, 3, 4, 5, 6], dtype=np.int32), np.array([0, 0, 0, 1, 1, 2], dtype=np.int32)), [2.0, 4.5, 6.0])
    assert_allclose(nb.group_nanmean(np.array([True, True, False]), np.array([0, 0, 0])), [2 / 3])
    # 1-D labels along an axis (test_grouped.py:264-303)
    values2d = np.arange(25.0).reshape(5, 5)
    assert_allclose(nb.group_nansum(values2d, np.arange(5), axis=1), values2d)
    assert_allclose(nb.group_nansum(values2d, np.arange(5), axis=0), values2d.T)
    got = nb.group_nansum(values2d, [0, 0, -1, 1, 1], axis=1)
    assert_allclose(got, np.stack([values2d[:, :2].sum(1), values2d[:, 3:].sum(1)], axis=-1))
    # narrow label dtypes and wrap-around integer sums
    for ldt in (np.int8, np.int16, np.int32, np.int64):
        assert_allclose(nb.group_nansum(np.arange(5.0), np.arange(5, dtype=ldt)), np.arange(5.0))
    big = np.full(10, 2**30, dtype=np.int32)
    assert nb.group_nansum(big, np.zeros(10, dtype=np.int32))[0] == np.int32(np.int64(10 * 2**30) % 2**32 - 2**32)


def test_group_zipf_bench_shape(gpu_lib):
    """A scaled-down version of the (64, 4M) 1000-label benchmark, uniform and
    Zipf(1.1) labels, checked for every function."""
    import numbagg_amd as nb
    from oracle import exact, restated

    for zipf in (False, True):
        v, lab = _data(4, 1_000_000, 1000, np.float32, seed=7, zipf=zipf)
        for func in EXACT_FUNCS:
            got = getattr(nb, func)(v, lab, num_labels=1000, axis=-1)
            assert_array_equal(got.view(np.uint32), restated.group(func, v, lab, 1000).view(np.uint32))
        for func in FLOAT_FUNCS:
            got = getattr(nb, func)(v, lab, num_labels=1000, axis=-1)
            want = exact.group(func, v, lab, 1000)
            assert_allclose(got, want, rtol=1e-6, atol=1e-6)


ALL_GROUP_FUNCS = EXACT_FUNCS + FLOAT_FUNCS


@pytest.mark.parametrize("func", ALL_GROUP_FUNCS)
def test_group_empty_core_axis(gpu_lib, func):
    """An empty reduced axis with explicit num_labels gives the empty-label
    outputs (0 for sum/count, NaN otherwise) in every axis mode, like the
    reference (which reduces nothing)."""
    import numbagg_amd as nb

    empty_label = 0.0 if func in ("group_nansum", "group_nancount") else np.nan
    for dtype in (np.float32, np.float64):
        v = np.zeros((4, 0), dtype=dtype)
        got = getattr(nb, func)(v, np.zeros(0, np.int64), axis=-1, num_labels=3)
        assert got.shape == (4, 3) and got.dtype == dtype
        np.testing.assert_array_equal(got, np.full((4, 3), empty_label, dtype=dtype))
        got = getattr(nb, func)(np.zeros((2, 0, 5), dtype=dtype), np.zeros((0, 5), np.int32), axis=(1, 2), num_labels=2)
        np.testing.assert_array_equal(got, np.full((2, 2), empty_label, dtype=dtype))
        got = getattr(nb, func)(np.zeros((0, 5), dtype=dtype), np.zeros((0, 5), np.int32), num_labels=2)
        np.testing.assert_array_equal(got, np.full(2, empty_label, dtype=dtype))


def test_group_negative_ddof_empty_label(gpu_lib):
    """ddof < 0 leaves count - ddof > 0 for an empty label; the reference's
    (0 - 0**2/0) / (0 - ddof) is still NaN there."""
    import numbagg_amd as nb
    from oracle import exact

    v, lab = _data(2, 5000, 30, np.float64, seed=2)
    lab[lab == 7] = 3  # label 7 empty
    for func in ("group_nanvar", "group_nanstd"):
        for dtype in (np.float32, np.float64):
            vv = v.astype(dtype)
            got = getattr(nb, func)(vv, lab, num_labels=30, axis=-1, ddof=-1)
            assert np.isnan(got[:, 7]).all()
            want = exact.group(func, vv, lab, 30, ddof=-1)
            ok = ~np.isnan(want)
            assert np.array_equal(np.isnan(got), np.isnan(want))
            assert_allclose(got[ok], want[ok], rtol=2e-7 if dtype == np.float32 else 1e-12)


REST_EXACT = ["group_nanmin", "group_nanmax", "group_nanany", "group_nanall"]


@pytest.mark.parametrize("func", REST_EXACT)
@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32, np.int64])
@pytest.mark.parametrize("shape,L,zipf", [((3, 1000), 12, False), ((4, 300_000), 1000, True), ((2, 70_001), 5000, False)])
def test_group_min_max_any_all_exact(gpu_lib, strategy, func, dtype, shape, L, zipf):
    """min/max copy the value at the first index of the extreme (so signed
    zeros and 64-bit integers come out as the reference's float64 running
    extreme leaves them); any/all are flags. All bit-exact."""
    from oracle import restated

    v, lab = _data(*shape, L, dtype, seed=shape[1] + 5 * L, zipf=zipf)
    if np.issubdtype(dtype, np.floating):
        v[:, 10:40] = 0.25  # ties
        v[0, 100], v[0, 101], v[1, 102] = -0.0, 0.0, -0.0
        v[:, 200:260] = 0.0  # zeros for any/all
        v[:, 300:340] = np.nan
    else:
        v[:, 200:260] = 0
        if dtype == np.int64:
            v[0, 500:504] = [2**53 + 1, 2**53, -(2**60) - 1, 2**60 + 3]  # round through float64
    got = _call(func, v, lab, num_labels=L, axis=-1)
    want = restated.group(func, v, lab.astype(np.int64), L)
    assert got.dtype == v.dtype
    assert_array_equal(got.view(np.uint8), want.view(np.uint8))


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32, np.int64])
@pytest.mark.parametrize("shape,L,zipf", [((3, 1000), 12, False), ((4, 300_000), 1000, True), ((2, 70_001), 5000, False)])
def test_group_prod_and_sum_of_squares(gpu_lib, strategy, dtype, shape, L, zipf):
    """Integer products and sums of squares wrap like the reference's
    fixed-width arithmetic (bit-exact against the restatement). Float ones are
    formed in float64 and rounded once: float32 within 1 ULP of the float64
    oracle plus 2 n 2**-53 of the value (float64 products of n factors in any
    order), float64 within 4 n 2**-53 relative."""
    from oracle import exact, restated

    rng = np.random.default_rng(shape[1] + L)
    v, lab = _data(*shape, L, dtype, seed=shape[1] + 7 * L, zipf=zipf)
    if np.issubdtype(dtype, np.floating):
        vp = (1.0 + 0.02 * rng.standard_normal(v.shape)).astype(dtype)  # products stay O(1)
        vp[np.isnan(v)] = np.nan
        vp[0, 7] = 0.0
    else:
        vp = v
    for func, data in (("group_nanprod", vp), ("group_nansum_of_squares", v)):
        got = _call(func, data, lab, num_labels=L, axis=-1)
        assert got.dtype == data.dtype
        if not np.issubdtype(dtype, np.floating):
            want = restated.group(func, data, lab.astype(np.int64), L)
            assert_array_equal(got, want)
            continue
        want = exact.group(func, data, lab, L)
        cnt = exact.group("group_nancount", data, lab, L)
        g64 = got.astype(np.float64)
        err = np.abs(g64 - want)
        if dtype == np.float32:
            ulp = np.spacing(np.abs(want.astype(np.float32))).astype(np.float64)
            assert np.all(err <= ulp + 2 * (cnt + 1) * 2.0**-53 * np.abs(want)), func
        else:
            assert np.all(err <= 4 * (cnt + 1) * 2.0**-53 * np.abs(want) + 1e-300), func


def test_group_rest_of_family_pandas_vectors(gpu_lib, golden):
    """The reference's pandas comparisons (conftest.py:444-487 on the (1, 500)
    fixture; test_grouped.py:116-135 on float64 / int32 / bool series)."""
    import numbagg_amd as nb

    g = golden["group_array_1x500"]
    codes = golden["group_codes"]
    for func in ("group_nanmin", "group_nanmax", "group_nanprod", "group_nansum_of_squares", "group_nanany",
                 "group_nanall"):
        assert_allclose(getattr(nb, func)(g, codes, num_labels=12, axis=-1), golden[func], err_msg=func)
    for dt in ("float64", "int32", "bool"):
        vals = golden[f"series_{dt}_values"]
        labels = golden[f"series_{dt}_labels"]
        for func in ("group_nanall", "group_nanany", "group_nancount", "group_nanfirst", "group_nanlast",
                     "group_nanmax", "group_nanmean", "group_nanmin", "group_nanprod", "group_nanstd",
                     "group_nansum", "group_nanvar", "group_nansum_of_squares"):
            if dt == "int32" and func == "group_nanprod":
                continue  # "group_nanprod result too large" (test_grouped.py:127-128)
            if dt == "bool" and func in ("group_nanstd", "group_nanvar"):
                with pytest.raises(TypeError, match="does not support boolean input"):
                    getattr(nb, func)(vals, labels)
                continue
            got = getattr(nb, func)(vals, labels)
            np.testing.assert_almost_equal(got, golden[f"series_{dt}_{func}"], err_msg=f"{dt} {func}")


@pytest.mark.parametrize("func", ["group_nanmin", "group_nanmax", "group_nanprod", "group_nansum_of_squares",
                                  "group_nanany", "group_nanall"])
def test_group_rest_of_family_edge_cases(gpu_lib, func):
    """test_grouped.py:229-262 (all-NaN label, no valid labels, single NaN
    label, all values NaN) and an empty reduced axis."""
    import numbagg_amd as nb
    from oracle import exact

    cases = [
        (np.array([1.0, 2.0, np.nan, np.nan]), np.array([0, 0, 1, 1])),
        (np.array([1.0, 2.0, 3.0, 4.0]), np.array([-1, -1, -1, -1])),
        (np.array([1.0, 2.0, np.nan]), np.array([0, 0, 1])),
        (np.array([np.nan, np.nan, np.nan]), np.array([0, 1, 2])),
    ]
    for v, lab in cases:
        L = max(int(lab.max()) + 1, 1)
        np.testing.assert_array_equal(getattr(nb, func)(v, lab, num_labels=L), exact.group(func, v[None, :], lab, L)[0])
    empty = {"group_nanprod": 1.0, "group_nanall": 1.0, "group_nansum_of_squares": 0.0, "group_nanany": 0.0}
    got = getattr(nb, func)(np.zeros((3, 0), np.float32), np.zeros(0, np.int32), axis=-1, num_labels=2)
    np.testing.assert_array_equal(got, np.full((3, 2), empty.get(func, np.nan), np.float32))
