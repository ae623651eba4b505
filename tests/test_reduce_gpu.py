"""GPU parity of the plain reductions (numbagg/funcs.py:23-228).

The reference's own comparison (numbagg/test/test_funcs.py:129-227): every
array of its generator (numbagg/test/util.py:26-120: NaN/inf patterns, byte
order, float16, 0-d input, zero-size shapes, random arrays up to 1e6
elements, non-contiguous views) along every axis and axis=None, against
NumPy's nan-functions: the same error messages where both raise, the same
dtype, equal results (nansum/nanmax/nanmin/arg*/nancount exactly — to
NumPy's own precision for float sums —, nanmean/nanvar/nanstd to the
reference's decimals). On top: the C restatement of the reference kernels,
float32 sums within 1 ULP of the float64 sum, and bitwise repeatability.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
from numpy.testing import assert_allclose, assert_array_almost_equal, assert_array_equal, assert_equal

pytestmark = pytest.mark.gpu


def _slow_count(x, axis=None):
    return np.sum(~np.isnan(x), axis=axis)


def _fsum(a, axis):
    """Exactly rounded float64 sums of the non-NaN values over `axis` (and of
    their magnitudes); +-inf handled as NumPy does."""
    a = np.asarray(a, dtype=np.float64)
    axes = tuple(range(a.ndim)) if axis is None else ((axis,) if np.isscalar(axis) else tuple(axis))
    axes = tuple(ax % max(a.ndim, 1) for ax in axes) if a.ndim else ()
    m = np.moveaxis(a, axes, tuple(range(a.ndim - len(axes), a.ndim))) if axes else a
    outer = m.shape[: a.ndim - len(axes)]
    inner = int(np.prod(m.shape[a.ndim - len(axes):], dtype=np.int64))
    flat = m.reshape(int(np.prod(outer, dtype=np.int64)), inner) if a.ndim else m.reshape(1, 1)
    s = np.empty(flat.shape[0])
    sa = np.empty(flat.shape[0])
    for i, row in enumerate(flat):
        row = row[~np.isnan(row)]
        fin = row[np.isfinite(row)]
        inf = row[~np.isfinite(row)]
        s[i] = np.sum(inf) if inf.size else math.fsum(fin)
        sa[i] = math.fsum(np.abs(fin))
    return s.reshape(outer), sa.reshape(outer)


def _check_sum(actual, want, sabs, tag):
    """float32: within 1 ULP of the exact sum; float64: within 2**-52 of it
    relative plus 2**-100 of the sum of |x| (double-double accumulation)."""
    dt = np.asarray(actual).dtype
    g = np.atleast_1d(np.asarray(actual, dtype=np.float64)).ravel()
    want = np.atleast_1d(np.asarray(want)).ravel()
    sabs = np.atleast_1d(np.asarray(sabs)).ravel()
    assert g.shape == want.shape, tag
    fin = np.isfinite(want)
    assert np.array_equal(g[~fin], want[~fin], equal_nan=True), tag
    err = np.abs(g[fin] - want[fin])
    if dt == np.float32:
        bar = np.spacing(np.abs(want[fin]).astype(np.float32)).astype(np.float64) + 1e-14 * sabs[fin]
    else:
        bar = 2.0**-52 * np.abs(want[fin]) + 2.0**-100 * sabs[fin]
    assert np.all(err <= bar), (tag, float(np.max(err - bar)))


def _functions():
    import numbagg_amd as nb

    yield nb.nansum, np.nansum, np.inf
    yield nb.nanmax, np.nanmax, np.inf
    yield nb.nanargmin, np.nanargmin, np.inf
    yield nb.nanargmax, np.nanargmax, np.inf
    yield nb.nanmin, np.nanmin, np.inf
    yield nb.nanmean, np.nanmean, 5
    yield nb.nanmean, np.nanmean, 1
    yield nb.nanstd, lambda x, axis=None: np.nanstd(x, axis=axis, ddof=1), 5
    yield nb.nanvar, lambda x, axis=None: np.nanvar(x, axis=axis, ddof=1), 5
    yield nb.nancount, _slow_count, np.inf


def _arrays(func_name, dtypes=(np.float32, np.float64)):
    """The reference's test arrays (numbagg/test/util.py:26-120)."""
    nan, inf = np.nan, np.inf
    out = [
        np.array([inf, nan]), np.array([inf, -inf]), np.array([nan, 2, 3]), np.array([-inf, 2, 3]),
    ]
    if func_name != "nanargmin":
        out.append(np.array([nan, inf]))
    out += [
        np.array([1, 2, 3], dtype=">f4"), np.array([1, 2, 3], dtype="<f4"), np.array([1, 2, 3], dtype=np.float16),
        np.array([1, 2, 3], dtype=np.float64) + 1e9, np.array([0, 0, 0], dtype=np.float64),
        np.array([1, nan, nan, 2]), np.array([2**31], dtype=np.int64),
        np.array([[1.0, 2], [3, 4]])[..., np.newaxis], np.array([1, 1, 1], dtype=np.float64),
    ]
    for v in (-9, 0, 9, -9.0, 0.0, 9.0, -inf, inf, nan):
        out.append(np.array(v, dtype=np.float64))
    ss = {
        0: (0, [(0,), (0, 0), (2, 0), (2, 0, 1)]), 1: (8, [(8,)]), 2: (12, [(2, 6), (3, 4)]),
        3: (16, [(2, 2, 4)]), 4: (24, [(1, 2, 3, 4)]), 5: (1000, [(1000,)]), 6: (10_000, [(100, 100)]),
        7: (1_000_000, [(1000, 1000)]), 8: (1_000_000, [(100, 10_000)]),
    }
    for seed in (1, 2):
        rs = np.random.RandomState(seed)
        for size, shapes in ss.values():
            for dtype in dtypes:
                a = np.arange(size, dtype=dtype)
                if func_name not in ("nanargmin", "nanargmax"):
                    a[rs.rand(*a.shape) < 0.2] = inf
                a[rs.rand(*a.shape) < 0.2] = nan
                idx = rs.rand(*a.shape) < 0.2
                a[idx] *= -1
                rs.shuffle(a)
                for shape in shapes:
                    out.append(a.reshape(shape))
    out.append(np.array([[1, 2], [3, 4]], dtype=np.int64)[:, [1]])
    for dtype in dtypes:
        a = np.arange(12).astype(dtype)
        for start in range(3):
            for step in range(1, 3):
                out.append(a[start::step])
    res = []
    for a in out:
        res.append(a)
        if a.ndim >= 2:
            res.append(a.T)
    return res


@pytest.mark.filterwarnings("ignore")
@pytest.mark.parametrize("k", range(10))
def test_numerical_results_identical(gpu_lib, k):
    numbagg_func, comp_func, decimal = list(_functions())[k]
    for arr in _arrays(numbagg_func.__name__):
        for axis in list(range(-arr.ndim, arr.ndim)) + [None]:
            with np.errstate(invalid="ignore", divide="ignore"):
                desired_arr = arr.copy()
                if desired_arr.dtype == np.float16:
                    desired_arr = desired_arr.astype(np.float32)
                try:
                    desired, desiredraised = comp_func(desired_arr, axis=axis), False
                except Exception as err:
                    desired, desiredraised = str(err), True
                try:
                    actual, actualraised = numbagg_func(arr.copy(), axis=axis), False
                except Exception as err:
                    if not desiredraised:
                        raise
                    actual, actualraised = str(err), True
            tag = f"{numbagg_func.__name__} {arr.dtype} {arr.shape} axis={axis}"
            if actualraised and desiredraised:
                if actual == "All-NaN slice encountered" and "empty sequence" in desired:
                    continue
                assert desired == actual, tag
            elif desiredraised and hasattr(actual, "size") and actual.size == 0:
                continue
            else:
                assert actualraised == desiredraised, tag
                actual, desired = np.asarray(actual), np.asarray(desired)
                if numbagg_func.__name__ == "nansum" and actual.dtype.kind == "f":
                    # float sums: the reference compares with NumPy's pairwise sum;
                    # here the sum is held to the exactly rounded one (math.fsum)
                    want, sabs = _fsum(desired_arr, axis)
                    _check_sum(actual, want, sabs, tag)
                elif decimal < np.inf and np.isfinite(arr).sum() > 0:
                    assert_array_almost_equal(actual, desired, decimal, tag)
                else:
                    assert_array_equal(actual, desired, tag)
                assert_equal(actual.dtype, desired.dtype, tag)


@pytest.mark.parametrize("func", ["nansum", "nanmean", "nancount", "allnan", "anynan", "nanvar", "nanstd",
                                  "nanmax", "nanmin", "nanargmax", "nanargmin"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32, np.int64])
@pytest.mark.parametrize("shape", [(7, 1), (3, 5000), (2, 300_001), (1, 2_000_003), (5000, 33)])
def test_against_restatement(gpu_lib, func, dtype, shape):
    """Bit-exact where the reference's result is exact (counts, flags,
    extremes and their indices, integer sums); float sums, means and
    variances within their bars of the float64 oracle."""
    import numbagg_amd as nb
    from oracle import exact, restated

    if dtype in (np.int32, np.int64) and func in ("nanmean", "nanvar", "nanstd"):
        pytest.skip("integer input runs the float64 loop (covered by the dtype test)")
    rng = np.random.default_rng(shape[1] + len(func))
    if np.issubdtype(dtype, np.floating):
        x = (rng.standard_normal(shape) * 10).astype(dtype)
        x[rng.random(shape) < 0.1] = np.nan
        x[0, : min(3, shape[1])] = np.nan
        if shape[0] > 2:
            x[2, :] = np.nan  # an all-NaN row
        x[rng.random(shape) < 0.01] = 0.0
    else:
        x = rng.integers(-(2**20), 2**20, size=shape).astype(dtype)
        x[0, : min(7, shape[1])] = x.max()  # ties
    if func in ("nanargmax", "nanargmin") and np.isnan(x).all(axis=-1).any():
        with pytest.raises(ValueError, match="All-NaN slice encountered"):
            getattr(nb, func)(x, axis=-1)
        x = x[~np.isnan(x).all(axis=-1)]
    got = getattr(nb, func)(x, axis=-1)
    want = restated.reduce(func, x)
    if func in ("nansum", "nanmean", "nanvar", "nanstd") and np.issubdtype(dtype, np.floating):
        ex = exact.reduce(func, x)
        assert np.array_equal(np.isnan(got), np.isnan(ex))
        ok = ~np.isnan(ex)
        if func == "nansum":
            scale = np.nansum(np.abs(x.astype(np.float64)), axis=-1)[ok]
            ulp = np.spacing(np.abs(ex[ok]).astype(dtype)).astype(np.float64)
            assert np.all(np.abs(got[ok] - ex[ok]) <= ulp + 1e-14 * scale)
        else:
            assert_allclose(got[ok], ex[ok], rtol=2e-7 if dtype == np.float32 else 1e-12)
    else:
        assert got.dtype == want.dtype
        assert_array_equal(got.view(np.uint8), want.view(np.uint8))


def test_multi_axis_and_layouts(gpu_lib):
    """axis tuples in any order, negative axes, non-contiguous and
    transposed inputs, torch tensors in and out (numbagg/utils.py:20-32)."""
    import torch

    import numbagg_amd as nb

    rng = np.random.default_rng(5)
    x = rng.standard_normal((4, 6, 50, 7))
    x[rng.random(x.shape) < 0.2] = np.nan
    for axis in [(1, 2), (2, 1), (-1, 0), (0, 1, 2, 3), (3,), 2, -3, None]:
        assert_allclose(nb.nanmean(x, axis=axis), np.nanmean(x, axis=axis), rtol=1e-12)
        assert_allclose(nb.nanvar(x, axis=axis), np.nanvar(x, axis=axis, ddof=1), rtol=1e-11)
        assert_array_equal(nb.nancount(x, axis=axis), _slow_count(x, axis=axis))
    assert_allclose(nb.nanstd(x, axis=(0, 2), ddof=0), np.nanstd(x, axis=(0, 2)), rtol=1e-11)
    xt = np.transpose(x, (2, 0, 3, 1))[:, ::2]
    want, sabs = _fsum(xt, (1, 3))
    _check_sum(nb.nansum(xt, axis=(1, 3)), want, sabs, "nansum transposed view")
    t = torch.from_numpy(x).cuda()
    r = nb.nanmean(t, axis=-1)
    assert isinstance(r, torch.Tensor) and r.is_cuda
    assert_allclose(r.cpu().numpy(), np.nanmean(x, axis=-1), rtol=1e-12)
    # 0-d results are NumPy scalars, as from a gufunc
    s = nb.nansum(x)
    assert np.isscalar(s) and isinstance(s, np.float64)
    assert nb.allnan(np.array([np.nan, np.nan])) and not nb.anynan(np.array([1.0, 2.0]))
    # the reference's allnan/anynan cases (test_funcs.py:261-327)
    assert_array_equal(nb.allnan(np.array([[np.nan, np.nan], [1.0, np.nan], [np.nan, np.nan]]), axis=1),
                       [True, False, True])
    assert_array_equal(nb.anynan(np.array([[1.0, 2.0], [1.0, np.nan], [3.0, 4.0]]), axis=1), [False, True, False])
    assert nb.allnan(np.zeros((3, 0)), axis=1).tolist() == [True] * 3
    assert nb.anynan(np.zeros((3, 0)), axis=1).tolist() == [False] * 3
    with pytest.raises(TypeError, match="must be arrays"):
        nb.nansum([1.0, 2.0])
    assert nb.count is nb.nancount


def test_dtype_selection(gpu_lib):
    """NumPy's loop choice over numbagg's signatures: bool/int8 -> int32 for
    nansum, float32 for nanmean (safe casts); int32 nanmax -> int64; float16
    -> float32; uint32 -> int64."""
    import numbagg_amd as nb

    b = np.array([True, False, True])
    assert nb.nansum(b).dtype == np.int32 and nb.nansum(b) == 2
    assert nb.nanmean(b).dtype == np.float32
    assert nb.nanmean(np.array([1, 2], np.int32)).dtype == np.float64
    assert nb.nanmax(np.array([1, 5], np.int32)).dtype == np.int64
    assert nb.nanmax(np.array([1, 5], np.float16)).dtype == np.float32
    assert nb.nansum(np.array([1, 2], np.uint32)).dtype == np.int64
    assert nb.nansum(np.array([2**31 - 1, 1], np.int32)) == np.int32(-(2**31))  # wraps like the reference
    assert nb.nancount(np.array([1.0, np.nan])).dtype == np.int64


def test_repeatable_bits(gpu_lib):
    """No atomics: two runs give the same bits."""
    import numbagg_amd as nb

    rng = np.random.default_rng(9)
    x = rng.standard_normal((3, 1_000_003)).astype(np.float32)
    for f in (nb.nansum, nb.nanmean, nb.nanvar):
        a, b = f(x, axis=-1), f(x, axis=-1)
        assert_array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("func", ["nanmax", "nanmin", "nanargmax", "nanargmin"])
def test_extreme_ties_signed_zeros_and_infinities(gpu_lib, func, dtype):
    """Repeated extremes spread over many chunks and lanes (the first index
    wins, as the reference's strict comparison keeps it), -0.0 beside +0.0
    and infinities: bit-exact against the C restatement."""
    import numbagg_amd as nb
    from oracle import restated

    rng = np.random.default_rng(7)
    n = 2_000_003
    x = rng.integers(-3, 4, size=(4, n)).astype(dtype)  # many ties of every value
    x[rng.random(x.shape) < 0.05] = np.nan
    x[1] = np.where(x[1] == 0, -0.0, x[1])  # signed zeros: row 1 all -0.0 where 0
    x[2, rng.integers(0, n, 50)] = np.inf
    x[2, rng.integers(0, n, 50)] = -np.inf
    x[3] = rng.choice(np.array([0.0, -0.0, np.nan], dtype=dtype), size=n)  # zeros of both signs only
    got = getattr(nb, func)(x, axis=-1)
    want = restated.reduce(func, x)
    assert got.dtype == want.dtype
    assert_array_equal(got.view(np.uint8), want.view(np.uint8))
