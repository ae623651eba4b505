"""GPU parity of the correlation / covariance matrices (numbagg/funcs.py:338-532,
numbagg/moving_matrix.py:16-432) against the exact oracle (pinned to pandas in
tests/test_oracle_matrix.py), pandas' golden vectors, and the known answers of
numbagg/test/test_matrix_functions.py.

Bars (e = exact value, s = sqrt(|e_ii e_jj|) the scale of a pair, read off
the exact matrix's diagonal; NaN masks must be identical):
  float32 outputs: |gpu - e| <= 1 ULP(float32 e) + 1e-12 s (corr: + 1e-12);
  float64 outputs: |gpu - e| <= 1e-12 s for static matrices, 1e-11 s for the
  moving and exponential ones (running float64 window sums / decayed state);
  corr: the same with s = 1. Windows with fewer than 10 pair observations get
  1e-8 s: a statistic formed from sums (Q - S^2 / n) loses digits in
  proportion to Q / (Q - S^2 / n), which is unbounded for a few close values
  (two observations make the correlation exactly +-1 and the sums leave a
  rounding residue of about 1e-11 there).
Every matrix is exactly symmetric and a finite correlation diagonal is exactly
1. Results are deterministic (no atomics): repeated calls are bitwise equal.
"""

from __future__ import annotations

import numpy as np
import pytest
from numpy.testing import assert_allclose

pytestmark = pytest.mark.gpu


def _scale(e, corr):
    if corr:
        return np.ones_like(e)
    d = np.diagonal(e, axis1=-2, axis2=-1)
    return np.sqrt(np.abs(d[..., :, None] * d[..., None, :]))


def _pair_counts(x, window=None):
    """Pair-complete observations per output element of the moving (window) or
    exponential (window None: all so far) matrices of x (batch, obs, vars)."""
    v = (~np.isnan(x)).astype(np.int64)
    pairs = v[..., :, None] * v[..., None, :]
    c = np.cumsum(pairs, axis=1)
    if window is not None:
        c[:, window:] = c[:, window:] - c[:, :-window]
    return c


def _check(got, exact, corr, rel64, tag, counts=None):
    got = np.asarray(got)
    assert got.shape == exact.shape, tag
    gnan, enan = np.isnan(got), np.isnan(exact)
    assert np.array_equal(gnan, enan), f"{tag}: NaN masks differ at {np.argwhere(gnan != enan)[:5].tolist()}"
    ok = ~enan
    g = got[ok].astype(np.float64)
    e = exact[ok]
    s = _scale(exact, corr)[ok]
    if got.dtype == np.float32:
        bar = np.spacing(np.abs(e).astype(np.float32)).astype(np.float64) + 1e-12 * np.where(np.isfinite(s), s, 0)
    else:
        rel = np.full(e.shape, rel64)
        if counts is not None:
            rel = np.where(counts[ok] < 10, 1e-8, rel)
        bar = rel * np.where(np.isfinite(s), s, 0) + 1e-300
    err = np.abs(g - e)
    worst = np.argmax(err - bar) if err.size else 0
    assert np.all(err <= bar), f"{tag}: error {err[worst]:.3g} > bar {bar[worst]:.3g}"
    # exact symmetry and unit correlation diagonal
    assert np.array_equal(got, np.swapaxes(got, -1, -2), equal_nan=True), tag
    if corr:
        d = np.diagonal(got, axis1=-2, axis2=-1)
        assert np.all(d[~np.isnan(d)] == 1.0), tag


def _static_data(batch, nv, nobs, dtype, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((batch, nv, nobs)) * rng.uniform(0.5, 3.0, (batch, nv, 1)) + 100.0
    x[rng.random(x.shape) < 0.1] = np.nan
    if nv >= 4:
        x[:, 1] = 7.0  # a constant variable
        x[:, 2] = np.nan  # an all-NaN variable
        x[:, 3, 1:] = np.nan  # a single observation
    return x.astype(dtype)


STATIC_SHAPES = [
    (1, 3, 500),  # one 16 x 16 tile, vector loads
    (2, 5, 1003),  # scalar loads, ragged tail
    (1, 20, 1000),  # 32 x 32 wave tile
    (1, 70, 3001),  # 3 x 3 tiles, off-diagonal tiles
    (3, 17, 50_000),  # observation splits, finishing kernel
    (1, 200, 20_000),  # many tiles and splits
    (600, 5, 700),  # few variables over a large batch: vector-unit kernel, staged chunks
    (520, 8, 1024),  # the most variables that kernel takes, whole chunks
    (512, 1, 40),  # one variable
    (513, 3, 513),  # whole chunks and one observation
]


@pytest.mark.parametrize("shape", STATIC_SHAPES)
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("func", ["nancorrmatrix", "nancovmatrix"])
def test_static_vs_exact(gpu_lib, func, dtype, shape):
    import numbagg_amd as nb
    from oracle import exact

    x = _static_data(*shape, dtype, seed=sum(shape))
    got = getattr(nb, func)(x)
    assert got.dtype == dtype
    _check(got, exact.matrix(func, x), func == "nancorrmatrix", 1e-12, f"{func} {shape} {dtype.__name__}")
    again = getattr(nb, func)(x)
    assert np.array_equal(got.view(np.uint8), again.view(np.uint8))


@pytest.mark.parametrize("tag", ["3x500", "20x1000"])
@pytest.mark.parametrize("func", ["nancorrmatrix", "nancovmatrix"])
def test_static_pandas_goldens(gpu_lib, golden, func, tag):
    import numbagg_amd as nb

    got = getattr(nb, func)(golden[f"matrix_array_{tag}"])
    want = golden[f"{func}_{tag}"]
    assert np.array_equal(np.isnan(got), np.isnan(want))
    # pandas' own rounding: absolute in the data's scale (variances ~ 1/12)
    assert_allclose(got, want, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("func", ["nancorrmatrix", "nancovmatrix"])
def test_static_known_answers(gpu_lib, func):
    """numbagg/test/test_matrix_functions.py:23-190, :730-759."""
    import numbagg_amd as nb

    f = getattr(nb, func)
    corr = func == "nancorrmatrix"
    data = np.array([[1, 2, 3, 4], [2, 4, 6, 8]], dtype=np.float64)
    r = f(data)
    assert r.shape == (2, 2)
    assert_allclose(r, r.T)
    if corr:
        assert_allclose(r, [[1.0, 1.0], [1.0, 1.0]])
    else:
        assert np.all(np.diag(r) >= 0)
    nans = np.array([[1, 2, np.nan, 4], [2, 4, 6, np.nan], [np.nan, 1, 2, 3]], dtype=np.float64)
    r = f(nans)
    assert r.shape == (3, 3)
    assert_allclose(r, r.T, equal_nan=True)
    if corr:
        assert_allclose(np.diag(r), [1.0, 1.0, 1.0])
    with pytest.raises(ValueError, match="requires at least a 2D array"):
        f(np.array([1, 2, 3, 4, 5], dtype=np.float64))
    rng = np.random.RandomState(42)
    data = rng.randn(5, 100)
    assert_allclose(f(data), np.corrcoef(data) if corr else np.cov(data), rtol=1e-10)
    const = np.array([[1, 1, 1, 1], [2, 2, 2, 2], [1, 2, 3, 4]], dtype=np.float64)
    r = f(const)
    if corr:
        assert np.isnan(r[0, 1]) and np.isnan(r[0, 2]) and r[2, 2] == 1.0
    else:
        assert r[0, 0] == 0.0 and r[1, 1] == 0.0 and r[0, 1] == 0.0
    d3 = rng.randn(2, 4, 10)
    d4 = rng.randn(2, 3, 4, 10)
    r3, r4 = f(d3), f(d4)
    assert r3.shape == (2, 4, 4) and r4.shape == (2, 3, 4, 4)
    for i in range(2):
        assert_allclose(r3[i], f(d3[i]), rtol=1e-10)
    if corr:
        assert np.all((r4 >= -1) & (r4 <= 1))
    with pytest.raises(TypeError):
        f(data, axis=0)
    assert f(rng.randn(3, 10).astype(np.float32)).dtype == np.float32
    assert f(rng.randn(3, 10)).dtype == np.float64
    assert f(np.arange(12).reshape(3, 4)).dtype == np.float64  # int64 -> the float64 loop
    empty = f(np.zeros((3, 0)))
    assert empty.shape == (3, 3) and np.all(np.isnan(empty))
    assert f(np.zeros((0, 5))).shape == (0, 0)


def test_static_cov_corr_relationship(gpu_lib):
    import numbagg_amd as nb

    data = np.random.RandomState(42).randn(4, 50)
    cov = nb.nancovmatrix(data)
    s = np.sqrt(np.diag(cov))
    assert_allclose(nb.nancorrmatrix(data), cov / np.outer(s, s), rtol=1e-10)


def _walk(batch, nobs, nv, dtype, seed, offset=1e4, p_nan=0.1):
    rng = np.random.default_rng(seed)
    x = np.cumsum(rng.standard_normal((batch, nobs, nv)), axis=1) * 0.05 + rng.standard_normal((batch, nobs, nv))
    x = x + offset
    x[rng.random(x.shape) < p_nan] = np.nan
    return x.astype(dtype)


MOVE_CASES = [
    ((1, 500, 3), 20, None),
    ((1, 500, 3), 20, 3),
    ((2, 3000, 2), 50, 0),  # four pairs: 16 time ranges per wave
    ((1, 2000, 9), 100, 1),  # 81 pairs: two waves per tile
    ((1, 5000, 16), 300, 150),
    ((2, 400, 5), 400, 10),  # window = the whole series
]


@pytest.mark.parametrize("case", MOVE_CASES)
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("func", ["move_corrmatrix", "move_covmatrix"])
def test_moving_vs_exact(gpu_lib, func, dtype, case):
    import numbagg_amd as nb
    from oracle import exact

    shape, window, mc = case
    x = _walk(*shape, dtype, seed=shape[1] + window)
    got = getattr(nb, func)(x, window=window, min_count=mc)
    assert got.dtype == dtype
    want = exact.matrix(func, x, window=window, min_count=mc)
    _check(got, want, func == "move_corrmatrix", 1e-11, f"{func} {case} {dtype.__name__}", _pair_counts(x, window))


@pytest.mark.parametrize("window,mc", [(20, None), (20, 3), (50, 1)])
@pytest.mark.parametrize("func", ["move_corrmatrix", "move_covmatrix"])
def test_moving_pandas_goldens(gpu_lib, golden, func, window, mc):
    import numbagg_amd as nb

    x = np.ascontiguousarray(golden["matrix_array_3x500"].T)
    got = getattr(nb, func)(x, window=window, min_count=mc)
    want = golden[f"{func}_w{window}_mc{mc}"]
    assert np.array_equal(np.isnan(got), np.isnan(want))
    assert_allclose(got, want, rtol=1e-10, atol=1e-13)


@pytest.mark.parametrize("func", ["move_corrmatrix", "move_covmatrix"])
def test_moving_known_answers(gpu_lib, func):
    """numbagg/test/test_matrix_functions.py:195-434, :730-759."""
    import numbagg_amd as nb

    f = getattr(nb, func)
    corr = func == "move_corrmatrix"
    data = np.array([[1, 2], [2, 4], [3, 6], [4, 8], [5, 10], [6, 12]], dtype=np.float64)
    r = f(data, window=3, min_count=2)
    assert r.shape == (6, 2, 2)
    for t in range(6):
        assert_allclose(r[t], r[t].T, equal_nan=True)
    if corr:
        for t in range(1, 6):
            assert_allclose(r[t], [[1.0, 1.0], [1.0, 1.0]], rtol=1e-10)
    with pytest.raises(ValueError, match="requires at least a 2D array"):
        f(np.array([1, 2, 3, 4, 5], dtype=np.float64), window=3)
    zero = np.array([[1, 2], [1, 2], [1, 2], [2, 3], [3, 4], [4, 5]], dtype=np.float64)
    r = f(zero, window=3, min_count=2)
    if corr:
        assert np.isnan(r[2, 0, 1])
    else:
        assert r[2, 0, 0] == 0.0 and r[2, 1, 1] == 0.0 and r[2, 0, 1] == 0.0
    assert not np.all(np.isnan(r[5]))
    rng = np.random.RandomState(42)
    d3, d4 = rng.randn(2, 20, 4), rng.randn(2, 3, 20, 4)
    r3 = f(d3, window=5)
    assert r3.shape == (2, 20, 4, 4) and f(d4, window=5).shape == (2, 3, 20, 4, 4)
    for i in range(2):
        assert_allclose(r3[i], f(d3[i], window=5), rtol=1e-10)
    data_moving = rng.randn(50, 4)
    if corr:
        # the reference asks rtol 1e-14 of one code path against itself; the two
        # kernels here sum in different orders: 1e-15 absolute for correlations
        assert_allclose(f(data_moving, window=50)[-1], nb.nancorrmatrix(data_moving.T), rtol=1e-14, atol=1e-15)
    mism = np.array([[1.0, 2.0], [2.0, 4.0], [3.0, np.nan], [np.nan, 8.0], [5.0, 10.0], [6.0, 12.0]])
    m = f(mism, window=6, min_count=2)[5]
    if corr:
        fin = m[np.isfinite(m)]
        assert np.all((fin >= -1.0) & (fin <= 1.0))
        assert_allclose(m[0, 1], 1.0, rtol=1e-10)
    else:
        assert m[0, 1] > 0
    with pytest.raises(ValueError, match="window not in valid range"):
        f(data, window=7)
    with pytest.raises(ValueError, match="min_count must be positive"):
        f(data, window=3, min_count=-1)
    with pytest.raises(TypeError):
        f(data, window=3.0)
    assert f(rng.randn(10, 3).astype(np.float32), window=5, min_count=3).dtype == np.float32
    assert f(rng.randn(10, 3), window=5, min_count=3).dtype == np.float64


EXP_CASES = [
    ((1, 500, 3), "scalar", 0.0),
    ((2, 3000, 2), "row", 0.0),
    ((2, 1000, 9), "full", 0.0),
    ((1, 4000, 16), "scalar", 0.5),
    ((3, 200, 5), "row", 0.9),
]


def _alpha(mode, batch, nobs, seed):
    rng = np.random.default_rng(seed)
    if mode == "scalar":
        return 0.1
    if mode == "row":
        return rng.uniform(0.01, 0.6, nobs)
    return rng.uniform(0.01, 0.6, (batch, nobs))


@pytest.mark.parametrize("case", EXP_CASES)
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("func", ["move_exp_nancorrmatrix", "move_exp_nancovmatrix"])
def test_exp_vs_exact(gpu_lib, func, dtype, case):
    import numbagg_amd as nb
    from oracle import exact

    shape, mode, mw = case
    x = _walk(*shape, dtype, seed=shape[1], offset=100.0)
    alpha = _alpha(mode, shape[0], shape[1], seed=shape[1])
    if mode != "scalar" and dtype == np.float32:
        alpha = alpha.astype(np.float32)  # float32 alpha array: the float32 loop
    got = getattr(nb, func)(x, alpha=alpha, min_weight=mw)
    loop = np.float64 if (mode == "scalar" or dtype == np.float64) else np.float32
    assert got.dtype == loop
    want = exact.matrix(func, x, alpha=alpha, min_weight=mw, loop_dtype=loop)
    _check(got, want, func == "move_exp_nancorrmatrix", 1e-11, f"{func} {case} {dtype.__name__}", _pair_counts(x))


@pytest.mark.parametrize("alpha", [0.5, 0.1])
@pytest.mark.parametrize("func", ["move_exp_nancorrmatrix", "move_exp_nancovmatrix"])
def test_exp_pandas_goldens(gpu_lib, golden, func, alpha):
    import numbagg_amd as nb

    x = np.ascontiguousarray(golden["matrix_array_3x500"].T)
    got = getattr(nb, func)(x, alpha=alpha)
    want = golden[f"{func}_a{alpha}"]
    assert np.array_equal(np.isnan(got), np.isnan(want))
    assert_allclose(got, want, rtol=1e-10, atol=1e-13)


@pytest.mark.parametrize("func", ["move_exp_nancorrmatrix", "move_exp_nancovmatrix"])
def test_exp_known_answers(gpu_lib, func):
    """numbagg/test/test_matrix_functions.py:436-728."""
    import numbagg_amd as nb

    f = getattr(nb, func)
    corr = func == "move_exp_nancorrmatrix"
    data = np.array([[1, 2], [2, 4], [3, 6], [4, 8]], dtype=np.float64)
    r = f(data, alpha=0.5)
    assert r.shape == (4, 2, 2)
    for t in range(4):
        assert_allclose(r[t], r[t].T, rtol=1e-10)
    if corr:
        assert_allclose(r[-1], [[1.0, 1.0], [1.0, 1.0]], rtol=1e-10)
    else:
        assert np.all(np.diag(r[-1]) >= 0)
    nans = np.array([[1, 2, np.nan], [2, 4, 1], [np.nan, 6, 2], [4, np.nan, 3]], dtype=np.float64)
    r = f(nans, alpha=0.3)
    assert r.shape == (4, 3, 3) and np.any(np.isfinite(r))
    hi, lo = f(data, alpha=0.1, min_weight=0.8), f(data, alpha=0.1, min_weight=0.1)
    assert np.sum(np.isnan(hi[0])) >= np.sum(np.isnan(lo[0]))
    sq = np.array([[1, 1], [2, 4], [3, 9], [4, 16], [5, 25]], dtype=np.float64)
    assert not np.allclose(f(sq, alpha=0.9)[-1], f(sq, alpha=0.1)[-1], rtol=1e-3)
    arr = f(data, alpha=np.array([0.1, 0.5, 0.9, 0.3]))
    assert arr.shape == (4, 2, 2) and not np.allclose(arr, f(data, alpha=0.5))
    rng = np.random.RandomState(42)
    d3, d4 = rng.randn(2, 20, 4), rng.randn(2, 3, 15, 4)
    r3 = f(d3, alpha=0.3)
    assert r3.shape == (2, 20, 4, 4) and f(d4, alpha=0.3).shape == (2, 3, 15, 4, 4)
    for i in range(2):
        assert_allclose(r3[i], f(d3[i], alpha=0.3), rtol=1e-10, equal_nan=True)
    with pytest.raises(ValueError, match="requires at least a 2D array"):
        f(np.array([1.0, 2.0, 3.0]), alpha=0.5)
    psd = f(rng.randn(50, 4), alpha=0.3)
    for t in range(50):
        m = psd[t]
        if not np.any(np.isnan(m)):
            assert np.all(np.linalg.eigvals(m) >= -1e-10)
            if corr:
                assert_allclose(np.diag(m), 1.0, rtol=1e-12)
                assert np.all((m >= -1 - 1e-10) & (m <= 1 + 1e-10))


@pytest.mark.parametrize("alpha", [0.1, 0.5, 0.9])
def test_exp_matrix_matches_pairwise_functions(gpu_lib, alpha):
    """numbagg/test/test_matrix_functions.py:620-706."""
    import numbagg_amd as nb

    rng = np.random.RandomState(42)
    a1, a2 = rng.randn(50), rng.randn(50) * 2 + 1
    a3 = rng.randn(50) * 0.8 - 0.2
    m = np.column_stack([a1, a2, a3])
    cov, corr = nb.move_exp_nancovmatrix(m, alpha=alpha), nb.move_exp_nancorrmatrix(m, alpha=alpha)
    for s1, s2, i, j in ((a1, a2, 0, 1), (a1, a3, 0, 2), (a2, a3, 1, 2)):
        assert_allclose(nb.move_exp_nancov(s1, s2, alpha=alpha), cov[:, i, j], rtol=1e-10)
        assert_allclose(nb.move_exp_nancorr(s1, s2, alpha=alpha), corr[:, i, j], rtol=1e-10)
        assert_allclose(cov[:, i, j], cov[:, j, i], rtol=1e-10)


def test_matrices_stay_on_device(gpu_lib):
    import torch

    import numbagg_amd as nb

    x = torch.randn(3, 4, 1000, device="cuda", dtype=torch.float64)
    r = nb.nancorrmatrix(x)
    assert isinstance(r, torch.Tensor) and r.is_cuda and r.shape == (3, 4, 4)
    assert_allclose(r.cpu().numpy(), nb.nancorrmatrix(x.cpu().numpy()), rtol=0, atol=0)


def test_matrix_edge_shapes(gpu_lib):
    """One variable, one observation, all-NaN variables, empty batches, ragged
    observation counts and float64 data at a large level."""
    import numbagg_amd as nb
    from oracle import exact

    rng = np.random.default_rng(9)
    for shape in ((1, 1, 50), (2, 3, 1), (1, 17, 4099), (3, 2, 5)):
        x = rng.standard_normal(shape)
        x[..., ::7] = np.nan
        if shape[1] > 2:
            x[:, 1] = np.nan
        for f in ("nancorrmatrix", "nancovmatrix"):
            _check(getattr(nb, f)(x), exact.matrix(f, x), f == "nancorrmatrix", 1e-12, f"{f} {shape}")
    assert nb.nancovmatrix(np.zeros((0, 3, 10))).shape == (0, 3, 3)
    big = rng.standard_normal((1, 2, 3000)) * 1e3 + 1e12
    _check(nb.nancovmatrix(big), exact.matrix("nancovmatrix", big), False, 1e-12, "large level")
    for shape, window in (((1, 1, 1), 1), ((1, 7, 1), 3), ((2, 5, 3), 5)):
        x = rng.standard_normal(shape)
        x[:, ::2, 0] = np.nan
        for f in ("move_corrmatrix", "move_covmatrix"):
            got = getattr(nb, f)(x, window=window, min_count=1)
            _check(got, exact.matrix(f, x, window=window, min_count=1), f == "move_corrmatrix", 1e-11,
                   f"{f} {shape}", _pair_counts(x, window))
        for f in ("move_exp_nancorrmatrix", "move_exp_nancovmatrix"):
            got = getattr(nb, f)(x, alpha=0.4)
            _check(got, exact.matrix(f, x, alpha=0.4), f == "move_exp_nancorrmatrix", 1e-11, f"{f} {shape}",
                   _pair_counts(x))
    assert nb.move_corrmatrix(np.zeros((0, 4, 2)), window=2).shape == (0, 4, 2, 2)
