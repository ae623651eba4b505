"""GPU parity of ffill/bfill: bit-exact against the oracle (closed form and the
reference restatement), through the public API and the C ABI."""

from __future__ import annotations

import numpy as np
import pytest

from ._data import bits as _bits
from ._data import runs_array

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("func", ["ffill", "bfill"])
@pytest.mark.parametrize("limit", [None, 0, 1, 16, 5000])
@pytest.mark.parametrize("shape", [(3, 500), (7, 4097), (2, 50_001), (1, 200_000), (300, 333), (12, 30_001)])
def test_fill_bit_exact(gpu_lib, dtype, func, limit, shape):
    import numbagg_amd as nb
    from oracle import exact, restated

    x = runs_array(*shape, seed=shape[1], dtype=dtype)
    got = getattr(nb, func)(x, limit=limit)
    assert got.dtype == x.dtype and got.shape == x.shape
    np.testing.assert_array_equal(_bits(got), _bits(exact.fill(func, x, limit)))
    np.testing.assert_array_equal(_bits(got), _bits(restated.fill(func, x, limit)))


@pytest.mark.parametrize("pack", ["1", "0"])
@pytest.mark.parametrize("fill_bytes", ["16", "32", "64", "128"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("func", ["ffill", "bfill"])
@pytest.mark.parametrize("limit", [None, 3, 700])
def test_fill_tile_sizes_bit_exact(gpu_lib, monkeypatch, pack, fill_bytes, dtype, func, limit):
    """Every tile size (bytes per thread), with the carry value packed in the
    look-back status word or loaded from the input, gives the same bits."""
    import numbagg_amd as nb
    from oracle import exact

    monkeypatch.setenv("NBG_FILL_TWO_PASS", "0")  # the ordered single pass's knobs
    monkeypatch.setenv("NBG_FILL_BYTES", fill_bytes)
    monkeypatch.setenv("NBG_FILL_PACK", pack)
    for shape in [(5, 1023), (3, 40_961), (64, 3000)]:
        x = runs_array(*shape, seed=shape[1] + 7, dtype=dtype)
        got = getattr(nb, func)(x, limit=limit)
        np.testing.assert_array_equal(_bits(got), _bits(exact.fill(func, x, limit)))


@pytest.mark.parametrize("span,prefetch", [("2", "0"), ("3", "1"), ("8", "0"), ("16", "1")])
@pytest.mark.parametrize("pack", ["1", "0"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("func", ["ffill", "bfill"])
@pytest.mark.parametrize("limit", [None, 5, 9000])
def test_fill_spans_bit_exact(gpu_lib, monkeypatch, span, prefetch, pack, dtype, func, limit):
    """Span workgroups (several tiles per ticket, the leading NaN run written
    after one look-back) give the same bits, including rows whose spans hold
    no valid value at all."""
    import numbagg_amd as nb
    from oracle import exact

    monkeypatch.setenv("NBG_FILL_TWO_PASS", "0")  # the ordered single pass's knobs
    monkeypatch.setenv("NBG_FILL_SPAN", span)
    monkeypatch.setenv("NBG_FILL_PREFETCH", prefetch)
    monkeypatch.setenv("NBG_FILL_PACK", pack)
    for shape in [(5, 1023), (3, 40_961), (16, 70_000)]:
        x = runs_array(*shape, seed=shape[1] + int(span), dtype=dtype)
        x[1, 100:30_000] = np.nan  # NaN stretches longer than whole spans
        x[2, :] = np.nan
        x[2, shape[1] // 2] = 1.0
        got = getattr(nb, func)(x, limit=limit)
        np.testing.assert_array_equal(_bits(got), _bits(exact.fill(func, x, limit)))


@pytest.mark.parametrize("two_pass", ["1", "0"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("func", ["ffill", "bfill"])
@pytest.mark.parametrize("limit", [None, 0, 3, 700, 5000])
def test_fill_two_pass_bit_exact(gpu_lib, monkeypatch, two_pass, dtype, func, limit):
    """The two-pass fill (independent tiles, then leading NaN runs filled from
    the tile summaries) and the ordered single pass give the same bits,
    including NaN stretches over many tiles, all-NaN rows, unaligned rows and
    partial last tiles."""
    import numbagg_amd as nb
    from oracle import exact

    monkeypatch.setenv("NBG_FILL_TWO_PASS", two_pass)
    for shape in [(5, 1023), (3, 40_961), (16, 70_000), (600, 3000)]:
        x = runs_array(*shape, seed=shape[1] + 11, dtype=dtype)
        rows = shape[0]
        x[1, 100 : min(30_000, shape[1] - 1)] = np.nan
        x[2, :] = np.nan
        x[rows - 1, :] = np.nan
        x[rows - 1, shape[1] // 2] = 1.0
        if rows > 4:
            x[3, : shape[1] - 2] = np.nan
            x[4, 2:] = np.nan
        got = getattr(nb, func)(x, limit=limit)
        np.testing.assert_array_equal(_bits(got), _bits(exact.fill(func, x, limit)))


@pytest.mark.parametrize("run", ["2", "5", "16", "100000"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("func", ["ffill", "bfill"])
@pytest.mark.parametrize("limit", [None, 7])
def test_fill_dispatch_runs_bit_exact(gpu_lib, monkeypatch, run, dtype, func, limit):
    """Dispatching runs of consecutive tiles of one row (up to row-major order)
    gives the same bits, with tile counts that are and are not multiples of
    the run."""
    import numbagg_amd as nb
    from oracle import exact

    monkeypatch.setenv("NBG_FILL_TWO_PASS", "0")  # the ordered single pass's knobs
    monkeypatch.setenv("NBG_FILL_RUN", run)
    for shape in [(5, 1023), (3, 40_961), (16, 70_000), (2, 65_536)]:
        x = runs_array(*shape, seed=shape[1] + len(run), dtype=dtype)
        x[1, 100:30_000] = np.nan
        got = getattr(nb, func)(x, limit=limit)
        np.testing.assert_array_equal(_bits(got), _bits(exact.fill(func, x, limit)))


@pytest.mark.parametrize("fill_bytes", ["16", "64"])
@pytest.mark.parametrize("func", ["ffill", "bfill"])
@pytest.mark.parametrize("two_pass", ["1", "0"])
def test_fill_long_nan_runs_cross_many_tiles(gpu_lib, monkeypatch, two_pass, fill_bytes, func):
    monkeypatch.setenv("NBG_FILL_TWO_PASS", two_pass)
    monkeypatch.setenv("NBG_FILL_BYTES", fill_bytes)
    # All-NaN stretches longer than several tiles force the look-back to walk
    # past tiles that hold no valid value.
    import numbagg_amd as nb
    from oracle import exact

    x = np.full((4, 60_000), np.nan, dtype=np.float32)
    x[0, 5] = 1.5
    x[1, 59_990] = -2.0
    x[2, 30_000] = 3.0
    for limit in (None, 10_000, 40_000):
        got = getattr(nb, func)(x, limit=limit)
        np.testing.assert_array_equal(_bits(got), _bits(exact.fill(func, x, limit)))


def test_fill_reference_pandas_vectors(gpu_lib, golden):
    import numbagg_amd as nb

    a = golden["array_3x500"]
    for func in ("ffill", "bfill"):
        for limit in (1, 3, None):
            np.testing.assert_array_equal(getattr(nb, func)(a, limit=limit), golden[f"{func}_l{limit}"])


def test_fill_axes_strides_and_dtypes(gpu_lib):
    import torch

    import numbagg_amd as nb
    from oracle import exact

    rng = np.random.default_rng(3)
    x = rng.standard_normal((6, 7, 900))
    x[rng.random(x.shape) < 0.4] = np.nan
    for axis in (0, 1, 2, -1):
        got = nb.ffill(x, limit=3, axis=axis)
        xm = np.moveaxis(x, axis, -1)
        want = exact.fill("ffill", xm.reshape(-1, xm.shape[-1]), 3).reshape(xm.shape)
        np.testing.assert_array_equal(got, np.moveaxis(want, -1, axis))
    # non-contiguous / unaligned rows on the device (scalar-load kernel path)
    t = torch.from_numpy(x.astype(np.float32)).cuda()[:, :, 1:]
    got = nb.bfill(t, limit=2).cpu().numpy()
    want = exact.fill("bfill", x.astype(np.float32)[:, :, 1:].reshape(-1, 899), 2).reshape(6, 7, 899)
    np.testing.assert_array_equal(got, want)
    # float16 goes through float32 exactly; integers are copied through
    h = x.astype(np.float16)
    got = nb.ffill(h)
    assert got.dtype == np.float16
    np.testing.assert_array_equal(got, exact.fill("ffill", h.reshape(-1, 900).astype(np.float32)).reshape(h.shape).astype(np.float16))
    i = rng.integers(-5, 5, size=(4, 33)).astype(np.int32)
    np.testing.assert_array_equal(nb.bfill(i, limit=1), i)



@pytest.mark.parametrize("two_pass", ["1", "0"])
@pytest.mark.parametrize("func", ["ffill", "bfill"])
def test_fill_huge_limit(gpu_lib, monkeypatch, two_pass, func):
    """Limits up to INT64_MAX fill like limit=None (the reference accepts any
    non-negative limit); on the two-pass path `carry + limit + 1` must not
    overflow."""
    import sys

    import numbagg_amd as nb
    from oracle import exact

    monkeypatch.setenv("NBG_FILL_TWO_PASS", two_pass)
    x = runs_array(16, 70_000, seed=5)
    x[1, 100:30_000] = np.nan
    want = exact.fill(func, x, None)
    for limit in (np.iinfo(np.int64).max, sys.maxsize - 1, 70_000, 1 << 40):
        got = getattr(nb, func)(x, limit=int(limit))
        np.testing.assert_array_equal(_bits(got), _bits(want))
