"""Pins the fill oracle (CPU): against the pandas golden vectors computed on the
reference's own test inputs, against the closed form, and against brute force."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import exact, restated


@pytest.mark.parametrize("func", ["ffill", "bfill"])
@pytest.mark.parametrize("limit", [1, 3, None])
def test_restated_fill_matches_pandas(golden, func, limit):
    a = golden["array_3x500"]
    want = golden[f"{func}_l{limit}"]
    got = restated.fill(func, a, limit)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(exact.fill(func, a, limit), want)


def _brute_fill(x, limit, reverse):
    # The reference's loop, written out per element (numbagg/funcs.py:294-326).
    out = np.empty_like(x)
    for r in range(x.shape[0]):
        order = range(x.shape[1] - 1, -1, -1) if reverse else range(x.shape[1])
        lives, cur = limit, np.nan
        for i in order:
            v = x[r, i]
            if np.isnan(v):
                if lives <= 0:
                    cur = np.nan
                lives -= 1
            else:
                lives, cur = limit, v
            out[r, i] = cur
    return out


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("limit", [0, 1, 2, 7, None])
def test_fill_oracles_agree_with_brute_force(dtype, limit):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((5, 97)).astype(dtype)
    x[rng.random(x.shape) < 0.5] = np.nan
    x[2] = np.nan  # an all-NaN row
    lim = x.shape[1] if limit is None else limit
    for func, rev in (("ffill", False), ("bfill", True)):
        want = _brute_fill(x, lim, rev)
        np.testing.assert_array_equal(restated.fill(func, x, limit), want)
        np.testing.assert_array_equal(exact.fill(func, x, limit), want)


def test_fill_unfilled_slots_are_canonical_nan():
    x = np.array([[np.nan, 1.0, np.nan, np.nan, np.nan]], dtype=np.float32)
    # a NaN with a non-canonical payload is never copied to the output
    x.view(np.uint32)[0, 0] = 0x7FC00123
    out = restated.fill("ffill", x, 1)
    assert out.view(np.uint32)[0, 0] == 0x7FC00000
    assert out.view(np.uint32)[0, 3] == 0x7FC00000
    assert out[0, 2] == 1.0


def test_planted_mistake_is_caught():
    # A plausible bug (off-by-one in the limit) must fail the brute-force check.
    rng = np.random.default_rng(5)
    x = rng.standard_normal((3, 50))
    x[rng.random(x.shape) < 0.5] = np.nan
    wrong = exact.fill("ffill", x, 2 + 1)
    assert not np.array_equal(wrong, _brute_fill(x, 2, False), equal_nan=True)
