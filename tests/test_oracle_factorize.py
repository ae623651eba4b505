"""Pins the factorisation oracle (CPU) to pandas.factorize, the function the
reference's benchmarks call (numbagg/test/conftest.py:175-189), on the edge
cases a hash-and-sort kernel can get wrong: NaN, signed zeros (the first one
seen represents both), -1 and the extreme integers, empty and all-NaN input,
and first-appearance order without sorting."""

from __future__ import annotations

import numpy as np
import pandas as pd
import pytest
from numpy.testing import assert_array_equal

from oracle import exact

CASES = [
    np.array([3, -1, 2, -1, 7, 3], dtype=np.int64),
    np.array([np.iinfo(np.int64).max, np.iinfo(np.int64).min, -1, 0, -1], dtype=np.int64),
    np.array([5, -1, 5, 2**31 - 1, -(2**31)], dtype=np.int32),
    np.array([-0.0, 0.0, 1.0, np.nan, -2.5, 1.0]),
    np.array([0.0, -0.0, np.inf, -np.inf, np.nan], dtype=np.float32),
    np.array([np.nan, np.nan]),
    np.array([], dtype=np.int32),
    np.random.default_rng(0).integers(-50, 50, 10_000).astype(np.int32),
]


@pytest.mark.parametrize("sort", [True, False])
@pytest.mark.parametrize("k", range(len(CASES)))
def test_matches_pandas(k, sort):
    v = CASES[k]
    codes, uniques = exact.factorize(v, sort=sort)
    pc, pu = pd.factorize(v, sort=sort)
    assert_array_equal(codes, pc)
    assert uniques.dtype == pu.dtype
    assert_array_equal(uniques.view(np.uint8), np.asarray(pu).view(np.uint8))  # signed zeros included


def test_planted_mistake():
    # "last occurrence represents" would give +0.0 for [-0.0, 0.0]
    _, u = exact.factorize(np.array([-0.0, 0.0]))
    assert np.signbit(u[0])
