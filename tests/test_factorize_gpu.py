"""GPU parity of the label factorisation against the oracle (pinned to
pandas.factorize): bit-exact codes and uniques for every dtype, NaN, signed
zeros, -1 and extreme integers, empty and all-NaN input, sort on and off, up
to the (4M,) label vector of the grouped benchmark (1000 uniform or Zipf(1.1)
labels), 2048 and 2050 uniques (either side of the LDS code table) and 1M
distinct values; the codes then drive a grouped reduction."""

from __future__ import annotations

import numpy as np
import pytest
from numpy.testing import assert_array_equal

pytestmark = pytest.mark.gpu


def _check(v, sort):
    import numbagg_amd as nb
    from oracle import exact

    codes, uniques = nb.factorize(v, sort=sort)
    wc, wu = exact.factorize(v, sort=sort)
    assert codes.dtype == np.int64
    assert_array_equal(codes, wc)
    assert uniques.dtype == v.dtype
    assert_array_equal(uniques.view(np.uint8), wu.view(np.uint8))


@pytest.mark.parametrize("sort", [True, False])
def test_edge_cases(gpu_lib, sort):
    cases = [
        np.array([3, -1, 2, -1, 7, 3], dtype=np.int64),
        np.array([np.iinfo(np.int64).max, np.iinfo(np.int64).min, -1, 0, -1, np.iinfo(np.int64).max], dtype=np.int64),
        np.array([5, -1, 5, 2**31 - 1, -(2**31)], dtype=np.int32),
        np.array([-0.0, 0.0, 1.0, np.nan, -2.5, 1.0]),
        np.array([0.0, -0.0, np.inf, -np.inf, np.nan], dtype=np.float32),
        np.array([np.nan, np.nan]),
        np.array([], dtype=np.int32),
        np.array([7], dtype=np.int64),
    ]
    for v in cases:
        _check(v, sort)


@pytest.mark.parametrize("dtype", [np.int32, np.int64, np.float32, np.float64])
@pytest.mark.parametrize("sort", [True, False])
def test_random_labels(gpu_lib, dtype, sort):
    rng = np.random.default_rng(3)
    # 2k uniques: 2048 fit the LDS code table (one of them the integer -1,
    # kept outside the hash table), 2050 do not; 1M distinct values overflow
    # every workgroup's LDS insert table and the small global table; 30000
    # uniques fill the small table to 46%
    cases = ((4_000_000, 1000, False), (4_000_000, 1000, True), (1_000_003, 10**9, False), (5000, 3, False),
             (300_000, 1024, False), (300_000, 1025, False), (300_000, 15_000, False))
    for n, k, zipf in cases:
        if zipf:
            p = np.arange(1, k + 1, dtype=np.float64) ** -1.1
            v = rng.choice(k, size=n, p=p / p.sum()).astype(dtype)
        else:
            v = rng.integers(-k, k, size=n).astype(dtype)
        if np.issubdtype(dtype, np.floating):
            v[rng.random(n) < 0.05] = np.nan
        _check(v, sort)


def test_codes_drive_group_by(gpu_lib):
    import numbagg_amd as nb
    from oracle import exact

    rng = np.random.default_rng(5)
    raw = rng.choice(np.array([10.5, -3.0, 7.25, np.nan]), size=200_000)
    vals = rng.standard_normal((3, 200_000)).astype(np.float32)
    codes, uniques = nb.factorize(raw)
    assert_array_equal(uniques, [-3.0, 7.25, 10.5])
    got = nb.group_nansum(vals, codes, num_labels=len(uniques), axis=-1)
    want = exact.group("group_nansum", vals, codes, len(uniques))
    np.testing.assert_allclose(got, want, rtol=1e-6)
    import torch

    t_codes, t_uniques = nb.factorize(torch.from_numpy(raw).cuda())
    assert t_codes.is_cuda and t_uniques.is_cuda
    assert_array_equal(t_codes.cpu().numpy(), codes)
