"""Seeded synthetic inputs shared by the tests (shapes of the benchmark configs,
scaled down)."""

from __future__ import annotations

import numpy as np


def runs_mask(rows: int, n: int, mean_nan: float = 50.0, mean_valid: float = 116.7, seed: int = 0) -> np.ndarray:
    """Boolean NaN mask made of alternating runs with geometric lengths
    (NaN runs mean ``mean_nan``, valid runs mean ``mean_valid``: 30% NaN)."""
    rng = np.random.default_rng(seed)
    mask = np.empty((rows, n), dtype=bool)
    k = int(n / (mean_nan + mean_valid) * 1.5) + 16
    for r in range(rows):
        nan_l = rng.geometric(1.0 / mean_nan, size=k)
        val_l = rng.geometric(1.0 / mean_valid, size=k)
        first_nan = rng.random() < mean_nan / (mean_nan + mean_valid)
        lengths = np.empty(2 * k, dtype=np.int64)
        states = np.empty(2 * k, dtype=bool)
        if first_nan:
            lengths[0::2], lengths[1::2] = nan_l, val_l
            states[0::2], states[1::2] = True, False
        else:
            lengths[0::2], lengths[1::2] = val_l, nan_l
            states[0::2], states[1::2] = False, True
        row = np.repeat(states, lengths)
        while row.size < n:
            row = np.concatenate([row, row])
        mask[r] = row[:n]
    return mask


def runs_array(rows: int, n: int, seed: int = 0, dtype=np.float32) -> np.ndarray:
    rng = np.random.default_rng(seed + 1)
    x = rng.standard_normal((rows, n)).astype(dtype)
    x[runs_mask(rows, n, seed=seed)] = np.nan
    return x


def walk_array(rows: int, n: int, seed: int = 0, p_nan: float = 0.02, offset: float = 1e4, dtype=np.float32) -> np.ndarray:
    """Random walk plus an offset (stresses cancellation), with i.i.d. NaNs."""
    rng = np.random.default_rng(seed)
    x = (np.cumsum(rng.standard_normal((rows, n)), axis=-1) + offset).astype(dtype)
    x[rng.random(x.shape) < p_nan] = np.nan
    return x


def bits(a: np.ndarray) -> np.ndarray:
    return a.view({4: np.uint32, 8: np.uint64, 2: np.uint16}[a.dtype.itemsize])
