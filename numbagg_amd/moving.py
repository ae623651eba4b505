"""Moving-window functions (numbagg/moving.py) on the gfx950 rolling kernels.

Window statistics are computed from float64 running sums for float32 data
(first moments; FP32 TwoSum pairs are a build-time alternative), float64
second/co-moments around a per-tile shift, and double-double sums for float64
data, so results are closer to the exact values than the reference's running
add/subtract accumulators, while the valid-count masking (`min_count`) is
identical.
"""

from __future__ import annotations

from .decorators import ndmove


@ndmove.wrap()
def move_mean(a, window, min_count, out):
    """Moving window mean; ``min_count`` is raised to at least 1 (numbagg/moving.py:12-55)."""


@ndmove.wrap()
def move_sum(a, window, min_count, out):
    """Moving window sum; ``min_count=0`` gives 0.0 for empty windows (numbagg/moving.py:58-99)."""


@ndmove.wrap()
def move_std(a, window, min_count, out):
    """Moving window standard deviation, ddof=1; ``min_count`` >= 2 (numbagg/moving.py:123-151)."""


@ndmove.wrap()
def move_var(a, window, min_count, out):
    """Moving window variance, ddof=1; ``min_count`` >= 2 (numbagg/moving.py:154-181)."""


@ndmove.wrap(nargs=2)
def move_cov(a, b, window, min_count, out):
    """Moving window covariance over pairwise-complete observations (numbagg/moving.py:184-220)."""


@ndmove.wrap(nargs=2)
def move_corr(a, b, window, min_count, out):
    """Moving window correlation over pairwise-complete observations (numbagg/moving.py:223-275)."""


__all__ = ["move_mean", "move_sum", "move_std", "move_var", "move_cov", "move_corr"]
