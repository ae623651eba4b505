"""Exponentially weighted moving functions (numbagg/moving_exp.py) on gfx950.

State is carried in float64, as in the reference (its accumulators start from
the literal 0.0); tiles of a row are chained with a decoupled look-back over
the affine map each tile applies to the state.
"""

from __future__ import annotations

from .decorators import ndmoveexp


@ndmoveexp.wrap()
def move_exp_nancount(a, alpha, min_weight, out):
    """Exponentially weighted count of non-NaN values (numbagg/moving_exp.py:12-38)."""


@ndmoveexp.wrap()
def move_exp_nanmean(a, alpha, min_weight, out):
    """Exponentially weighted moving mean (numbagg/moving_exp.py:41-72)."""


@ndmoveexp.wrap()
def move_exp_nansum(a, alpha, min_weight, out):
    """Exponentially weighted moving sum; NaN before the first value (numbagg/moving_exp.py:75-103)."""


@ndmoveexp.wrap()
def move_exp_nanvar(a, alpha, min_weight, out):
    """Exponentially weighted moving variance, bias-corrected (numbagg/moving_exp.py:106-156)."""


@ndmoveexp.wrap()
def move_exp_nanstd(a, alpha, min_weight, out):
    """Square root of the bias-corrected exponential variance (numbagg/moving_exp.py:159-224)."""


@ndmoveexp.wrap(nargs=2)
def move_exp_nancov(a1, a2, alpha, min_weight, out):
    """Exponentially weighted moving covariance (numbagg/moving_exp.py:227-273)."""


@ndmoveexp.wrap(nargs=2)
def move_exp_nancorr(a1, a2, alpha, min_weight, out):
    """Exponentially weighted moving correlation (numbagg/moving_exp.py:276-335)."""


__all__ = [
    "move_exp_nancount",
    "move_exp_nanmean",
    "move_exp_nansum",
    "move_exp_nanvar",
    "move_exp_nanstd",
    "move_exp_nancov",
    "move_exp_nancorr",
]
