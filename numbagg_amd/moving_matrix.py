"""Moving-window and exponentially weighted correlation / covariance matrices
(numbagg/moving_matrix.py) on the gfx950 matrix kernels.

Inputs are (..., obs, vars); outputs (..., obs, vars, vars). Statistics are
pairwise complete and formed from float64 sums of shifted values, so results
are closer to the exact values than the reference's accumulators in the data
dtype; the validity rules (min_count, min_weight, bias) are the reference's.
"""

from __future__ import annotations

from .decorators import ndmoveexpmatrix, ndmovematrix


@ndmovematrix.wrap()
def move_corrmatrix(a, window, min_count, out):
    """Moving window correlation matrix; NaN where a pair has fewer than
    max(min_count, 2) observations or a variance is not positive
    (numbagg/moving_matrix.py:16-114)."""


@ndmovematrix.wrap()
def move_covmatrix(a, window, min_count, out):
    """Moving window covariance matrix (ddof 1); NaN below max(min_count, 1)
    observations or with a single one (numbagg/moving_matrix.py:117-201)."""


@ndmoveexpmatrix.wrap()
def move_exp_nancorrmatrix(a, alpha, min_weight, out):
    """Exponentially weighted correlation matrix (numbagg/moving_matrix.py:204-328)."""


@ndmoveexpmatrix.wrap()
def move_exp_nancovmatrix(a, alpha, min_weight, out):
    """Exponentially weighted, bias-corrected covariance matrix
    (numbagg/moving_matrix.py:331-432)."""


__all__ = ["move_corrmatrix", "move_covmatrix", "move_exp_nancorrmatrix", "move_exp_nancovmatrix"]
