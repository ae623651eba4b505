"""ffill / bfill (numbagg/funcs.py:294-326) on the gfx950 fill kernel, and the
plain reductions (numbagg/funcs.py:23-228) on the gfx950 reduction kernel."""

from __future__ import annotations

import numpy as np

from .decorators import ndaggregate, ndfill, ndmatrix, ndreduce

_NUMERIC = [np.int32, np.int64, np.float32, np.float64]
_FLOAT = [np.float32, np.float64]


@ndaggregate.wrap(loops=_NUMERIC, out_dtype=np.bool_)
def allnan(a, out):
    """True where every value along the axes is NaN (numbagg/funcs.py:23-35)."""


@ndaggregate.wrap(loops=_NUMERIC, out_dtype=np.bool_)
def anynan(a, out):
    """True where any value along the axes is NaN (numbagg/funcs.py:38-50)."""


@ndaggregate.wrap(loops=_NUMERIC, out_dtype=np.int64)
def nancount(a, out):
    """Number of non-NaN values along the axes (numbagg/funcs.py:53-66)."""


@ndaggregate.wrap(loops=_NUMERIC)
def nansum(a, out):
    """Sum of the non-NaN values along the axes; float sums are formed in float64
    (double-double for float64 input) and rounded once (numbagg/funcs.py:69-82)."""


@ndaggregate.wrap(loops=_FLOAT)
def nanmean(a, out):
    """Mean of the non-NaN values along the axes (numbagg/funcs.py:85-102)."""


@ndaggregate.wrap(loops=_FLOAT, supports_ddof=True)
def nanvar(a, ddof, out):
    """Variance of the non-NaN values with ``ddof`` (default 1), from per-lane
    Chan merges of two-pass batch moments (numbagg/funcs.py:105-130)."""


@ndaggregate.wrap(loops=_FLOAT, supports_ddof=True)
def nanstd(a, ddof, out):
    """Standard deviation of the non-NaN values with ``ddof`` (numbagg/funcs.py:133-157)."""


@ndreduce.wrap(index=True)
def nanargmax(a):
    """Flat index of the maximum over the axes, ignoring NaN; the first index wins
    ties (numbagg/funcs.py:160-178)."""


@ndreduce.wrap(index=True)
def nanargmin(a):
    """Flat index of the minimum over the axes, ignoring NaN; the first index wins
    ties (numbagg/funcs.py:181-197)."""


@ndreduce.wrap(index=False)
def nanmax(a):
    """Maximum over the axes, ignoring NaN; NaN where every value is NaN
    (numbagg/funcs.py:200-216)."""


@ndreduce.wrap(index=False)
def nanmin(a):
    """Minimum over the axes, ignoring NaN; NaN where every value is NaN
    (numbagg/funcs.py:219-234)."""


count = nancount


@ndmatrix.wrap()
def nancorrmatrix(a, out):
    """Correlation matrix of (..., vars, obs) data over pairwise-complete
    observations: float64 sums on the matrix cores, NaN where a pair has fewer
    than two observations or a variance is not positive (numbagg/funcs.py:338-441)."""


@ndmatrix.wrap()
def nancovmatrix(a, out):
    """Covariance matrix (ddof 1) of (..., vars, obs) data over
    pairwise-complete observations (numbagg/funcs.py:444-532)."""


@ndfill.wrap()
def bfill(a, limit, out):
    """Backward fill missing values.

    HIP kernel ``nbg_bfill``: each NaN takes the next valid value along the axis
    if it lies at most ``limit`` positions away (reference: numbagg/funcs.py:294-310).
    """


@ndfill.wrap()
def ffill(a, limit, out):
    """Forward fill missing values.

    HIP kernel ``nbg_ffill``: each NaN takes the previous valid value along the
    axis if it lies at most ``limit`` positions away (reference: numbagg/funcs.py:313-326).
    """


__all__ = [
    "allnan",
    "anynan",
    "bfill",
    "count",
    "ffill",
    "nanargmax",
    "nanargmin",
    "nancorrmatrix",
    "nancount",
    "nancovmatrix",
    "nanmax",
    "nanmean",
    "nanmin",
    "nanstd",
    "nansum",
    "nanvar",
]
