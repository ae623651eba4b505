"""Grouped reductions (numbagg/grouped.py) on the gfx950 group-by kernels.

Per-label accumulators are privatised in LDS when the label count fits, or kept
in global memory otherwise. Float sums, sums of squares and products are
accumulated in float64 and rounded once to the output dtype (integer ones wrap
like the reference's fixed-width arithmetic); counts and indices are exact
integers; argmax/argmin and min/max keep the reference's first-index tie rule;
first/last/min/max are exact value copies.
"""

from __future__ import annotations

from .decorators import groupndreduce


@groupndreduce.wrap(supports_ints=False)
def group_nanmean(values, labels, out):
    """Mean of the non-NaN values of each label; NaN for empty labels (numbagg/grouped.py:7-27)."""


@groupndreduce.wrap()
def group_nansum(values, labels, out):
    """Sum of the non-NaN values of each label; 0 for empty labels (numbagg/grouped.py:30-40)."""


@groupndreduce.wrap()
def group_nancount(values, labels, out):
    """Number of non-NaN values of each label (numbagg/grouped.py:43-51)."""


@groupndreduce.wrap()
def group_nanargmax(values, labels, out):
    """Flat index of each label's maximum; the first index wins ties (numbagg/grouped.py:54-75)."""


@groupndreduce.wrap()
def group_nanargmin(values, labels, out):
    """Flat index of each label's minimum; the first index wins ties (numbagg/grouped.py:78-92)."""


@groupndreduce.wrap()
def group_nanfirst(values, labels, out):
    """First non-NaN value of each label (numbagg/grouped.py:95-110)."""


@groupndreduce.wrap()
def group_nanlast(values, labels, out):
    """Last non-NaN value of each label (numbagg/grouped.py:113-121)."""


@groupndreduce.wrap(supports_bool=False, supports_ints=False, supports_ddof=True)
def group_nanvar(values, labels, ddof, out):
    """Variance of each label's non-NaN values with ``ddof`` (numbagg/grouped.py:148-176)."""


@groupndreduce.wrap(supports_bool=False, supports_ints=False, supports_ddof=True)
def group_nanstd(values, labels, ddof, out):
    """Standard deviation of each label's non-NaN values (numbagg/grouped.py:179-208)."""


@groupndreduce.wrap()
def group_nanmin(values, labels, out):
    """Minimum of the non-NaN values of each label; NaN for empty labels (numbagg/grouped.py:211-225)."""


@groupndreduce.wrap()
def group_nanmax(values, labels, out):
    """Maximum of the non-NaN values of each label; NaN for empty labels (numbagg/grouped.py:228-244)."""


@groupndreduce.wrap()
def group_nanprod(values, labels, out):
    """Product of the non-NaN values of each label; 1 for empty labels (numbagg/grouped.py:124-132)."""


@groupndreduce.wrap()
def group_nansum_of_squares(values, labels, out):
    """Sum of the squares of the non-NaN values of each label (numbagg/grouped.py:135-145)."""


@groupndreduce.wrap()
def group_nanany(values, labels, out):
    """1 where a label holds a non-NaN non-zero value, else 0 (numbagg/grouped.py:247-257)."""


@groupndreduce.wrap()
def group_nanall(values, labels, out):
    """0 where a label holds a non-NaN zero, else 1 (numbagg/grouped.py:260-270)."""


__all__ = [
    "group_nanmean",
    "group_nansum",
    "group_nancount",
    "group_nanargmax",
    "group_nanargmin",
    "group_nanfirst",
    "group_nanlast",
    "group_nanvar",
    "group_nanstd",
    "group_nanmin",
    "group_nanmax",
    "group_nanprod",
    "group_nansum_of_squares",
    "group_nanany",
    "group_nanall",
]
