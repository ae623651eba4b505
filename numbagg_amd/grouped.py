"""Grouped reductions (numbagg/grouped.py) on the gfx950 group-by kernels.

Per-label accumulators are privatised in LDS when the label count fits, or kept
in global memory otherwise. Float sums are accumulated in float64 and rounded
once to the output dtype; counts and indices are exact integers; argmax/argmin
keep the reference's first-index tie rule; first/last are exact value copies.
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
]
