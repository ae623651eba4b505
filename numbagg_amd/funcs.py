"""ffill / bfill (numbagg/funcs.py:294-326) on the gfx950 fill kernel."""

from __future__ import annotations

from .decorators import ndfill


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


__all__ = ["bfill", "ffill"]
