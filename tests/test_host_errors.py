"""Host-layer argument handling (CPU): the reference's validation errors are
raised before any device work, with the reference's messages
(numbagg/decorators.py:305-341, :374-414, :465-487, :558-674)."""

from __future__ import annotations

import numpy as np
import pytest

import numbagg_amd as nb


def test_fill_errors():
    with pytest.raises(ValueError, match="limit"):
        nb.ffill(np.ones(3), limit=-1)
    with pytest.raises(TypeError, match="Unsupported dtype"):
        nb.ffill(np.array(["a", "b"]))


def test_repr_and_names():
    assert repr(nb.ffill) == "numbagg.ffill"
    assert nb.move_mean.__name__ == "move_mean"
    assert "window" in str(nb.move_mean.__signature__)
    assert nb.group_nanvar.supports_ddof and not nb.group_nansum.supports_ddof
    assert not nb.group_nanmean.supports_ints and nb.group_nansum.supports_ints
    assert not nb.group_nanvar.supports_bool
