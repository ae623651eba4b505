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


def test_move_errors():
    # numbagg/test/test_moving.py:168-177
    array = np.random.RandomState(0).rand(3, 500)
    with pytest.raises(TypeError):
        nb.move_mean(array, window=0.5)
    with pytest.raises(ValueError):
        nb.move_mean(array, window=-1)
    with pytest.raises(ValueError, match="window not in valid range"):
        nb.move_mean(array, window=array.shape[-1] + 1)
    with pytest.raises(ValueError, match="min_count must be positive"):
        nb.move_mean(array, window=1, min_count=-1)
    x = np.arange(10.0)
    assert nb.move_mean(x, window=2, axis=()) is x
    with pytest.raises(ValueError, match="only one axis"):
        nb.move_mean(x[None], window=2, axis=(0, 1))
    with pytest.raises(ValueError, match="empty tuple"):
        nb.move_corr(x, x, window=2, axis=())


def test_move_exp_errors():
    x = np.arange(10.0)
    assert nb.move_exp_nanmean(x, alpha=0.5, axis=()) is x
    with pytest.raises(ValueError, match="Only one axis"):
        nb.move_exp_nanmean(x[None], alpha=0.5, axis=(0, 1))
    with pytest.raises(ValueError, match="empty tuple"):
        nb.move_exp_nancorr(x, x, alpha=0.5, axis=())


def test_group_errors():
    values2d = np.arange(25.0).reshape(5, 5)
    labels = np.arange(5)
    with pytest.raises(ValueError, match="axis required"):
        nb.group_nansum(values2d, labels)
    with pytest.raises(ValueError, match="must have same shape"):
        nb.group_nansum(values2d, labels[:4], axis=0)
    with pytest.raises(ValueError, match="must have same shape"):
        nb.group_nansum(values2d, labels[:4], axis=(0,))
    with pytest.raises(TypeError, match="labels must be an integer array"):
        nb.group_nansum(values2d, labels.astype(float), axis=1)
    with pytest.raises(TypeError, match="does not support boolean input"):
        nb.group_nanvar(np.array([True, False]), np.array([0, 0]), num_labels=1)


def test_spin_timeout_increase_raises(monkeypatch):
    # The synchronising NumPy path reports an expired device-side wait instead
    # of returning its (invalid) outputs; the counter is simulated on the host.
    import torch

    from numbagg_amd import _device, _lib

    counts = iter([0, 0, 2, 2])
    monkeypatch.setattr(_lib, "spin_timeouts", lambda: next(counts))
    monkeypatch.setattr(torch.cuda, "current_device", lambda: 0)
    monkeypatch.setattr(_device, "_spin_seen", {})
    _device.check_spin_timeouts()
    _device.check_spin_timeouts()
    with pytest.raises(RuntimeError, match="wait"):
        _device.check_spin_timeouts()
    _device.check_spin_timeouts()  # no new expiry since the last check


def test_out_argument_delivery():
    """out= as the reference's gufuncs take it (numbagg/decorators.py:341, :414,
    :487, :731, :807, :1089): exact shape, casting 'same_kind', the out array
    returned; other keywords are rejected."""
    res = np.arange(6.0).reshape(2, 3)
    dest = np.empty((2, 3), dtype=np.float32)
    assert nb.move_mean._deliver(res, dest) is dest
    np.testing.assert_array_equal(dest, res.astype(np.float32))
    assert nb.ffill._deliver(res, (dest,)[0]) is dest
    with pytest.raises(TypeError, match="same_kind"):
        nb.move_mean._deliver(res, np.empty((2, 3), dtype=np.int64))
    with pytest.raises(ValueError, match="broadcast shape"):
        nb.move_mean._deliver(res, np.empty((3, 2)))
    assert nb.move_mean._take_out({"out": (dest,)}) is dest
    with pytest.raises(ValueError, match="exactly one"):
        nb.move_mean._take_out({"out": (dest, dest)})
    with pytest.raises(TypeError, match="ArrayType"):
        nb.move_mean._take_out({"out": [1, 2]})
    with pytest.raises(TypeError, match="unsupported keyword"):
        nb.move_mean._take_out({"out": dest, "bogus": 1})
    with pytest.raises(TypeError):
        nb.nancorrmatrix(np.ones((3, 4)), axis=0)
    with pytest.raises(ValueError, match="requires at least a 2D array"):
        nb.move_covmatrix(np.ones(5), window=2)
    with pytest.raises(ValueError, match="window not in valid range"):
        nb.move_corrmatrix(np.ones((5, 2)), window=6)
