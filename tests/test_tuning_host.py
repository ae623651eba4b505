"""Host-side logic of the ragged strategy tuner (CPU): series chunking and the
per-series batch statistics that feed the cost model."""

from __future__ import annotations

import types

import numpy as np
import pytest
import torch


def _fake_batch(lengths, ks, seed=0):
    rng = np.random.default_rng(seed)
    lengths = np.asarray(lengths, dtype=np.int64)
    ks = np.asarray(ks, dtype=np.int64)
    offs = np.concatenate([[0], np.cumsum(lengths)])
    loffs = np.concatenate([[0], np.cumsum(ks)])
    vals = rng.standard_normal(int(offs[-1])).astype(np.float32)
    vals[rng.random(vals.size) < 0.1] = np.nan
    labs = np.concatenate([rng.integers(-1, k + 2, n) for n, k in zip(lengths, ks)]).astype(np.int32)
    rb = types.SimpleNamespace(nseries=len(lengths), lengths=lengths, num_labels=ks, offsets=offs,
                               label_offsets=loffs)
    return rb, vals, labs


def test_series_chunks_cover_in_order():
    from numbagg_amd.tuning import series_chunks

    lengths = np.array([5, 7, 100, 1, 1, 3, 50, 2])
    chunks = list(series_chunks(lengths, max_elements=10))
    assert chunks[0][0] == 0 and chunks[-1][1] == len(lengths)
    for (a0, a1), (b0, b1) in zip(chunks, chunks[1:]):
        assert a1 == b0 and a0 < a1
    for s0, s1 in chunks:
        assert s1 - s0 == 1 or lengths[s0:s1].sum() <= 10
    assert list(series_chunks(np.array([], dtype=np.int64))) == []


@pytest.mark.parametrize("max_elements", [1 << 27, 64, 1])
def test_batch_statistics_matches_per_series(max_elements):
    from numbagg_amd.tuning import batch_statistics

    rb, vals, labs = _fake_batch([30, 200, 1, 77, 500, 9], [3, 1, 5, 40, 2, 1])
    skew, nanf = batch_statistics(rb, torch.from_numpy(vals), torch.from_numpy(labs), max_elements=max_elements)
    for s in range(rb.nseries):
        o0, o1 = rb.offsets[s], rb.offsets[s + 1]
        v, lab, k = vals[o0:o1], labs[o0:o1], rb.num_labels[s]
        ok = (lab >= 0) & (lab < k)
        top = np.bincount(lab[ok], minlength=k).max() if ok.any() else 0
        assert skew[s] == pytest.approx(top / (o1 - o0))
        assert nanf[s] == pytest.approx(np.isnan(v).mean())


def test_features_shape_and_logs():
    from numbagg_amd.tuning import features

    f = features([1024, 8], [16, 1], 0.5, 0.05, 50)
    assert f.shape == (2, 5) and f.dtype == np.float32
    np.testing.assert_allclose(f[:, 0], [10, 3])
    np.testing.assert_allclose(f[:, 1], [4, 0])
    np.testing.assert_allclose(f[:, 4], np.log2(50), rtol=1e-6)


@pytest.mark.parametrize("sample", [1, 16, 100])
def test_batch_statistics_prefix_sample(sample):
    from numbagg_amd.tuning import batch_statistics

    rb, vals, labs = _fake_batch([30, 200, 1, 77, 500, 9], [3, 1, 5, 40, 2, 1])
    skew, nanf = batch_statistics(rb, torch.from_numpy(vals), torch.from_numpy(labs), sample=sample)
    for s in range(rb.nseries):
        o0 = rb.offsets[s]
        o1 = min(rb.offsets[s + 1], o0 + sample)
        v, lab, k = vals[o0:o1], labs[o0:o1], rb.num_labels[s]
        ok = (lab >= 0) & (lab < k)
        top = np.bincount(lab[ok], minlength=k).max() if ok.any() else 0
        assert skew[s] == pytest.approx(top / (o1 - o0))
        assert nanf[s] == pytest.approx(np.isnan(v).mean())


def test_cost_model_load_missing_or_unreadable(tmp_path, monkeypatch):
    """No GPU needed: a missing or unreadable model file loads as None (the
    launch choice then falls back to the exhaustive tuner)."""
    from numbagg_amd.tuning import CostModel

    assert CostModel.load(tmp_path / "missing.npz") is None
    bad = tmp_path / "bad.npz"
    bad.write_bytes(b"\x00garbage")
    assert CostModel.load(bad) is None
    monkeypatch.setenv("NBG_COST_MODEL", str(bad))
    assert CostModel.load() is None


def test_launch_candidates_add_the_uniform_launch():
    """Each close slot count is timed with the model's strategies and, when
    they are mixed, with every series on the strategy holding most elements."""
    from numbagg_amd.tuning import STRATEGY_GLOBAL, STRATEGY_LDS, launch_candidates

    lengths = np.array([1000, 10, 10, 500])
    all_lds = np.full(4, STRATEGY_LDS, np.uint8)
    mixed = np.array([STRATEGY_LDS, STRATEGY_GLOBAL, STRATEGY_GLOBAL, STRATEGY_LDS], np.uint8)
    mostly_global = np.array([STRATEGY_GLOBAL, STRATEGY_LDS, STRATEGY_LDS, STRATEGY_GLOBAL], np.uint8)
    c = launch_candidates([(1.0, 4096, all_lds)], lengths)
    assert len(c) == 1 and c[0][0] == 4096 and np.array_equal(c[0][1], all_lds)
    c = launch_candidates([(1.0, 4096, mixed), (1.5, 1024, mostly_global)], lengths)
    assert [cap for cap, _ in c] == [4096, 4096, 1024, 1024]
    assert np.array_equal(c[0][1], mixed) and np.array_equal(c[1][1], all_lds)
    assert np.array_equal(c[2][1], mostly_global) and np.all(c[3][1] == STRATEGY_GLOBAL)
