"""Ragged batched path: one persistent launch vs per-series plain launches.

move_mean outputs are bitwise equal to per-series plain move_mean calls (same
tiles, same arithmetic); group means match the per-series plain
group_nanmean calls (float64 sums rounded once; the summation order differs,
so equality is checked bitwise with at most 1-ULP slack) and the float64 oracle.
Graph replays are checked against the eager result, with eager work and
synchronisations between replays.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make_batch(nseries, lo, hi, klo, khi, seed, device="cuda"):
    import torch

    rng = np.random.default_rng(seed)
    lengths = np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), nseries))).astype(np.int64)
    ks = np.round(np.exp(rng.uniform(np.log(klo), np.log(khi), nseries))).astype(np.int64)
    total = int(lengths.sum())
    vals = rng.standard_normal(total).astype(np.float32)
    vals[rng.random(total) < 0.05] = np.nan
    labs = np.empty(total, dtype=np.int32)
    off = 0
    for L, K in zip(lengths, ks):
        u = rng.random(L)
        # Zipf(1.1)-like skew over K labels (continuous inverse CDF)
        x = (1.0 - u * (1.0 - float(K) ** -0.1)) ** -10.0
        labs[off : off + L] = np.clip(np.floor(x).astype(np.int64) - 1, 0, K - 1)
        off += L
    return lengths, ks, torch.from_numpy(vals).to(device), torch.from_numpy(labs).to(device)


def _bits(t):
    return t.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("strategy,lds_labels", [("heuristic", 1024), ("lds", 1024), ("global", 1024),
                                                 ("lds", 4096), ("lds", 100), ("lds", 8192)])
def test_ragged_matches_plain_launches(gpu_lib, strategy, lds_labels):
    from numbagg_amd.ragged import RaggedMoveMeanGroupMean

    lengths, ks, v, lab = make_batch(120, 100, 60_000, 10, 3000, seed=1)
    rb = RaggedMoveMeanGroupMean(lengths, ks, window=50, min_count=50, strategy=strategy, lds_labels=lds_labels)
    mm, gm = rb(v, lab)
    pm, pg = rb.plain(v, lab)
    np.testing.assert_array_equal(_bits(mm), _bits(pm))
    a, b = _bits(gm).astype(np.int64), _bits(pg).astype(np.int64)
    assert np.all(np.abs(a - b) <= 1)
    assert np.mean(a == b) > 0.999


@pytest.mark.parametrize("grab", ["1", "3", "17", "64"])
def test_ragged_grab_sizes(gpu_lib, monkeypatch, grab):
    """Any number of tiles per queue grab (series split over grabs or several
    series in one grab, LDS slots flushed per grab) gives the same results."""
    from numbagg_amd.ragged import RaggedMoveMeanGroupMean

    monkeypatch.setenv("NBG_RAGGED_GRAB", grab)
    lengths, ks, v, lab = make_batch(150, 100, 80_000, 10, 9000, seed=int(grab) + 5)
    rb = RaggedMoveMeanGroupMean(lengths, ks, window=50, min_count=50, strategy="lds", lds_labels=1024)
    mm, gm = rb(v, lab)
    pm, pg = rb.plain(v, lab)
    np.testing.assert_array_equal(_bits(mm), _bits(pm))
    a, b = _bits(gm).astype(np.int64), _bits(pg).astype(np.int64)
    assert np.all(np.abs(a - b) <= 1)
    assert np.mean(a == b) > 0.999


def test_ragged_grab_out_of_range(gpu_lib, monkeypatch):
    from numbagg_amd.ragged import RaggedMoveMeanGroupMean

    lengths, ks, v, lab = make_batch(4, 100, 1000, 10, 20, seed=3)
    rb = RaggedMoveMeanGroupMean(lengths, ks, window=50, min_count=50)
    for bad in ("0", "65"):
        monkeypatch.setenv("NBG_RAGGED_GRAB", bad)
        with pytest.raises(RuntimeError, match="NBG_RAGGED_GRAB"):
            rb(v, lab)


@pytest.mark.parametrize("strategy,khi", [("heuristic", 2000), ("lds", 50_000), ("global", 50_000)])
def test_ragged_against_oracle(gpu_lib, strategy, khi):
    """Includes series with more labels than the LDS slots: with "lds" their
    low labels are privatised and the rest go to global atomics."""
    from numbagg_amd.ragged import RaggedMoveMeanGroupMean
    from oracle import exact

    lengths, ks, v, lab = make_batch(40, 100, 20_000, 10, khi, seed=2)
    rb = RaggedMoveMeanGroupMean(lengths, ks, window=50, min_count=40, strategy=strategy)
    mm, gm = rb(v, lab)
    vn, ln, mmn, gmn = v.cpu().numpy(), lab.cpu().numpy(), mm.cpu().numpy(), gm.cpu().numpy()
    for s in range(len(lengths)):
        o, n = rb.offsets[s], lengths[s]
        want = exact.move("move_mean", vn[o : o + n], 50, 40)[0]
        got = mmn[o : o + n]
        assert np.array_equal(np.isnan(got), np.isnan(want))
        ok = ~np.isnan(want)
        assert np.allclose(got[ok], want[ok], rtol=2e-7, atol=1e-7)
        lo, k = rb.label_offsets[s], ks[s]
        gwant = exact.group("group_nanmean", vn[o : o + n], ln[o : o + n], int(k))[0]
        ggot = gmn[lo : lo + k]
        assert np.array_equal(np.isnan(ggot), np.isnan(gwant))
        ok = ~np.isnan(gwant)
        assert np.allclose(ggot[ok], gwant[ok], rtol=1e-6, atol=1e-7)


def test_ragged_graph_replays(gpu_lib):
    import torch

    from numbagg_amd.ragged import RaggedMoveMeanGroupMean

    lengths, ks, v, lab = make_batch(60, 100, 30_000, 10, 2000, seed=3)
    rb = RaggedMoveMeanGroupMean(lengths, ks, window=50)
    em, eg = rb(v, lab)
    em, eg = _bits(em), _bits(eg).astype(np.int64)
    replay, mo, go = rb.graph(v, lab)
    for i in range(4):
        mo.fill_(0)
        go.fill_(0)
        replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(_bits(mo), em)
        assert np.all(np.abs(_bits(go).astype(np.int64) - eg) <= 1)
        # eager work between replays
        _ = rb(v, lab)
        torch.cuda.synchronize()


def test_ragged_errors(gpu_lib):
    from numbagg_amd.ragged import RaggedMoveMeanGroupMean

    with pytest.raises(ValueError, match="window not in valid range"):
        RaggedMoveMeanGroupMean([100, 50], [3, 3], window=60)
    with pytest.raises(ValueError, match="min_count must be positive"):
        RaggedMoveMeanGroupMean([100, 50], [3, 3], window=10, min_count=-1)


def test_ragged_lds_labels_range(gpu_lib):
    from numbagg_amd.ragged import RaggedMoveMeanGroupMean

    with pytest.raises(ValueError, match="lds_labels"):
        RaggedMoveMeanGroupMean([100, 200], [5, 6], window=10, lds_labels=0)
    with pytest.raises(ValueError, match="lds_labels"):
        RaggedMoveMeanGroupMean([100, 200], [5, 6], window=10, lds_labels=100_000)


def test_model_choice_with_measured_tie_break(gpu_lib, monkeypatch):
    """The cost model ranks every LDS slot count; slot counts predicted within
    the tie margin are timed on the batch (with the model's strategies and
    with every series on its majority strategy) and the faster kept, so the
    choice is always one of those candidates."""
    from numbagg_amd import tuning
    from numbagg_amd.ragged import RaggedMoveMeanGroupMean

    lengths, ks, v, lab = make_batch(60, 100, 40_000, 10, 6000, seed=4)
    rb = RaggedMoveMeanGroupMean(lengths, ks, window=50, min_count=50)
    model = tuning.CostModel(seed=2)  # untrained: arbitrary but valid predictions
    skew, nanf = tuning.batch_statistics(rb, v, lab)
    ranked = model.rank_launches(tuning.features(rb.lengths, rb.num_labels, skew, nanf, rb.window), rb.lengths)
    assert sorted(c[1] for c in ranked) == sorted(tuning.CAPACITIES)
    assert all(a[0] <= b[0] for a, b in zip(ranked, ranked[1:]))
    for margin in (0.0, 1e9):  # no tie-break / every candidate timed
        monkeypatch.setattr(tuning, "TIE_MARGIN", margin)
        cap, st = tuning.choose_launch(rb, v, lab, method="model", model=model)
        close = ranked[:1] if margin == 0.0 else ranked
        cands = tuning.launch_candidates(close, rb.lengths)
        assert any(c == cap and np.array_equal(st, s) for c, s in cands)
        if margin == 0.0:
            assert cap == ranked[0][1]
        rb.set_strategy(st)
        rb.set_lds_labels(cap)
        mm, gm = rb(v, lab)
        pm, pg = rb.plain(v, lab)
        np.testing.assert_array_equal(_bits(mm), _bits(pm))


def test_missing_or_corrupt_model_falls_back_to_exhaustive(gpu_lib, monkeypatch, tmp_path):
    """A missing, unreadable or stale cost model gives CostModel.load() None and
    choose_launch(method="model") the exhaustive tuner's choice."""
    from numbagg_amd import tuning
    from numbagg_amd.ragged import RaggedMoveMeanGroupMean

    lengths, ks, v, lab = make_batch(12, 100, 20_000, 10, 3000, seed=5)
    rb = RaggedMoveMeanGroupMean(lengths, ks, window=50, min_count=50)
    skew, nanf = tuning.batch_statistics(rb, v, lab, sample=tuning.SAMPLE_PER_SERIES)
    want = tuning.exhaustive_launch(rb, v, lab, skew, nanf)
    corrupt = tmp_path / "corrupt.npz"
    corrupt.write_bytes(b"not a model")
    stale = tmp_path / "stale.npz"
    m = tuning.CostModel(seed=1)
    m.meta = {"arch": "gfx000", "kernel": "0", "configs": [list(c) for c in tuning.CONFIGS]}
    m.save(stale)
    calls = []
    real = tuning.exhaustive_launch

    def spy(*args, **kwargs):
        calls.append(1)
        return real(*args, **kwargs)

    monkeypatch.setattr(tuning, "exhaustive_launch", spy)
    for path in (tmp_path / "missing.npz", corrupt, stale):
        monkeypatch.setenv("NBG_COST_MODEL", str(path))
        assert tuning.CostModel.load() is None, path.name
        n0 = len(calls)
        cap, st = tuning.choose_launch(rb, v, lab, method="model")
        assert len(calls) == n0 + 1, path.name  # the exhaustive tuner decided
        assert cap in tuning.CAPACITIES and st.shape == (rb.nseries,)
        # a valid launch: the plain launches' results
        rb.set_strategy(st)
        rb.set_lds_labels(cap)
        mm, _ = rb(v, lab)
        pm, _ = rb.plain(v, lab)
        np.testing.assert_array_equal(_bits(mm), _bits(pm))
    assert want[0] in tuning.CAPACITIES
