"""The plain-kernel autotuner on the GPU: it times every candidate setting of a
family's knobs, records the fastest for the call's shape class, and the host
operators apply it; every setting keeps the parity bars (bit-exact functions
bit-exact)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_tune_records_and_applies(gpu_lib, tmp_path):
    import numbagg_amd as nb
    from numbagg_amd import plain_tuning
    from oracle import exact

    from ._data import runs_array

    db = plain_tuning.reload_database(tmp_path / "db.json")
    try:
        rng = np.random.default_rng(0)
        v = torch.from_numpy(rng.standard_normal((8, 300_000)).astype(np.float32)).cuda()
        lab = torch.from_numpy(rng.integers(0, 100, 300_000).astype(np.int32)).cuda()
        best, timings = plain_tuning.tune("group_nansum", lambda: nb.group_nansum(v, lab, num_labels=100, axis=-1),
                                          "float32", 8, 300_000, 100, reps=2, db=db)
        assert best in plain_tuning.candidates("group_sum")
        assert len(timings) == len(plain_tuning.candidates("group_sum"))
        key = plain_tuning.shape_key(plain_tuning.device_arch(), "group_nansum", "float32", 8, 300_000, 100)
        assert db.get(key) == best
        got = nb.group_nansum(v, lab, num_labels=100, axis=-1).cpu().numpy()  # the recorded setting applies
        vn, ln = v.cpu().numpy(), lab.cpu().numpy()
        want = exact.group("group_nansum", vn, ln, 100)
        sabs = exact.group("group_nansum", np.abs(vn), ln, 100)
        ulp = np.spacing(np.abs(want).astype(np.float32)).astype(np.float64)
        assert np.all(np.abs(got - want) <= ulp + 1e-14 * sabs)

        x = torch.from_numpy(runs_array(16, 70_000, seed=3)).cuda()
        want = exact.fill("ffill", x.cpu().numpy(), 9)
        for cand in plain_tuning.candidates("fill"):
            with plain_tuning.knobs_set(cand):
                got = nb.ffill(x, limit=9).cpu().numpy()
            np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
        best, _ = plain_tuning.tune("ffill", lambda: nb.ffill(x, limit=9), "float32", 16, 70_000, reps=2, db=db)
        np.testing.assert_array_equal(nb.ffill(x, limit=9).cpu().numpy().view(np.uint32), want.view(np.uint32))
        db.save()
        assert plain_tuning.PlainTuningDB(tmp_path / "db.json").get(key) == db.get(key)
    finally:
        plain_tuning.reload_database()


@pytest.mark.parametrize("func", ["move_mean", "move_std"])
def test_every_move_setting_gives_the_same_bits(gpu_lib, func):
    """Spans and prefetch change only which workgroup evaluates a tile."""
    import numbagg_amd as nb
    from numbagg_amd import plain_tuning

    from ._data import walk_array

    x = walk_array(5, 90_001, seed=4)
    outs = []
    for cand in plain_tuning.candidates("move"):
        with plain_tuning.knobs_set(cand):
            outs.append(getattr(nb, func)(x, window=700, min_count=300))
    for o in outs[1:]:
        np.testing.assert_array_equal(o.view(np.uint32), outs[0].view(np.uint32))
