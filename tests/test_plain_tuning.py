"""The plain-kernel autotuner's host side (CPU): thread-local knob overrides
through the C ABI (no device work), the database's versioning, and the
application of recorded settings around calls."""

from __future__ import annotations

import json
import threading

import pytest

from numbagg_amd import _build, _lib, plain_tuning


@pytest.fixture(scope="module")
def lib():
    if _build.library_is_stale():
        _build.build_library()
    return _lib.load()


def test_overrides_are_thread_local(lib, monkeypatch):
    monkeypatch.setenv("NBG_MOVE_SPAN", "5")
    assert lib.nbg_tune_get(b"NBG_MOVE_SPAN", 0) == 5  # the environment without an override
    seen = {}
    with plain_tuning.knobs_set({"NBG_MOVE_SPAN": 3, "NBG_GROUP_ROWS": 4}):
        assert lib.nbg_tune_get(b"NBG_MOVE_SPAN", 0) == 3
        assert lib.nbg_tune_get(b"NBG_GROUP_ROWS", 2) == 4

        def other():
            seen["span"] = lib.nbg_tune_get(b"NBG_MOVE_SPAN", 0)
            seen["rows"] = lib.nbg_tune_get(b"NBG_GROUP_ROWS", 2)

        t = threading.Thread(target=other)
        t.start()
        t.join()
    assert seen == {"span": 5, "rows": 2}  # another thread keeps the environment and defaults
    assert lib.nbg_tune_get(b"NBG_MOVE_SPAN", 0) == 5  # cleared on exit
    assert lib.nbg_tune_set(b"bad name", 1) == 1


def test_database_versioning_and_keys(tmp_path):
    key = plain_tuning.shape_key("gfx950", "move_mean", "float32", 256, 2_000_000, 1000)
    assert key == "gfx950|move_mean|float32|r8|n20|x9"
    path = tmp_path / "db.json"
    db = plain_tuning.PlainTuningDB(path)
    fill_key = plain_tuning.shape_key("gfx950", "ffill", "float64", 1, 1 << 24)
    db.put(key, {"NBG_MOVE_SPAN": 4}, {"{}": 1.0})
    db.put(fill_key, {"NBG_FILL_TWO_PASS": 0}, {"{}": 1.0})
    db.save()
    again = plain_tuning.PlainTuningDB(path)
    assert again.get(key) == {"NBG_MOVE_SPAN": 4}
    raw = json.loads(path.read_text())
    raw["kernel_versions"]["move"] = "stale"
    path.write_text(json.dumps(raw))
    stale = plain_tuning.PlainTuningDB(path)
    assert stale.get(key) is None  # the move kernels' sources changed: ignored
    assert stale.get(fill_key) == {"NBG_FILL_TWO_PASS": 0}  # other families keep theirs
    assert plain_tuning.kernel_version("group") == plain_tuning.kernel_version("group_sum")
    assert plain_tuning.kernel_version("move") != plain_tuning.kernel_version("fill")
    path.write_text("{corrupt")
    assert plain_tuning.PlainTuningDB(path).entries == {}


def test_candidates_and_families():
    assert plain_tuning.family_of("move_mean") == "move"
    assert plain_tuning.family_of("move_exp_nanvar") == "exp"
    assert plain_tuning.family_of("group_nansum") == "group_sum"
    assert plain_tuning.family_of("group_nanargmax") == "group"
    assert plain_tuning.family_of("ffill") == "fill"
    assert plain_tuning.family_of("move_corrmatrix") is None
    assert plain_tuning.family_of("move_exp_nancovmatrix") is None
    assert plain_tuning.family_of("nansum") is None
    assert len(plain_tuning.candidates("group_sum")) == 3 * 2 * 4
    assert {"NBG_FILL_TWO_PASS": 0} in plain_tuning.candidates("fill")


def test_applied_sets_recorded_knobs(lib, tmp_path, monkeypatch):
    monkeypatch.delenv("NBG_GROUP_ROWS", raising=False)
    db = plain_tuning.reload_database(tmp_path / "db.json")
    arch = plain_tuning.device_arch()
    db.put(plain_tuning.shape_key(arch, "group_nansum", "float32", 64, 4_000_000, 1000), {"NBG_GROUP_ROWS": 4}, {})
    try:
        with plain_tuning.applied("group_nansum", "float32", 64, 4_000_000, 1000):
            assert lib.nbg_tune_get(b"NBG_GROUP_ROWS", 2) == 4
        assert lib.nbg_tune_get(b"NBG_GROUP_ROWS", 2) == 2
        with plain_tuning.applied("group_nansum", "float32", 3, 100, 7):  # no entry: defaults
            assert lib.nbg_tune_get(b"NBG_GROUP_ROWS", 2) == 2
        with plain_tuning.knobs_set({"NBG_GROUP_ROWS": 1}):  # explicit overrides win
            with plain_tuning.applied("group_nansum", "float32", 64, 4_000_000, 1000):
                assert lib.nbg_tune_get(b"NBG_GROUP_ROWS", 2) == 1
    finally:
        plain_tuning.reload_database()
