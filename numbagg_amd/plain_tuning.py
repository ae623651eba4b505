"""Autotuner for the plain kernels: tile spans, rows per workgroup, vectors
per thread, workgroups per CU and path choices, per call shape.

Each family of kernels exposes tuning knobs (``NBG_*`` names, read by the C
library through ``nbg_tune_set`` overrides of the calling thread, else the
environment, else the kernel's default). ``tune`` times every candidate
setting of a family's knobs on the given call with HIP events and records
the fastest in a JSON database keyed by GPU architecture, the function,
dtype and the call's shape class (log2 rows and length, and the window or
label count), and versioned per family by a hash of that family's kernel
sources. The host operators apply the recorded setting of a call's shape
class around the C call (``applied``); shapes without an entry run the
kernels' defaults. Entries of a family whose sources changed since they were
measured are ignored, as are entries of another architecture.

Knob settings change launch shapes only: results stay within the parity
bars (bit-exact functions stay bit-exact; float sums and moments may differ
in the last place where workgroups merge partials in another order).
"""

from __future__ import annotations

import contextlib
import hashlib
import itertools
import json
import math
import pathlib
import threading

from . import _lib

DATA_DIR = pathlib.Path(__file__).resolve().parent / "tuning_data"
DB_PATH = DATA_DIR / "plain_tuning_db.json"

# Candidate values per knob, by kernel family (0 for NBG_MOVE_SPAN: the
# launch's own choice from the grid size).
FAMILIES: dict[str, dict[str, list[int]]] = {
    "move": {"NBG_MOVE_SPAN": [0, 2, 3, 4, 6, 8], "NBG_MOVE_PREFETCH": [1, 0]},
    "group_sum": {
        "NBG_GROUP_ROWS": [1, 2, 4],
        "NBG_GROUP_UNROLL": [2, 4],
        "NBG_GROUP_BLOCKS_PER_CU": [2, 3, 4, 8],
    },
    "group": {"NBG_GROUP_BLOCKS_PER_CU": [2, 3, 4, 8]},
    "fill": {"NBG_FILL_TWO_PASS": [1, 0]},
    "exp": {"NBG_EXP_TWO_PASS": [0, 1]},
}
# Functions whose main pass is the grouped sum pass (its shape knobs apply).
GROUP_SUM_FUNCS = {"group_nansum", "group_nanmean"}


def family_of(func: str) -> str | None:
    if func.endswith("matrix"):  # the matrix kernels have no tuning knobs here
        return None
    if func.startswith("move_exp_"):
        return "exp"
    if func.startswith("move_"):
        return "move"
    if func in GROUP_SUM_FUNCS:
        return "group_sum"
    if func.startswith("group_"):
        return "group"
    if func in ("ffill", "bfill"):
        return "fill"
    return None


# The sources each family's kernels are built from.
FAMILY_SOURCES: dict[str, tuple[str, ...]] = {
    "move": ("moving.hip", "rolling.hpp", "common.hpp"),
    "group_sum": ("grouped.hip", "common.hpp"),
    "group": ("grouped.hip", "common.hpp"),
    "fill": ("fill.hip", "common.hpp"),
    "exp": ("moving_exp.hip", "common.hpp"),
}


def kernel_version(family: str) -> str:
    """Hash of the sources of ``family``'s kernels."""
    h = hashlib.sha256()
    csrc = pathlib.Path(__file__).resolve().parent / "csrc"
    for name in FAMILY_SOURCES[family]:
        h.update(name.encode())
        h.update((csrc / name).read_bytes())
    return h.hexdigest()[:16]


def kernel_versions() -> dict[str, str]:
    return {fam: kernel_version(fam) for fam in FAMILY_SOURCES}


def _func_of(key: str) -> str:
    parts = key.split("|")
    return parts[1] if len(parts) > 1 else ""


def _bucket(v: int) -> int:
    return int(math.floor(math.log2(v))) if v > 0 else -1


def shape_key(arch: str, func: str, dtype: str, rows: int, n: int, extra: int = 0) -> str:
    """Shape class of a call: log2 buckets of rows, length and the window or
    label count."""
    return f"{arch}|{func}|{dtype}|r{_bucket(rows)}|n{_bucket(n)}|x{_bucket(extra)}"


class PlainTuningDB:
    """The recorded settings (JSON: {"kernel_versions": {family: hash},
    "entries": {key: {...}}})."""

    def __init__(self, path: pathlib.Path | str = DB_PATH) -> None:
        self.path = pathlib.Path(path)
        self.versions = kernel_versions()
        self.entries: dict[str, dict] = {}
        self._lock = threading.Lock()
        try:
            with open(self.path) as f:
                raw = json.load(f)
        except (OSError, ValueError):
            raw = {}
        if not isinstance(raw, dict):
            raw = {}
        recorded = raw.get("kernel_versions") or {}
        for key, rec in dict(raw.get("entries") or {}).items():
            fam = family_of(_func_of(key))
            if fam is not None and recorded.get(fam) == self.versions[fam]:
                self.entries[key] = rec

    def get(self, key: str) -> dict | None:
        rec = self.entries.get(key)
        return None if rec is None else rec.get("knobs")

    def put(self, key: str, knobs: dict, timings: dict) -> None:
        with self._lock:
            self.entries[key] = {"knobs": knobs, "timings_ms": timings}

    def save(self) -> None:
        self.path.parent.mkdir(parents=True, exist_ok=True)
        with open(self.path, "w") as f:
            json.dump({"kernel_versions": self.versions, "entries": self.entries}, f, indent=1, sort_keys=True)


_db: PlainTuningDB | None = None
_db_lock = threading.Lock()
_arch: str | None = None


def database() -> PlainTuningDB:
    global _db
    if _db is None:
        with _db_lock:
            if _db is None:
                _db = PlainTuningDB()
    return _db


def reload_database(path: pathlib.Path | str = DB_PATH) -> PlainTuningDB:
    global _db
    with _db_lock:
        _db = PlainTuningDB(path)
    return _db


def device_arch() -> str:
    global _arch
    if _arch is None:
        import torch

        try:
            _arch = torch.cuda.get_device_properties(torch.cuda.current_device()).gcnArchName.split(":")[0]
        except Exception:
            _arch = "unknown"
    return _arch


def set_knobs(knobs: dict) -> None:
    lib = _lib.load()
    for name, value in knobs.items():
        if lib.nbg_tune_set(name.encode(), int(value)) != 0:
            raise RuntimeError(_lib.last_error())


def clear_knobs() -> None:
    _lib.load().nbg_tune_clear()


_local = threading.local()


@contextlib.contextmanager
def knobs_set(knobs: dict):
    """Set this thread's knob overrides for the block (cleared afterwards).
    Calls made inside keep these overrides: the recorded settings are not
    applied over them."""
    set_knobs(knobs)
    _local.explicit = getattr(_local, "explicit", 0) + 1
    try:
        yield
    finally:
        _local.explicit -= 1
        clear_knobs()


@contextlib.contextmanager
def applied(func: str, dtype, rows: int, n: int, extra: int = 0):
    """Apply the recorded setting of this call's shape class, if any (unless
    explicit overrides are in force on this thread)."""
    db = database()
    if not db.entries or getattr(_local, "explicit", 0):
        yield
        return
    knobs = db.get(shape_key(device_arch(), func, str(dtype), rows, n, extra))
    if not knobs:
        yield
        return
    set_knobs(knobs)
    try:
        yield
    finally:
        clear_knobs()


def candidates(family: str) -> list[dict]:
    space = FAMILIES[family]
    names = list(space)
    return [dict(zip(names, vals)) for vals in itertools.product(*(space[k] for k in names))]


def tune(func: str, call, dtype, rows: int, n: int, extra: int = 0, reps: int = 5,
         db: PlainTuningDB | None = None) -> tuple[dict, dict]:
    """Time ``call()`` (one call of ``func`` on device data) under every
    candidate setting of its family and record the fastest.

    Returns (chosen knobs, {setting: ms}). Timings are the minimum over three
    rounds of ``reps`` back-to-back calls between HIP events (host overhead
    hidden behind the queued launches).
    """
    import torch

    family = family_of(func)
    if family is None:
        raise ValueError(f"no tuning knobs for {func}")
    db = db or database()
    stream = torch.cuda.current_stream()
    timings: dict[str, float] = {}
    best, best_ms = {}, math.inf
    for cand in candidates(family):
        with knobs_set(cand):
            call()
            torch.cuda.synchronize()
            ms = math.inf
            for _ in range(3):
                a = torch.cuda.Event(enable_timing=True)
                b = torch.cuda.Event(enable_timing=True)
                a.record(stream)
                for _ in range(reps):
                    call()
                b.record(stream)
                torch.cuda.synchronize()
                ms = min(ms, a.elapsed_time(b) / reps)
        timings[json.dumps(cand, sort_keys=True)] = round(ms, 5)
        if ms < best_ms:
            best, best_ms = cand, ms
    db.put(shape_key(device_arch(), func, str(dtype), rows, n, extra), best, timings)
    return best, timings


__all__ = [
    "FAMILIES",
    "PlainTuningDB",
    "applied",
    "candidates",
    "clear_knobs",
    "database",
    "family_of",
    "kernel_version",
    "kernel_versions",
    "knobs_set",
    "reload_database",
    "set_knobs",
    "shape_key",
    "tune",
]
