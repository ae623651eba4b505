"""Ragged batched serving path: move_mean + group_nanmean over many series.

Thousands of float32 series of uneven lengths, each with its own int32 labels
and label count, are processed by one persistent HIP launch
(``nbg_ragged_move_mean_group_nanmean``) instead of two plain calls per series.
The reference semantics are the per-series plain calls
``move_mean(series, window=w, min_count=m)`` (numbagg/moving.py:12-55) and
``group_nanmean(series, labels, num_labels=K)`` (numbagg/grouped.py:7-27);
move_mean outputs are bitwise equal to them.

The per-series group-by strategy comes from ``strategy``: STRATEGY_LDS keeps
the accumulators of labels below the launch's ``lds_labels`` slots in LDS
(all labels when the series has no more; the rest go to global atomics),
STRATEGY_GLOBAL sends every label to global atomics. ``lds_labels``
(default LDS_MAX_LABELS = 1024) trades LDS slots against resident workgroups. "heuristic" (LDS when the labels fit, global
otherwise), "lds", "global", or an explicit per-series array (what the
autotuner or the cost model in ``numbagg_amd.tuning`` produce) select them. The launch sequence can be
captured once in a HIP graph and replayed (``graph``).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _device as dev
from . import _lib

LDS_MAX_LABELS = 1024
STRATEGY_LDS = 0
STRATEGY_GLOBAL = 1


def heuristic_strategy(lengths: np.ndarray, num_labels: np.ndarray) -> np.ndarray:
    """The label-count rule: LDS-privatised accumulators when all of a
    series' labels fit the default LDS slots, global atomics otherwise."""
    k = np.asarray(num_labels)
    return np.where(k <= LDS_MAX_LABELS, STRATEGY_LDS, STRATEGY_GLOBAL).astype(np.uint8)


class RaggedMoveMeanGroupMean:
    """move_mean(window, min_count) and group_nanmean for ragged float32 series.

    lengths: (S,) series lengths; num_labels: (S,) label counts.
    Call with device tensors ``values`` (sum(lengths),) float32 and ``labels``
    (sum(lengths),) int32; returns (move_out, group_out) where group_out has
    sum(num_labels) entries, series s at label_offsets[s]:label_offsets[s+1].
    """

    def __init__(self, lengths, num_labels, *, window: int, min_count: int | None = None, strategy="heuristic",
                 lds_labels: int = LDS_MAX_LABELS, device=None):
        dev.require_gpu()
        lengths = np.asarray(lengths, dtype=np.int64)
        num_labels = np.asarray(num_labels, dtype=np.int64)
        if lengths.ndim != 1 or lengths.shape != num_labels.shape:
            raise ValueError("lengths and num_labels must be 1-D of the same length")
        if np.any(lengths <= 0):
            raise ValueError("every series needs at least one element")
        if np.any(num_labels < 0):
            raise ValueError("num_labels must be non-negative")
        if min_count is None:
            min_count = window
        if min_count < 0:
            raise ValueError(f"min_count must be positive: {min_count}")
        if not 0 < window <= int(lengths.min()):
            raise ValueError(f"window not in valid range: {window}")
        self.window, self.min_count = int(window), int(min_count)
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.lengths, self.num_labels = lengths, num_labels
        self.nseries = int(lengths.size)
        self.offsets = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
        self.label_offsets = np.concatenate([[0], np.cumsum(num_labels)]).astype(np.int64)
        self.tile = int(_lib.query("nbg_ragged_tile_size"))
        tiles = (lengths + self.tile - 1) // self.tile
        self.tile_offsets = np.concatenate([[0], np.cumsum(tiles)]).astype(np.int64)
        self.total = int(self.offsets[-1])
        self.total_labels = int(self.label_offsets[-1])
        self.total_tiles = int(self.tile_offsets[-1])
        if self.total_tiles >= 2**31 - 2**20:
            raise ValueError("too many tiles for one ragged launch")
        tile_series = np.repeat(np.arange(self.nseries, dtype=np.int32), tiles)
        self.set_strategy(strategy)
        self.set_lds_labels(lds_labels)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(self.device)  # noqa: E731
        self._d_offsets = t(self.offsets)
        self._d_label_offsets = t(self.label_offsets)
        self._d_tile_offsets = t(self.tile_offsets)
        self._d_tile_series = t(tile_series)
        ws = _lib.query("nbg_ragged_workspace_bytes", self.nseries, self.total_labels)
        self._ws = torch.empty(int(ws), dtype=torch.uint8, device=self.device)

    def set_lds_labels(self, lds_labels: int) -> None:
        """LDS label slots of the launch (labels below it are privatised for
        STRATEGY_LDS series)."""
        cap = int(_lib.query("nbg_ragged_max_lds_labels"))
        if not 1 <= int(lds_labels) <= cap:
            raise ValueError(f"lds_labels must be in [1, {cap}]: {lds_labels}")
        self.lds_labels = int(lds_labels)

    def set_strategy(self, strategy) -> None:
        if isinstance(strategy, str):
            if strategy == "heuristic":
                arr = heuristic_strategy(self.lengths, self.num_labels)
            elif strategy == "lds":
                arr = np.full(self.nseries, STRATEGY_LDS, dtype=np.uint8)
            elif strategy == "global":
                arr = np.full(self.nseries, STRATEGY_GLOBAL, dtype=np.uint8)
            else:
                raise ValueError(f"unknown strategy {strategy!r}")
        else:
            arr = np.asarray(strategy, dtype=np.uint8)
            if arr.shape != (self.nseries,):
                raise ValueError("strategy array must have one entry per series")
        if np.any(arr > STRATEGY_GLOBAL):
            raise ValueError("strategy entries must be STRATEGY_LDS (0) or STRATEGY_GLOBAL (1)")
        self.strategy = arr
        self._d_strategy = torch.from_numpy(arr.copy()).to(self.device if hasattr(self, "device") else "cuda")

    def __call__(self, values: torch.Tensor, labels: torch.Tensor, move_out=None, group_out=None):
        if values.dtype != torch.float32 or labels.dtype != torch.int32:
            raise TypeError("values must be float32 and labels int32 device tensors")
        if values.numel() != self.total or labels.numel() != self.total:
            raise ValueError("values/labels length must equal sum(lengths)")
        values = values.contiguous()
        labels = labels.contiguous()
        if move_out is None:
            move_out = torch.empty_like(values)
        if group_out is None:
            group_out = torch.empty(self.total_labels, dtype=torch.float32, device=values.device)
        _lib.call(
            "nbg_ragged_move_mean_group_nanmean",
            dev.ptr(values), dev.ptr(labels), dev.ptr(self._d_offsets), self.nseries,
            dev.ptr(self._d_label_offsets), self.total_labels, dev.ptr(self._d_tile_offsets),
            dev.ptr(self._d_tile_series), self.total_tiles, dev.ptr(self._d_strategy), self.lds_labels,
            self.window, self.min_count, dev.ptr(move_out), dev.ptr(group_out),
            dev.ptr(self._ws), self._ws.numel(), dev.stream_handle(),
        )
        return move_out, group_out

    def graph(self, values: torch.Tensor, labels: torch.Tensor):
        """Capture the launch sequence in a HIP graph; returns (replay, move_out,
        group_out). Replays read `values`/`labels` in place."""
        move_out = torch.empty_like(values)
        group_out = torch.empty(self.total_labels, dtype=torch.float32, device=values.device)
        self(values, labels, move_out, group_out)  # warm-up outside capture
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self(values, labels, move_out, group_out)
        return g.replay, move_out, group_out

    def plain(self, values: torch.Tensor, labels: torch.Tensor):
        """Reference path: per-series plain calls through the C ABI (two calls
        per series, on views of the same device buffers)."""
        move_out = torch.empty_like(values)
        group_out = torch.empty(self.total_labels, dtype=torch.float32, device=values.device)
        gws = None
        stream = dev.stream_handle()
        esz = 4
        for s in range(self.nseries):
            o, n = int(self.offsets[s]), int(self.lengths[s])
            lo, k = int(self.label_offsets[s]), int(self.num_labels[s])
            _lib.call(
                "nbg_move_mean", ctypes.c_int(_lib.NBG_F32), ctypes.c_void_p(values.data_ptr() + o * esz),
                ctypes.c_void_p(move_out.data_ptr() + o * esz), 1, n, n, n, self.window, self.min_count, stream,
            )
            if k == 0:
                continue
            need = _lib.query("nbg_group_workspace_bytes", _lib.NBG_F32, 1, n, k, 0)
            if gws is None or gws.numel() < need:
                gws = torch.empty(int(need), dtype=torch.uint8, device=values.device)
            _lib.call(
                "nbg_group_nanmean", ctypes.c_int(_lib.NBG_F32), ctypes.c_void_p(values.data_ptr() + o * esz),
                ctypes.c_int(_lib.NBG_I32), ctypes.c_void_p(labels.data_ptr() + o * 4),
                ctypes.c_void_p(group_out.data_ptr() + lo * esz), 1, n, n, k, ctypes.c_int(0),
                dev.ptr(gws), gws.numel(), stream,
            )
        return move_out, group_out
