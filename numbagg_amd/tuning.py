"""Strategy selection for the ragged batched path: autotuner and cost model.

For each series of a ragged batch the kernel can fold values into its label
accumulators through LDS-privatised slots for the labels below the launch's
LDS slot count (flushed when the workgroup moves to another series; larger
labels still go to global atomics) or through global atomics only. Which is faster depends on
the series' length, label count, label skew (how much of the mass the low
labels carry, and how often lanes collide on one address), NaN fraction (and
window, through the tile's rolling work).

* ``Autotuner`` measures every launch configuration exhaustively: (LDS label
  slots in CAPACITIES) x (strategy), on homogeneous synthetic batches (one
  feature point each), timed with HIP events; ns/element go to a JSON timing
  DB versioned by GPU architecture and a hash of the kernel source. More
  slots keep more labels out of global atomics but leave fewer workgroups
  resident, which slows the rolling part: neither wins everywhere.
* ``CostModel`` is a 5 -> 64 -> 64 -> 4 fp32 MLP predicting log ns/element of
  each configuration from (log2 length, log2 labels, skew, NaN fraction, log2
  window). It is trained on the GPU by this library's own kernels
  (nbg_dense_forward / nbg_dense_backward / nbg_mse_grad / nbg_adam_step) on
  the autotuner's timings; inference uses nbg_dense_forward.
* ``choose_launch`` picks the launch's LDS label slots and per-series
  strategies with the cost model (slot counts whose predicted times are
  within ``TIE_MARGIN`` (2x) of the best are timed once on the batch itself,
  each with the model's per-series strategies and with every series on the
  model's majority strategy), the exhaustive tuner (the fallback, and the
  reference point of the model's regret) or a label-count heuristic.
"""

from __future__ import annotations

import ctypes
import hashlib
import json
import math
import pathlib
import types

import numpy as np
import torch

from . import _device as dev
from . import _lib
from .ragged import LDS_MAX_LABELS, STRATEGY_GLOBAL, STRATEGY_LDS, RaggedMoveMeanGroupMean

DATA_DIR = pathlib.Path(__file__).resolve().parent / "tuning_data"
DB_PATH = DATA_DIR / "ragged_timing_db.json"
MODEL_PATH = DATA_DIR / "ragged_cost_model.npz"
N_FEATURES = 5
STRATEGIES = (STRATEGY_LDS, STRATEGY_GLOBAL)
# LDS label slots per launch the tuner considers, and the launch
# configurations it times: (slots, strategy), in the cost model's output order.
CAPACITIES = (LDS_MAX_LABELS, 4096)
CONFIGS = tuple((c, st) for c in CAPACITIES for st in STRATEGIES)
# Slot counts whose predicted time is within this fraction of the best are
# settled by timing each once on the batch (two launches). The model is
# trained on homogeneous batches and misranks slot counts on mixed ones
# (measured on MI355X, 20,000 mixed series of 2.2e9 elements: it predicted
# 1024 slots best, which took 15.4 ms against 9.9 ms at 4096), so both are
# timed unless one is predicted twice as fast. The model keeps the per-series
# strategies, the combinatorial part of the choice.
TIE_MARGIN = 1.0


def kernel_version() -> str:
    """Hash of the sources whose timings the DB and the model describe."""
    h = hashlib.sha256()
    csrc = pathlib.Path(__file__).resolve().parent / "csrc"
    for name in ("ragged.hip", "rolling.hpp", "common.hpp"):
        h.update((csrc / name).read_bytes())
    return h.hexdigest()[:16]


def device_arch() -> str:
    try:
        return torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]
    except Exception:  # pragma: no cover - no device
        return "unknown"


def features(lengths, num_labels, skew, nan_fraction, window) -> np.ndarray:
    """(S, 5) float32 feature matrix."""
    lengths = np.asarray(lengths, dtype=np.float64)
    num_labels = np.asarray(num_labels, dtype=np.float64)
    s = np.broadcast_to(np.asarray(skew, dtype=np.float64), lengths.shape)
    f = np.broadcast_to(np.asarray(nan_fraction, dtype=np.float64), lengths.shape)
    w = np.broadcast_to(np.asarray(window, dtype=np.float64), lengths.shape)
    return np.stack(
        [np.log2(lengths), np.log2(np.maximum(num_labels, 1)), s, f, np.log2(np.maximum(w, 1))], axis=1
    ).astype(np.float32)


def series_chunks(lengths: np.ndarray, max_elements: int = 1 << 27):
    """Consecutive series ranges [s0, s1) holding at most `max_elements`
    elements each (a single longer series gets a range of its own), so that
    per-element temporaries stay far below 2**31 elements."""
    s0, acc = 0, 0
    for i, n in enumerate(np.asarray(lengths, dtype=np.int64)):
        if acc and acc + int(n) > max_elements:
            yield s0, i
            s0, acc = i, 0
        acc += int(n)
    if s0 < len(lengths):
        yield s0, len(lengths)


# Elements per series the launch choice reads for its statistics (a prefix of
# each series): the features need label skew and NaN share, not every element.
SAMPLE_PER_SERIES = 4096


def _prefix_sample(rb, values: torch.Tensor, labels: torch.Tensor, sample: int):
    """(batch view, values, labels) of the first min(length, sample) elements
    of every series."""
    slen = np.minimum(np.asarray(rb.lengths, dtype=np.int64), sample)
    offs = np.concatenate([[0], np.cumsum(slen)]).astype(np.int64)
    dev_ = values.device
    n = int(offs[-1])
    sl = torch.from_numpy(slen).to(dev_)
    base = torch.repeat_interleave(torch.from_numpy(np.asarray(rb.offsets[:-1], dtype=np.int64)).to(dev_), sl)
    within = torch.arange(n, device=dev_) - torch.repeat_interleave(torch.from_numpy(offs[:-1]).to(dev_), sl)
    idx = base + within
    view = types.SimpleNamespace(nseries=rb.nseries, lengths=slen, offsets=offs, num_labels=rb.num_labels,
                                 label_offsets=rb.label_offsets)
    return view, values[idx], labels[idx]


def batch_statistics(rb: RaggedMoveMeanGroupMean, values: torch.Tensor, labels: torch.Tensor,
                     max_elements: int = 1 << 27, sample: int | None = None):
    """Per-series skew (share of the most frequent label) and NaN fraction,
    computed over chunks of whole series; with ``sample``, over the first
    ``sample`` elements of each series."""
    if sample is not None:
        rb, values, labels = _prefix_sample(rb, values, labels, sample)
    dev_ = values.device
    skew = np.zeros(rb.nseries)
    nanf = np.zeros(rb.nseries)
    for s0, s1 in series_chunks(rb.lengths, max_elements):
        o0, o1 = int(rb.offsets[s0]), int(rb.offsets[s1])
        if o1 == o0:
            continue
        lens = torch.from_numpy(rb.lengths[s0:s1]).to(dev_)
        seg = torch.repeat_interleave(torch.arange(s1 - s0, device=dev_), lens)
        k = torch.from_numpy(rb.num_labels[s0:s1]).to(dev_)
        loff = torch.from_numpy(rb.label_offsets[s0:s1] - rb.label_offsets[s0]).to(dev_)
        nlab = int(rb.label_offsets[s1] - rb.label_offsets[s0])
        lab = labels[o0:o1].long()
        v = values[o0:o1]
        ok = (lab >= 0) & (lab < k[seg])
        glab = (loff[seg] + lab.clamp(min=0))[ok]
        counts = torch.bincount(glab, minlength=nlab).float()
        if nlab:
            top = torch.segment_reduce(counts, "max", lengths=k, unsafe=True)
            top = torch.nan_to_num(top, nan=0.0, neginf=0.0)
        else:
            top = torch.zeros(s1 - s0, device=dev_)
        L = lens.float()
        skew[s0:s1] = (top / L).cpu().numpy()
        nan_cnt = torch.zeros(s1 - s0, device=dev_).index_add_(0, seg, torch.isnan(v).float())
        nanf[s0:s1] = (nan_cnt / L).cpu().numpy()
        del seg, lab, v, ok, glab, counts
    return skew, nanf


def synthetic_labels(length: int, k: int, zipf: float, gen: torch.Generator, device) -> torch.Tensor:
    """int32 labels over k classes: uniform (zipf == 0) or Zipf(zipf)-like via a
    continuous inverse CDF."""
    u = torch.rand(length, generator=gen, device=device, dtype=torch.float64)
    if zipf == 0:
        return torch.clamp((u * k).long(), max=k - 1).to(torch.int32)
    a = zipf - 1.0
    x = (1.0 - u * (1.0 - float(k) ** -a)) ** (-1.0 / a)
    return torch.clamp(torch.floor(x).long() - 1, 0, k - 1).to(torch.int32)


class Autotuner:
    """Exhaustive strategy timing of the ragged kernel, cached in a JSON DB."""

    def __init__(self, db_path: pathlib.Path | str = DB_PATH, target_elements: int = 1 << 25, reps: int = 3,
                 max_labels: int = 1 << 26):
        self.db_path = pathlib.Path(db_path)
        # Batches of about target_elements (long enough for steady-state
        # timings) with at most max_labels labels in total.
        self.target_elements = target_elements
        self.max_labels = max_labels
        self.reps = reps
        self.meta = {"arch": device_arch(), "kernel": kernel_version(), "configs": [list(c) for c in CONFIGS]}
        self.entries: list[dict] = []
        if self.db_path.exists():
            try:
                db = json.loads(self.db_path.read_text())
                if db.get("meta") == self.meta:
                    self.entries = db.get("entries", [])
            except (OSError, ValueError):
                self.entries = []

    def save(self) -> None:
        self.db_path.parent.mkdir(parents=True, exist_ok=True)
        self.db_path.write_text(json.dumps({"meta": self.meta, "entries": self.entries}, indent=1))

    def _time(self, rb, values, labels, strategy, lds_labels: int = LDS_MAX_LABELS, move_out=None,
              group_out=None) -> float:
        """ns per element of one launch; `strategy` is one value or a per-series
        array. Outputs go to `move_out` / `group_out` when given (a large batch's
        fresh output allocation costs more than the launches timed)."""
        st = np.broadcast_to(np.asarray(strategy, dtype=np.uint8), (rb.nseries,))
        rb.set_strategy(np.ascontiguousarray(st))
        rb.set_lds_labels(lds_labels)
        mo = move_out if move_out is not None else torch.empty_like(values)
        go = group_out if group_out is not None else torch.empty(rb.total_labels, dtype=torch.float32,
                                                                 device=values.device)
        rb(values, labels, mo, go)
        stream = torch.cuda.current_stream()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(self.reps):
            rb(values, labels, mo, go)
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) * 1e6 / self.reps / values.numel()  # ns per element

    def measure(self, length: int, k: int, zipf: float, nan: float, window: int, seed: int = 0) -> dict:
        """Time every launch configuration on a homogeneous batch for one
        feature point."""
        dev.require_gpu()
        device = torch.device("cuda")
        gen = torch.Generator(device=device).manual_seed(seed)
        s = max(1, min(self.target_elements // length, self.max_labels // max(k, 1)))
        lengths = np.full(s, length, dtype=np.int64)
        ks = np.full(s, k, dtype=np.int64)
        total = int(lengths.sum())
        values = torch.randn(total, generator=gen, device=device, dtype=torch.float32)
        values[torch.rand(total, generator=gen, device=device) < nan] = float("nan")
        labels = synthetic_labels(total, k, zipf, gen, device)
        rb = RaggedMoveMeanGroupMean(lengths, ks, window=min(window, length), min_count=1)
        skew, nanf = batch_statistics(rb, values, labels)
        ns = [self._time(rb, values, labels, st, cap) for cap, st in CONFIGS]
        entry = {
            "length": length, "num_labels": k, "zipf": zipf, "nan": nan, "window": window,
            "skew": float(np.mean(skew)), "nan_fraction": float(np.mean(nanf)), "ns": ns,
        }
        self.entries.append(entry)
        return entry

    def run_grid(self, lengths=None, ks=None, zipfs=(0.0, 1.1, 2.0), nans=(0.05, 0.4), windows=(20, 200), seed=0,
                 progress=None):
        """Measure every grid point; ``progress(i, entry)`` after each."""
        lengths = lengths or [2**e for e in range(7, 21, 2)]
        ks = ks or [8, 32, 128, 512, 1024, 4096, 16384, 65536]
        i = 0
        for L in lengths:
            for k in ks:
                for z in zipfs:
                    for f in nans:
                        for w in windows:
                            entry = self.measure(L, k, z, f, w, seed=seed + i)
                            i += 1
                            if progress is not None:
                                progress(i, entry)
        return self.entries

    def training_set(self):
        """Features (n, 5), targets (n, len(CONFIGS)) = log ns/element per
        configuration, and the points' label counts."""
        rows = [e for e in self.entries if all(math.isfinite(v) for v in e["ns"])]
        X = features([e["length"] for e in rows], [e["num_labels"] for e in rows], [e["skew"] for e in rows],
                     [e["nan_fraction"] for e in rows], [e["window"] for e in rows])
        T = np.log(np.array([e["ns"] for e in rows], dtype=np.float64)).astype(np.float32)
        return X, T, np.array([e["num_labels"] for e in rows])


class CostModel:
    """5 -> 64 -> 64 -> n_out fp32 MLP on this library's HIP kernels; by
    default one output (log ns/element) per launch configuration in CONFIGS."""

    def __init__(self, seed: int = 0, device=None, n_out: int = len(CONFIGS)):
        dev.require_gpu()
        self.DIMS = (N_FEATURES, 64, 64, int(n_out))
        self.device = torch.device(device or "cuda")
        rng = np.random.default_rng(seed)
        self.params: list[torch.Tensor] = []
        for fan_in, fan_out in zip(self.DIMS[:-1], self.DIMS[1:]):
            w = rng.standard_normal((fan_out, fan_in)) * math.sqrt(2.0 / fan_in)
            self.params += [torch.tensor(w, dtype=torch.float32, device=self.device),
                            torch.zeros(fan_out, dtype=torch.float32, device=self.device)]
        self.x_mean = np.zeros(N_FEATURES, np.float32)
        self.x_std = np.ones(N_FEATURES, np.float32)
        self.t_mean = np.zeros(self.DIMS[-1], np.float32)
        self.t_std = np.ones(self.DIMS[-1], np.float32)
        self.meta: dict = {}
        self._adam = None
        self.step = 0

    # ---- kernels -----------------------------------------------------------
    def _forward(self, x: torch.Tensor):
        acts = [x]
        h = x
        nl = len(self.DIMS) - 1
        for li in range(nl):
            w, b = self.params[2 * li], self.params[2 * li + 1]
            y = torch.empty((h.shape[0], w.shape[0]), dtype=torch.float32, device=self.device)
            act = 1 if li < nl - 1 else 0
            _lib.call("nbg_dense_forward", dev.ptr(h), dev.ptr(w), dev.ptr(b), dev.ptr(y), h.shape[0],
                      w.shape[1], w.shape[0], act, dev.stream_handle())
            acts.append(y)
            h = y
        return acts

    def _backward(self, acts, dy: torch.Tensor):
        grads = [None] * len(self.params)
        nl = len(self.DIMS) - 1
        for li in reversed(range(nl)):
            w = self.params[2 * li]
            x, y = acts[li], acts[li + 1]
            dw = torch.empty_like(w)
            db = torch.empty_like(self.params[2 * li + 1])
            dz = torch.empty_like(y)
            dx = torch.empty_like(x) if li > 0 else None
            act = 1 if li < nl - 1 else 0
            _lib.call("nbg_dense_backward", dev.ptr(x), dev.ptr(w), dev.ptr(y), dev.ptr(dy), dev.ptr(dx),
                      dev.ptr(dw), dev.ptr(db), dev.ptr(dz), x.shape[0], w.shape[1], w.shape[0], act,
                      dev.stream_handle())
            grads[2 * li], grads[2 * li + 1] = dw, db
            dy = dx
        return grads

    def loss_and_grads(self, x: torch.Tensor, t: torch.Tensor):
        acts = self._forward(x)
        y = acts[-1]
        dy = torch.empty_like(y)
        loss = torch.zeros(1, dtype=torch.float32, device=self.device)
        _lib.call("nbg_mse_grad", dev.ptr(y), dev.ptr(t), dev.ptr(dy), dev.ptr(loss), y.numel(), dev.stream_handle())
        return loss, self._backward(acts, dy)

    def _adam_step(self, grads, lr: float):
        if self._adam is None:
            self._adam = [(torch.zeros_like(p), torch.zeros_like(p)) for p in self.params]
        self.step += 1
        for p, g, (m, v) in zip(self.params, grads, self._adam):
            _lib.call("nbg_adam_step", dev.ptr(p), dev.ptr(g), dev.ptr(m), dev.ptr(v), p.numel(),
                      ctypes.c_float(lr), ctypes.c_float(0.9), ctypes.c_float(0.999), ctypes.c_float(1e-8),
                      self.step, dev.stream_handle())

    # ---- training / inference -----------------------------------------------
    def fit(self, X: np.ndarray, T: np.ndarray, epochs: int = 2000, lr: float = 3e-3, seed: int = 0) -> list[float]:
        X = np.asarray(X, np.float32)
        T = np.asarray(T, np.float32)
        self.x_mean, self.x_std = X.mean(0), X.std(0) + 1e-6
        self.t_mean, self.t_std = T.mean(0), T.std(0) + 1e-6
        x = torch.from_numpy((X - self.x_mean) / self.x_std).to(self.device)
        t = torch.from_numpy((T - self.t_mean) / self.t_std).to(self.device)
        history = []
        for ep in range(epochs):
            loss, grads = self.loss_and_grads(x, t)
            self._adam_step(grads, lr)
            if ep % 100 == 0 or ep == epochs - 1:
                history.append(float(loss.item()))
        return history

    def predict(self, X: np.ndarray) -> np.ndarray:
        """Predicted targets (log ns/element per output), shape (n, n_out)."""
        X = np.asarray(X, np.float32)
        x = torch.from_numpy((X - self.x_mean) / self.x_std).to(self.device)
        y = self._forward(x)[-1].cpu().numpy()
        return y * self.t_std + self.t_mean

    def choose(self, X: np.ndarray, num_labels: np.ndarray | None = None) -> np.ndarray:
        """Index of the output predicted fastest, per row (uint8)."""
        return np.argmin(self.predict(X), axis=1).astype(np.uint8)

    def rank_launches(self, X: np.ndarray, lengths: np.ndarray) -> list[tuple[float, int, np.ndarray]]:
        """(predicted time, LDS label slots, per-series strategies) for each
        slot count, cheapest first: every series takes its faster predicted
        strategy and the batch costs the sum of length x ns/element."""
        pred = np.exp(self.predict(X).astype(np.float64))
        lengths = np.asarray(lengths, dtype=np.float64)
        out = []
        for cap in CAPACITIES:
            cols = [CONFIGS.index((cap, st)) for st in STRATEGIES]
            t = pred[:, cols]
            pick = np.argmin(t, axis=1)
            cost = float((lengths * t[np.arange(len(t)), pick]).sum())
            out.append((cost, cap, np.asarray(STRATEGIES, dtype=np.uint8)[pick]))
        return sorted(out, key=lambda c: c[0])

    def choose_launch(self, X: np.ndarray, lengths: np.ndarray) -> tuple[int, np.ndarray]:
        """(LDS label slots, per-series strategies) with the least predicted time."""
        _, cap, st = self.rank_launches(X, lengths)[0]
        return int(cap), st

    def save(self, path: pathlib.Path | str = MODEL_PATH) -> None:
        path = pathlib.Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        arrays = {f"p{i}": p.cpu().numpy() for i, p in enumerate(self.params)}
        np.savez(path, x_mean=self.x_mean, x_std=self.x_std, t_mean=self.t_mean, t_std=self.t_std,
                 meta=np.array(json.dumps(self.meta)), **arrays)

    @classmethod
    def load(cls, path: pathlib.Path | str | None = None, device=None) -> "CostModel | None":
        """The saved model, or None when missing, unreadable or built for another
        architecture / kernel version (callers then fall back to the tuner)."""
        import os

        path = pathlib.Path(path or os.environ.get("NBG_COST_MODEL", MODEL_PATH))
        if not path.exists():
            return None
        try:
            with np.load(path, allow_pickle=False) as z:
                meta = json.loads(str(z["meta"]))
                m = cls(device=device, n_out=int(z["p5"].shape[0]))
                m.params = [torch.from_numpy(z[f"p{i}"].copy()).to(m.device) for i in range(len(m.params))]
                m.x_mean, m.x_std, m.t_mean, m.t_std = z["x_mean"], z["x_std"], z["t_mean"], z["t_std"]
                m.meta = meta
        except (OSError, ValueError, KeyError):
            return None
        if meta.get("arch") != device_arch() or meta.get("kernel") != kernel_version():
            return None
        if meta.get("configs") != [list(c) for c in CONFIGS]:
            return None
        return m


def exhaustive_launch(rb: RaggedMoveMeanGroupMean, values, labels, skew, nanf, reps: int = 2) -> tuple[int, np.ndarray]:
    """Exhaustive choice of (LDS label slots, per-series strategies). Series are
    bucketed by (log2 length, log2 labels, skew, NaN fraction); for each slot
    count, each bucket's own sub-batch is timed with both strategies and the
    faster one is kept; the whole batch is then timed once per slot count and
    the faster launch wins."""
    key = np.stack([
        np.round(np.log2(rb.lengths)), np.round(np.log2(np.maximum(rb.num_labels, 1))),
        np.round(skew * 10), np.round(nanf * 10)], axis=1)
    _, bucket = np.unique(key, axis=0, return_inverse=True)
    bucket = bucket.ravel()
    tuner = Autotuner(db_path=pathlib.Path("/nonexistent/db.json"), reps=reps)
    samples = []
    for b in np.unique(bucket):
        idx = np.nonzero(bucket == b)[0]
        # time a sample of the bucket: up to 64 series / about 8M elements
        budget, take = 1 << 23, []
        for i in idx[:64]:
            take.append(i)
            budget -= int(rb.lengths[i])
            if budget <= 0:
                break
        sample = np.array(take)
        sel = np.concatenate([np.arange(rb.offsets[i], rb.offsets[i + 1]) for i in sample])
        sel_t = torch.from_numpy(sel).to(values.device)
        sub = RaggedMoveMeanGroupMean(rb.lengths[sample], rb.num_labels[sample], window=rb.window, min_count=rb.min_count)
        samples.append((idx, sub, values[sel_t].contiguous(), labels[sel_t].contiguous()))
    best = None
    for cap in CAPACITIES:
        out = np.full(rb.nseries, STRATEGY_LDS, dtype=np.uint8)
        for idx, sub, sv, sl in samples:
            t = [tuner._time(sub, sv, sl, st, cap) for st in STRATEGIES]
            out[idx] = STRATEGIES[int(np.argmin(t))]
        t_all = tuner._time(rb, values, labels, out, cap)
        if best is None or t_all < best[0]:
            best = (t_all, cap, out)
    return int(best[1]), best[2]


def launch_candidates(close, lengths) -> list[tuple[int, np.ndarray]]:
    """The launches to time for the model's close slot counts: at each, the
    model's per-series strategies and, where they differ from it, every series
    on the strategy the model gives most elements. Trained on homogeneous
    batches, the model can send series of a mixed batch to the slower strategy
    (measured on MI355X, 20,000 mixed series: 1,425 of them to global atomics,
    10.9 ms against 9.8 ms with all on LDS slots); timing the uniform launch
    beside it keeps such a choice from costing more than the fixed one."""
    w = np.asarray(lengths, dtype=np.float64)
    out = []
    for _, cap, st in close:
        out.append((int(cap), st))
        lds_share = float(w[st == STRATEGY_LDS].sum()) / max(float(w.sum()), 1.0)
        uni = np.full(len(st), STRATEGY_LDS if lds_share >= 0.5 else STRATEGY_GLOBAL, dtype=np.uint8)
        if not np.array_equal(uni, st):
            out.append((int(cap), uni))
    return out


def choose_launch(rb: RaggedMoveMeanGroupMean, values, labels, method: str = "model", model: CostModel | None = None,
                  move_out=None, group_out=None):
    """(LDS label slots, per-series strategies) for `rb` by "model" (falls back
    to "exhaustive" when no valid model is available), "exhaustive" or
    "heuristic" (default slots, the label-count rule). Launches timed for the
    choice write into `move_out` / `group_out` when given."""
    from .ragged import heuristic_strategy

    if method == "heuristic":
        return LDS_MAX_LABELS, heuristic_strategy(rb.lengths, rb.num_labels)
    skew, nanf = batch_statistics(rb, values, labels, sample=SAMPLE_PER_SERIES)
    if method == "model":
        model = model or CostModel.load()
        if model is not None:
            ranked = model.rank_launches(features(rb.lengths, rb.num_labels, skew, nanf, rb.window), rb.lengths)
            # Candidates predicted within TIE_MARGIN of the best are timed on the
            # batch itself (one launch each) and the faster is kept: the model
            # narrows the choice, a measurement settles close calls.
            close = [c for c in ranked if c[0] <= ranked[0][0] * (1.0 + TIE_MARGIN)]
            cands = launch_candidates(close, rb.lengths)
            if len(cands) == 1:
                return cands[0]
            tuner = Autotuner(db_path=pathlib.Path("/nonexistent/db.json"), reps=1)
            timed = [(tuner._time(rb, values, labels, st, cap, move_out, group_out), cap, st) for cap, st in cands]
            _, cap, st = min(timed, key=lambda c: c[0])
            return int(cap), st
        method = "exhaustive"
    if method == "exhaustive":
        return exhaustive_launch(rb, values, labels, skew, nanf)
    raise ValueError(f"unknown method {method!r}")
