"""Build the strategy timing DB and train the cost model (needs a GPU).

    python -m numbagg_amd.tuning_data.make_tuning_data [--out DIR]

1. The autotuner times every launch configuration of the ragged kernel
   ((LDS label slots) x (group-by strategy)) on a grid of homogeneous
   synthetic batches (length x labels x skew x NaN fraction x window) and
   writes ragged_timing_db.json.
2. The cost model (5 -> 64 -> 64 -> 4 MLP) is trained on 80% of the points by
   the library's HIP kernels; its regret on the held-out 20% is measured
   against the exhaustive choice (the fastest measured configuration) and
   compared with each fixed configuration and the label-count heuristic.
3. Weights go to ragged_cost_model.npz, the evaluation to
   ragged_cost_model_report.json.
"""

from __future__ import annotations

import argparse
import json
import pathlib
import sys
import time

import numpy as np


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(pathlib.Path(__file__).resolve().parent))
    ap.add_argument("--epochs", type=int, default=3000)
    args = ap.parse_args(argv)
    out = pathlib.Path(args.out)
    out.mkdir(parents=True, exist_ok=True)

    from numbagg_amd.tuning import Autotuner, CostModel, device_arch, kernel_version

    t0 = time.time()
    tuner = Autotuner(db_path=out / "ragged_timing_db.json")
    tuner.entries = []
    def progress(i, entry):
        if i % 24 == 0:
            print(f"# autotuner: {i} points, {time.time() - t0:.0f}s (length {entry['length']}, "
                  f"labels {entry['num_labels']})", file=sys.stderr, flush=True)

    tuner.run_grid(progress=progress)
    tuner.save()
    print(f"# autotuner: {len(tuner.entries)} points in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    from numbagg_amd.tuning import CONFIGS, LDS_MAX_LABELS, STRATEGY_GLOBAL, STRATEGY_LDS

    X, T, K = tuner.training_set()
    ns = np.exp(T.astype(np.float64))
    rng = np.random.default_rng(0)
    perm = rng.permutation(len(X))
    ntr = int(0.8 * len(X))
    tr, te = perm[:ntr], perm[ntr:]

    model = CostModel(seed=0)
    hist = model.fit(X[tr], T[tr], epochs=args.epochs, lr=3e-3)
    choice = model.choose(X[te])

    def total(sel):
        return float(ns[te, sel].sum())

    best = float(ns[te].min(axis=1).sum())
    heur = np.array([CONFIGS.index((LDS_MAX_LABELS, STRATEGY_LDS if k <= LDS_MAX_LABELS else STRATEGY_GLOBAL))
                     for k in K[te]])
    report = {
        "arch": device_arch(),
        "kernel": kernel_version(),
        "configs": [list(c) for c in CONFIGS],
        "points": int(len(X)),
        "train": int(ntr),
        "heldout": int(len(te)),
        "loss_first": hist[0],
        "loss_last": hist[-1],
        "heldout_regret_model": total(choice) / best - 1,
        "heldout_regret_heuristic": total(heur) / best - 1,
        "heldout_regret_fixed": {f"{c}/{'lds' if st == STRATEGY_LDS else 'global'}": total(np.full(len(te), i)) / best - 1
                                 for i, (c, st) in enumerate(CONFIGS)},
        "heldout_accuracy_model": float(np.mean(choice == ns[te].argmin(axis=1))),
        "best_config_counts": {f"{c}/{'lds' if st == STRATEGY_LDS else 'global'}": int(np.sum(ns.argmin(axis=1) == i))
                               for i, (c, st) in enumerate(CONFIGS)},
    }
    model.meta = {"arch": report["arch"], "kernel": report["kernel"], "configs": report["configs"],
                  "points": int(len(X))}
    model.save(out / "ragged_cost_model.npz")
    (out / "ragged_cost_model_report.json").write_text(json.dumps(report, indent=1))
    print(json.dumps(report))
    return 0


if __name__ == "__main__":
    sys.exit(main())
