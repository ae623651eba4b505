"""Build the strategy timing DB and train the cost model (needs a GPU).

    python -m numbagg_amd.tuning_data.make_tuning_data [--out DIR]

1. The autotuner times both group-by strategies of the ragged kernel on a grid
   of homogeneous synthetic batches (length x labels x skew x NaN fraction x
   window) and writes ragged_timing_db.json.
2. The cost model (5 -> 64 -> 64 -> 2 MLP) is trained on 80% of the points by
   the library's HIP kernels; its regret on the held-out 20% is measured
   against the exhaustive choice (the faster measured strategy) and compared
   with always-global, always-LDS and the label-count heuristic.
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

    from numbagg_amd.tuning import Autotuner, CostModel, device_arch, features, kernel_version

    t0 = time.time()
    tuner = Autotuner(db_path=out / "ragged_timing_db.json")
    tuner.entries = []
    tuner.run_grid()
    tuner.save()
    print(f"# autotuner: {len(tuner.entries)} points in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    rows = tuner.entries
    lds = np.array([e["ns_lds"] for e in rows])
    glb = np.array([e["ns_global"] for e in rows])
    X = features([e["length"] for e in rows], [e["num_labels"] for e in rows], [e["skew"] for e in rows],
                 [e["nan_fraction"] for e in rows], [e["window"] for e in rows])
    T = np.log(np.stack([lds, glb], 1)).astype(np.float32)
    K = np.array([e["num_labels"] for e in rows])
    rng = np.random.default_rng(0)
    perm = rng.permutation(len(rows))
    ntr = int(0.8 * len(rows))
    tr, te = perm[:ntr], perm[ntr:]

    model = CostModel(seed=0)
    hist = model.fit(X[tr], T[tr], epochs=args.epochs, lr=3e-3)
    choice = model.choose(X[te], K[te])

    def total(sel):
        return float(np.where(sel == 0, lds[te], glb[te]).sum())

    best = float(np.minimum(lds[te], glb[te]).sum())
    report = {
        "arch": device_arch(),
        "kernel": kernel_version(),
        "points": len(rows),
        "train": int(ntr),
        "heldout": int(len(te)),
        "loss_first": hist[0],
        "loss_last": hist[-1],
        "heldout_regret_model": total(choice) / best - 1,
        "heldout_regret_always_global": total(np.ones(len(te), np.uint8)) / best - 1,
        "heldout_regret_always_lds": total(np.zeros(len(te), np.uint8)) / best - 1,
        "heldout_regret_heuristic": total(np.where(K[te] <= 256, 0, 1).astype(np.uint8)) / best - 1,
        "heldout_accuracy_model": float(np.mean(choice == (lds[te] > glb[te]).astype(np.uint8))),
    }
    model.meta = {"arch": report["arch"], "kernel": report["kernel"], "points": len(rows)}
    model.save(out / "ragged_cost_model.npz")
    (out / "ragged_cost_model_report.json").write_text(json.dumps(report, indent=1))
    print(json.dumps(report))
    return 0


if __name__ == "__main__":
    sys.exit(main())
