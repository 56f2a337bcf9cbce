"""Train the partition pre-sizing MLP on the GPU and write gkc_amd/models/presize_mlp.npz.

    python -m gkc_amd.train_presize [--seeds 20] [--scale 1.0] [--epochs 300] [--out PATH]

Training pairs: (sketch, exact log total / log distinct) per partition, from
the plain counter on the short-read (C1), chunked short-read (C2) and
long-read HPC (C3) workloads over genome seeds 1..S.  Held-out evaluation:
seeds 101..105 of the same workloads and of the skewed-GC workload (C5).
The margin is the 99th percentile of true/predicted distinct on the
training pairs, so that about 1 % of the partitions take the
count-then-allocate fallback.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import learned


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", type=int, default=20)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--eval-scale", type=float, default=1.0)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=learned.DEFAULT_MODEL)
    ap.add_argument("--report", default=None)
    args = ap.parse_args()
    log = lambda *a: print("[train]", *a, file=sys.stderr, flush=True)  # noqa: E731
    t0 = time.time()
    X, T = learned.collect_pairs([1, 2, 3], range(1, args.seeds + 1), scale=args.scale, log=log)
    log(f"{len(X)} training pairs in {time.time() - t0:.0f}s")
    x_mean, x_std = X.mean(0), X.std(0) + 1e-3
    t_mean, t_std = T.mean(0), T.std(0) + 1e-6
    Xn = (X - x_mean) / x_std
    Tn = (T - t_mean) / t_std
    mlp = learned.MLP(seed=1234)
    t1 = time.time()
    losses = mlp.train(Xn, Tn, epochs=args.epochs, batch=256, lr=args.lr, seed=99)
    log(f"trained {args.epochs} epochs in {time.time() - t1:.1f}s, loss {losses[0]:.4f} -> {losses[-1]:.5f}")
    pred = learned.Predictor(mlp.get_params(), x_mean, x_std, t_mean, t_std, margin=1.0)
    # margin from the training distribution of true / predicted distinct
    y = mlp.forward_bf16(Xn) * t_std + t_mean
    ratio = np.expm1(T[:, 1]) / np.maximum(np.expm1(y[:, 1]), 1.0)
    pred.margin = float(max(1.0, np.quantile(ratio, 0.99)))
    report = {"pairs": int(len(X)), "epochs": args.epochs, "loss_first": float(losses[0]),
              "loss_last": float(losses[-1]), "margin": pred.margin, "eval": {}}
    for cid in [1, 2, 3, 5]:
        Xe, Te = learned.collect_pairs([cid], range(101, 106), scale=args.eval_scale, log=log)
        tot, dist = pred.predict(Xe)
        tt, td = np.expm1(Te[:, 0]), np.expm1(Te[:, 1])
        caps = np.minimum(np.ceil(dist * pred.margin) + 16, tt)
        report["eval"][f"C{cid}"] = {
            "partitions": int(len(Xe)),
            "mape_total_pct": learned.mape(tot, tt),
            "mape_distinct_pct": learned.mape(dist, td),
            "under_predicted_pct": float(np.mean(caps < td) * 100.0),
            "table_capacity_over_exact": float(caps.sum() / td.sum()),
            "table_capacity_over_total_bound": float(caps.sum() / tt.sum()),
        }
        log(f"C{cid}: {report['eval'][f'C{cid}']}")
    pred.info = report
    pred.save(args.out)
    rep = args.report or os.path.splitext(args.out)[0] + "_report.json"
    with open(rep, "w") as f:
        json.dump(report, f, indent=1)
    log(f"saved {args.out}")
    print(json.dumps(report))
    return 0


if __name__ == "__main__":
    sys.exit(main())
