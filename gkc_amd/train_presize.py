"""Train the partition pre-sizing MLP on the GPU and write gkc_amd/models/presize_mlp.npz.

    python -m gkc_amd.train_presize [--seeds 20] [--scale 1.0] [--epochs 300] [--out PATH]

Training pairs: (sketch, exact log total / log distinct-over-total) per partition, from
the plain counter on the short-read (C1), chunked short-read (C2) and
long-read HPC (C3) workloads over genome seeds 1..S.  Held-out evaluation:
seeds 101..105 of the same workloads and of the skewed-GC workload (C5).
The margin is the 99.5th percentile of true/predicted distinct on the
training pairs times a safety factor (1.05) for inputs unlike the training
workloads.  The binned count lays its table out densely from the exact
counts, so a count falls back to count-then-allocate only when the caps'
sum is short of the table; the report gives, per held-out run, whether it
is (`runs_short`), next to the per-partition under-prediction rate (what
the full-LSD count, with its per-partition capacities, would see).
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
    ap.add_argument("--margin-quantile", type=float, default=0.995)
    ap.add_argument("--margin-safety", type=float, default=1.05,
                    help="extra factor on the margin for inputs unlike the training workloads")
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
    # margin from the training distribution of true / predicted distinct ratio
    y = mlp.forward_bf16(Xn) * t_std + t_mean
    ratio = np.exp(T[:, 1] - np.minimum(y[:, 1], 0.0))
    pred.margin = float(max(1.0, np.quantile(ratio, args.margin_quantile)) * args.margin_safety)
    report = {"pairs": int(len(X)), "epochs": args.epochs, "loss_first": float(losses[0]),
              "loss_last": float(losses[-1]), "margin": pred.margin, "eval": {}}
    for cid in [1, 2, 3, 5]:
        runs = [learned.collect_pairs([cid], [g], scale=args.eval_scale, log=log) for g in range(101, 106)]
        tt_all, td_all, tot_all, dist_all, caps_all, short = [], [], [], [], [], 0
        for Xe, Te in runs:
            tt = np.rint(np.expm1(Te[:, 0]))
            td = np.rint(tt * np.exp(Te[:, 1]))
            tot, dist = pred.predict(Xe, tt)
            caps = np.minimum(np.ceil(dist * pred.margin) + 16, tt)
            short += int(caps.sum() < td.sum())
            for a, v in ((tt_all, tt), (td_all, td), (tot_all, tot), (dist_all, dist), (caps_all, caps)):
                a.append(v)
        tt, td, tot, dist, caps = (np.concatenate(a) for a in (tt_all, td_all, tot_all, dist_all, caps_all))
        report["eval"][f"C{cid}"] = {
            "partitions": int(len(tt)),
            "runs": len(runs),
            "runs_short": short,
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
