"""Trains the bracket-position network on records the exact method labels.

    python -m robotfd_amd.train_bracket [--pairs 200000] [--out robotfd_amd/data/bracket_net.bin]

Pairs come from the random-walk generator of the batch workload (N(0,1)
steps, dims 2/3 at random) with lengths log-uniform in [60, 2000] per curve,
shorter than the evaluation pairs so that labelling 200,000 of them stays
cheap; the features are scale-free (curves divided by DFD).  Every pair is
solved by the plain batched pipeline (DFD, VE-FD, bisection to 1e-9
relative); pairs whose bracket is open become records.  The network trains on
the GPU (Adam, binary cross-entropy); the certification policy is calibrated
on a held-out tenth of the records.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import workloads as W
from .batch import Batch
from .frechet import Context
from .learned import (DEFAULT_WEIGHTS, BracketNet, calibrate_policy, end_flags, records_from_run,
                      save_records)


def walk_pairs(count: int, seed: int, lo: int = 60, hi: int = 2000):
    rng = np.random.default_rng(seed)
    lens = np.exp(rng.uniform(np.log(lo), np.log(hi), size=(count, 2))).astype(np.int64)
    dims = np.where(rng.uniform(size=count) < 0.5, 2, 3)
    return [(W.random_walk(rng, int(lens[k, 0]), int(dims[k])),
             W.random_walk(rng, int(lens[k, 1]), int(dims[k]))) for k in range(count)]


def label(count: int, seed: int, ctx: Context, chunk: int = 20000, log=print):
    Xs, ys, vs, ds, dist, steps, lows, highs = [], [], [], [], [], [], [], []
    done = 0
    while done < count:
        n = min(chunk, count - done)
        pairs = walk_pairs(n, seed + done)
        with Batch(pairs, ctx=ctx) as b:
            out = b.run(rows_per_lane=2)
            X, y, keep = records_from_run(b, out)
            lo, hi = end_flags(b, rows_per_lane=2)
        Xs.append(X[keep]); ys.append(y[keep]); vs.append(out["vefd"][keep]); ds.append(out["dfd"][keep])
        dist.append(out["distance"][keep]); steps.append(out["iterations"][keep])
        lows.append(lo[keep]); highs.append(hi[keep])
        done += n
        log(f"labelled {done}/{count} pairs, {sum(len(v) for v in ys)} records", flush=True)
    cat = np.concatenate
    return (cat(Xs), cat(ys), cat(vs), cat(ds), cat(dist), cat(steps), cat(lows), cat(highs))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=200_000)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--records", default="")
    args = ap.parse_args()
    ctx = Context.default()
    t0 = time.time()
    X, y, vefd, dfd, dist, steps, at_low, at_high = label(args.pairs, args.seed, ctx)
    if args.records:
        save_records(args.records, X, y, vefd, dfd, dist)
    rng = np.random.default_rng(args.seed)
    idx = rng.permutation(len(y))
    nval = max(1, len(y) // 10)
    va, tr = idx[:nval], idx[nval:]
    net = BracketNet(seed=args.seed, ctx=ctx)
    losses = net.train(X[tr], y[tr], epochs=args.epochs, batch=128, lr=1e-3, seed=args.seed)
    pred = net.forward(X[va])
    tau, delta, gain = calibrate_policy(pred, y[va], steps[va], at_low[va], at_high[va])
    net.policy = (tau, delta)
    net.save(args.out)
    summary = {"pairs": args.pairs, "records": int(len(y)), "train": int(len(tr)), "val": int(nval),
               "loss_first": float(losses[0]), "loss_last": float(losses[-1]), "tau": tau,
               "delta": delta, "val_sweeps_saved": gain, "val_sweeps_plain": int(steps[va].sum()),
               "seconds": time.time() - t0, "weights": args.out}
    print(json.dumps(summary), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
