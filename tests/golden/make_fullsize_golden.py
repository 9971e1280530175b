"""Reference values of the full-size workloads, computed by the CPU oracle.

    python tests/golden/make_fullsize_golden.py [--threads 8] [--only c2,c3,c4]

Writes, next to this script:

* fullsize.json: the 100,000 x 120,000 noisy-copy pair and the 200,000 x
  200,000 random-walk pair (seed 0): endpoint bound, DFD, VE-FD, the distance
  to 1e-9 relative, and the terminal rightUB of decisions at DFD (1 -/+ 0.5e-9).
  The sweeps run on the oracle's row-pipelined form (or_sweep_mt), which the
  oracle tests pin bit for bit to the single-thread restatement.
* batch4096.npz: for each of the 4,096 pairs of the batch workload (seed 0)
  the endpoint bound, DFD, VE-FD, the final bracket, the distance to 1e-9
  relative and the bisection steps: or_distance's loop with its decisions on
  the pipelined oracle at one thread (one pair per process).

Inputs are regenerated from robotfd_amd.workloads and identified by their
SHA-256 (tests/golden/workload_sha256.json).  Floats are stored exactly
(JSON repr round-trips, npz holds the FP64 bits).
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from robotfd_amd import workloads as W  # noqa: E402

REL = 1e-9


def mt_distance(P, Q, threads: int) -> dict:
    """frechet.cpp:18-75 with the rel 1e-9 stop, every sweep on the
    row-pipelined oracle."""
    lb0 = O.endpoint_lb(P, Q)
    t = time.time()
    dfd = O.sweep_mt(P, Q, 2, 0.0, threads)
    print(f"  dfd {dfd!r} ({time.time() - t:.0f} s)", flush=True)
    t = time.time()
    vefd = O.sweep_mt(P, Q, 0, lb0, threads)
    print(f"  vefd {vefd!r} ({time.time() - t:.0f} s)", flush=True)
    lb, ub, it = vefd, dfd, 0
    while ub - lb > REL * ub:
        split = (ub + lb) / 2.0
        reach = O.sweep_mt(P, Q, 1, split, threads) == 1.0
        if reach:
            ub = split
        else:
            lb = split
        it += 1
    decisions = []
    for f in (1 - 0.5e-9, 1 + 0.5e-9):
        r = dfd * f
        t = time.time()
        ubr = O.sweep_mt(P, Q, 1, r, threads)
        print(f"  decide {r!r} -> {ubr!r} ({time.time() - t:.0f} s)", flush=True)
        decisions.append([r, ubr])
    return {"endpoint_lb": lb0, "dfd": dfd, "vefd": vefd, "lb": lb, "ub": ub,
            "distance": (ub + lb) / 2.0, "iterations": it, "decisions": decisions}


def pair_distance(P, Q):
    """or_distance (frechet.cpp:18-75, rel 1e-9 stop, VE-FD gate in 64 bits)
    with each bisection decision on the pipelined oracle at one thread, which
    skips the interval arithmetic of unreachable cells (pinned bit-identical
    to or_decide_ub by the oracle tests)."""
    lb0 = O.endpoint_lb(P, Q)
    dfd = O.dfd(P, Q)
    vefd = O.vefd(P, Q, lb0) if len(P) * len(Q) > 10000 else lb0
    lb, ub, it = vefd, dfd, 0
    while ub - lb > REL * ub:
        split = (ub + lb) / 2.0
        if O.sweep_mt(P, Q, 1, split, 1) == 1.0:
            ub = split
        else:
            lb = split
        it += 1
    return np.array([(ub + lb) / 2.0, lb0, dfd, vefd, lb, ub]), it


def one_pair(args):
    k, P, Q = args
    out, it = pair_distance(P, Q)
    return k, out, it


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--only", default="c2,c3,c4")
    a = ap.parse_args()
    only = set(a.only.split(","))
    O.build()
    path = os.path.join(HERE, "fullsize.json")
    gold = json.load(open(path)) if os.path.exists(path) else {}
    if "c2" in only:
        print("C2 100k x 120k noisy copy", flush=True)
        P, Q = W.robot_pair(100_000, 120_000, seed=0)
        gold["c2"] = {"workload": "robot_pair(100000, 120000, seed=0)", "sha256": [W.sha256(P), W.sha256(Q)],
                      **mt_distance(P, Q, a.threads)}
        json.dump(gold, open(path, "w"), indent=1)
    if "c3" in only:
        print("C3 200k x 200k random walks", flush=True)
        P, Q = W.long_3d(200_000, 200_000, seed=0)
        gold["c3"] = {"workload": "long_3d(200000, 200000, seed=0)", "sha256": [W.sha256(P), W.sha256(Q)],
                      **mt_distance(P, Q, a.threads)}
        json.dump(gold, open(path, "w"), indent=1)
    if "c4" in only:
        print("C4 4096 pairs", flush=True)
        pairs = W.batch_pairs(4096, seed=0)
        res = np.zeros((len(pairs), 6))
        iters = np.zeros(len(pairs), dtype=np.int64)
        t = time.time()
        # Largest pairs first, so the pool ends evenly.
        order = sorted(range(len(pairs)), key=lambda k: -len(pairs[k][0]) * len(pairs[k][1]))
        # The loop above restates or_distance: check it on the smallest pairs.
        for k in order[-6:]:
            out, it = pair_distance(*pairs[k])
            ref = O.distance(*pairs[k], stop_mode=1, eps=REL, max_trace=1)
            assert out[0] == ref["distance"] and it == ref["iterations"], k
        with mp.Pool(a.threads) as pool:
            for done, (k, out, it) in enumerate(pool.imap_unordered(
                    one_pair, ((k, pairs[k][0], pairs[k][1]) for k in order), chunksize=4)):
                res[k] = out
                iters[k] = it
                if done % 512 == 0:
                    print(f"  {done} pairs, {time.time() - t:.0f} s", flush=True)
        sha = W.sha256(np.concatenate([np.concatenate([P.ravel(), Q.ravel()]) for P, Q in pairs]))
        np.savez(os.path.join(HERE, "batch4096.npz"), distance=res[:, 0], endpoint_lb=res[:, 1],
                 dfd=res[:, 2], vefd=res[:, 3], lb=res[:, 4], ub=res[:, 5], iterations=iters,
                 sizes=np.array([(len(P), len(Q), P.shape[1]) for P, Q in pairs], dtype=np.int64),
                 points_sha256=np.array(sha))
        print(f"  open pairs: {int((iters > 0).sum())}, {time.time() - t:.0f} s", flush=True)


if __name__ == "__main__":
    main()
