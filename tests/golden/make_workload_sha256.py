"""Writes workload_sha256.json: SHA-256 of the FP64 bytes of every benchmark
workload, so a run on another machine can check that it regenerated the same
inputs (same numpy PCG64 stream and arithmetic).

    python tests/golden/make_workload_sha256.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from robotfd_amd import workloads as W  # noqa: E402


def digests() -> dict:
    out = {}
    P, Q = W.small_2d(0)
    out["small_2d_2000x2500_seed0"] = [W.sha256(P), W.sha256(Q)]
    P, Q = W.long_3d(200_000, 200_000, seed=0)
    out["long_3d_200000x200000_seed0"] = [W.sha256(P), W.sha256(Q)]
    P, Q = W.robot_pair(100_000, 120_000, seed=0)
    out["robot_3d_100000x120000_seed0"] = [W.sha256(P), W.sha256(Q)]
    pairs = W.batch_pairs(4096, seed=0)
    h = [W.sha256(np.concatenate([P.ravel(), Q.ravel()])) for P, Q in pairs]
    import hashlib
    out["batch_4096_seed0"] = hashlib.sha256("".join(h).encode()).hexdigest()
    out["batch_4096_seed0_lengths"] = hashlib.sha256(
        np.array([[len(P), len(Q), P.shape[1]] for P, Q in pairs], dtype="<i8").tobytes()).hexdigest()
    return out


if __name__ == "__main__":
    with open(os.path.join(HERE, "workload_sha256.json"), "w") as f:
        json.dump(digests(), f, indent=1)
        f.write("\n")
