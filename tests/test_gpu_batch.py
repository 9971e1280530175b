"""Batched pipeline parity: every pair of a batch against the CPU oracle,
bit-exact, for every tile height, with and without the hipGraph, and across
graph replays."""
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O
from robotfd_amd import Context, workloads as W
from robotfd_amd.batch import Batch

pytestmark = pytest.mark.gpu


def mixed_pairs(count, seed, lo=60, hi=900):
    rng = np.random.default_rng(seed)
    pairs = []
    for k in range(count):
        m, n = (np.exp(rng.uniform(np.log(lo), np.log(hi), 2))).astype(int)
        d = 2 if rng.uniform() < 0.5 else 3
        if k % 3 == 0:
            P, Q = W.robot_pair(max(m, 20), max(n, 20), seed=seed * 1000 + k)
            if d == 2:
                P, Q = np.ascontiguousarray(P[:, :2]), np.ascontiguousarray(Q[:, :2])
        else:
            P = np.cumsum(rng.standard_normal((m, d)), axis=0)
            Q = np.cumsum(rng.standard_normal((n, d)), axis=0)
        pairs.append((P, Q))
    pairs.append((np.array([[0.0, 0.0]]), np.array([[1.0, 0.0], [0.0, 2.0]])))  # 1-point curve
    pairs.append((np.array([[0.0, 0.0], [1.0, 1.0]]), np.array([[0.5, 0.0], [1.0, 2.0]])))
    return pairs


@pytest.fixture(scope="module")
def batch_and_ref(gpu):
    pairs = mixed_pairs(40, seed=5)
    refs = [O.distance(P, Q, stop_mode=1, eps=1e-9) for P, Q in pairs]
    b = Batch(pairs, ctx=gpu)
    yield b, refs
    b.close()


@pytest.mark.parametrize("rows,depth", [(1, 1), (2, 0), (3, 0), (4, 0), (4, 2), (2, 3)])
def test_batch_matches_oracle_per_pair(batch_and_ref, rows, depth):
    """Also pins speculative bisection: walking the evaluated tree must give
    the sequential trajectory (same iterations, same final bracket)."""
    b, refs = batch_and_ref
    print("batch rows", rows, "depth", depth, flush=True)
    out = b.run(stop="rel", eps=1e-9, rows_per_lane=rows, spec_depth=depth)
    for k, ref in enumerate(refs):
        assert out["dfd"][k] == ref["dfd"], k
        assert out["vefd"][k] == ref["vefd"], k
        assert out["iterations"][k] == ref["iterations"], k
        assert out["distance"][k] == ref["distance"], k
    assert sum(r["iterations"] for r in refs) > 0  # some pairs bisect


def test_batch_graph_replay_and_stream_agree(batch_and_ref):
    b, refs = batch_and_ref
    a = b.run(rows_per_lane=2, use_graph=True)
    c = b.run(rows_per_lane=2, use_graph=True)
    d = b.run(rows_per_lane=2, use_graph=False)
    for key in ("distance", "dfd", "vefd", "iterations", "lb", "ub"):
        assert np.array_equal(a[key], c[key]), key
        assert np.array_equal(a[key], d[key]), key
    assert (a["latency_ms"] > 0).all()


def ctx_with(**env):
    saved = {k: os.environ.get(k) for k in env}
    os.environ.update({k: str(v) for k, v in env.items()})
    try:
        return Context()
    finally:
        for k, v in saved.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


@pytest.mark.parametrize("bcfg", [0, 21, 42, 63, 77, 137, 1310])
def test_batch_bound_round_configs(batch_and_ref, bcfg):
    """Bounds as separate DFD and VE-FD rounds (0) or as one fused round in
    each configuration (rows per lane, producer waves)."""
    _, refs = batch_and_ref
    ctx = ctx_with(RFD_BATCH_BOUNDS=bcfg)
    try:
        with Batch(mixed_pairs(40, seed=5), ctx=ctx) as b:
            print("batch bounds cfg", bcfg, flush=True)
            out = b.run(stop="rel", eps=1e-9, rows_per_lane=2, spec_depth=1)
        for k, ref in enumerate(refs):
            assert out["dfd"][k] == ref["dfd"], k
            assert out["vefd"][k] == ref["vefd"], k
            assert out["iterations"][k] == ref["iterations"], k
            assert out["distance"][k] == ref["distance"], k
    finally:
        ctx.close()


def test_batch_pruning_does_not_change_results(batch_and_ref):
    """Pruning applies to the separate VE-FD round."""
    ctx = ctx_with(RFD_BATCH_BOUNDS=0)
    try:
        with Batch(mixed_pairs(40, seed=5), ctx=ctx) as b:
            on = b.run(rows_per_lane=4, prune=1)
            off = b.run(rows_per_lane=4, prune=0)
        for key in ("distance", "dfd", "vefd", "iterations"):
            assert np.array_equal(on[key], off[key]), key
    finally:
        ctx.close()


def test_batch_reference_stop_rule(batch_and_ref):
    b, refs = batch_and_ref
    pairs = mixed_pairs(40, seed=5)
    eps = float(np.finfo(np.float32).eps)
    out = b.run(stop="abs", eps=eps, rows_per_lane=2)
    for k in range(0, len(pairs), 5):
        ref = O.distance(*pairs[k], stop_mode=0, eps=eps)
        assert out["distance"][k] == ref["distance"], k
        assert out["iterations"][k] == ref["iterations"], k


def test_batch_autotuner_picks_a_valid_height(batch_and_ref):
    b, _ = batch_and_ref
    best, ms = b.tune()
    assert best in (1, 2, 3, 4)
    assert all(t > 0 for t in ms)
    assert ms[best - 1] == min(ms)


def test_batch_bound_round_tall_pairs():
    """The fused bound round on pairs of many bands (832 rows each at 13 rows
    per lane) and few columns, 2D and 3D, whose bands the round's work list
    spreads apart: bit-exact bounds and distances against the oracle."""
    rng = np.random.default_rng(11)
    pairs = []
    for k in range(10):
        m = int(rng.integers(1500, 5200))
        n = int(rng.integers(40, 400))
        d = 2 + (k % 2)
        P = np.cumsum(rng.standard_normal((m, d)), axis=0)
        Q = np.cumsum(rng.standard_normal((n, d)), axis=0)
        if k >= 8:
            P, Q = Q, P  # wide pairs too
        pairs.append((P, Q))
    refs = [O.distance(P, Q, stop_mode=1, eps=1e-9) for P, Q in pairs]
    ctx = ctx_with(RFD_BATCH_BOUNDS=1310)
    try:
        with Batch(pairs, ctx=ctx) as b:
            print("batch tall pairs", [(len(P), len(Q), P.shape[1]) for P, Q in pairs], flush=True)
            out = b.run(stop="rel", eps=1e-9, rows_per_lane=2, spec_depth=1)
            again = b.run(stop="rel", eps=1e-9, rows_per_lane=2, spec_depth=0)
        for k, ref in enumerate(refs):
            assert out["dfd"][k] == ref["dfd"], k
            assert out["vefd"][k] == ref["vefd"], k
            assert out["iterations"][k] == ref["iterations"], k
            assert out["distance"][k] == ref["distance"], k
            assert again["distance"][k] == ref["distance"], k
    finally:
        ctx.close()


@pytest.mark.parametrize("env,rows", [(dict(RFD_BATCH_DEC=0), 1), (dict(RFD_BATCH_DEC=0), 3),
                                      (dict(RFD_MAX_WAVES=2), 2), (dict(RFD_MAX_WAVES=2), 4)])
def test_batch_decision_kernels_match_oracle(env, rows):
    """Decision rounds on band_sweep (RFD_BATCH_DEC=0) and on the windowed
    sweep with two workgroups taking every item: every pair bit-exact
    against the oracle, sequential and speculative."""
    ctx = ctx_with(**env)
    try:
        pairs = mixed_pairs(24, seed=9)
        refs = [O.distance(P, Q, stop_mode=1, eps=1e-9) for P, Q in pairs]
        assert sum(r["iterations"] for r in refs) > 0
        with Batch(pairs, ctx=ctx) as b:
            for depth in (1, 0):
                print("batch decisions", env, "depth", depth, flush=True)
                out = b.run(stop="rel", eps=1e-9, rows_per_lane=rows, spec_depth=depth)
                for k, ref in enumerate(refs):
                    assert out["iterations"][k] == ref["iterations"], (env, depth, k)
                    assert out["distance"][k] == ref["distance"], (env, depth, k)
    finally:
        ctx.close()
