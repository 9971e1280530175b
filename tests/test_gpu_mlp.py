"""The network's forward, gradients and training on the GPU against the CPU
network oracle, the batch features against the feature oracle, and the
learned batch run: certified results within 1e-9 of plain bisection, a forced
fallback, and the trained network's gain in decision sweeps."""
import os

import numpy as np
import pytest

from oracle import mlp_oracle as M
from oracle import oracle as O
from robotfd_amd import workloads as W
from robotfd_amd.batch import Batch
from robotfd_amd.learned import (BracketNet, batch_features, calibrate_policy, end_flags,
                                 records_from_run, run_learned)

pytestmark = pytest.mark.gpu


def layer_blocks():
    for l in range(M.LAYERS):
        K = M.IN if l == 0 else M.H
        yield f"W{l + 1}", M.w_off(l), M.w_off(l) + K * M.H
        yield f"b{l + 1}", M.b_off(l), M.b_off(l) + M.H
    yield "w5", M.W5, M.W5 + M.H
    yield "b5", M.B5, M.B5 + 1


def test_forward_matches_oracle(gpu):
    net = BracketNet(seed=3, ctx=gpu)
    params = net.get_params()
    rng = np.random.default_rng(0)
    X = (rng.standard_normal((150, M.FEAT)) * 0.7).astype(np.float32)
    got = net.forward(X)
    ref, _ = M.forward(params, X)
    assert np.max(np.abs(got - ref)) < 3e-3


def test_gradients_match_oracle(gpu):
    net = BracketNet(seed=4, ctx=gpu)
    params = net.get_params()
    rng = np.random.default_rng(1)
    X = (rng.standard_normal((130, M.FEAT)) * 0.7).astype(np.float32)
    y = rng.uniform(size=130).astype(np.float32)
    g, loss = net.gradients(X, y)
    gr = M.gradients(params, X, y)
    assert loss == pytest.approx(M.loss(params, X, y), rel=2e-3)
    for name, a, b in layer_blocks():
        err = np.linalg.norm(g[a:b] - gr[a:b]) / max(np.linalg.norm(gr[a:b]), 1e-12)
        assert err < 3e-2, (name, err)


def test_params_roundtrip_and_file(gpu, tmp_path):
    net = BracketNet(seed=5, ctx=gpu)
    p = net.get_params()
    p2 = p * 0.5
    net.set_params(p2)
    assert np.array_equal(net.get_params(), p2)
    net.policy = (0.1, 0.02)
    path = str(tmp_path / "w.bin")
    net.save(path)
    other = BracketNet(seed=9, ctx=gpu, path=path)
    assert np.array_equal(other.get_params(), p2)
    assert other.policy == (0.1, 0.02)


def test_training_lowers_loss(gpu):
    rng = np.random.default_rng(2)
    X = (rng.standard_normal((2048, M.FEAT)) * 0.5).astype(np.float32)
    y = (1 / (1 + np.exp(-(X[:, 0] * 2 - X[:, 5] + 0.5 * X[:, 384])))).astype(np.float32)
    net = BracketNet(seed=6, ctx=gpu)
    _, before = net.gradients(X, y)
    losses = net.train(X, y, epochs=25, batch=256, lr=2e-3, seed=1)
    _, after = net.gradients(X, y)
    assert after < before - 0.05
    assert losses[-1] < losses[0]


def noisy_pairs(count, seed, lo=120, hi=500):
    rng = np.random.default_rng(seed)
    pairs = []
    for k in range(count):
        m, n = (np.exp(rng.uniform(np.log(lo), np.log(hi), 2))).astype(int)
        P, Q = W.robot_pair(int(m), int(n), seed=seed * 10000 + k)
        if k % 2:
            P, Q = np.ascontiguousarray(P[:, :2]), np.ascontiguousarray(Q[:, :2])
        pairs.append((P, Q))
    return pairs


def test_batch_features_match_oracle(gpu):
    pairs = noisy_pairs(12, seed=3)
    with Batch(pairs, ctx=gpu) as b:
        out = b.run(rows_per_lane=2)
        X = batch_features(b)
    for k, (P, Q) in enumerate(pairs):
        ref = M.features(P, Q, out["vefd"][k], out["dfd"][k], O.endpoint_lb(P, Q))
        np.testing.assert_allclose(X[k], ref, rtol=1e-5, atol=1e-6)


@pytest.fixture(scope="module")
def trained(gpu):
    """Network trained on records the exact method labels, its policy
    calibrated on separate validation pairs."""
    train = noisy_pairs(1500, seed=11, lo=60, hi=300)
    with Batch(train, ctx=gpu) as b:
        out = b.run(rows_per_lane=2)
        X, y, keep = records_from_run(b, out)
    net = BracketNet(seed=7, ctx=gpu)
    net.train(X[keep], y[keep], epochs=60, batch=64, lr=1e-3, seed=2)
    val = noisy_pairs(800, seed=15, lo=60, hi=300)
    with Batch(val, ctx=gpu) as b:
        vout = b.run(rows_per_lane=2, spec_depth=1)
        Xv, yv, kv = records_from_run(b, vout)
        lo, hi = end_flags(b, rows_per_lane=2)
    print("val at low end", lo[kv].mean(), "at high end", hi[kv].mean(), flush=True)
    pred = net.forward(Xv[kv])
    tau, delta, gain = calibrate_policy(pred, yv[kv], vout["iterations"][kv], lo[kv], hi[kv])
    print("records", int(keep.sum()), "val open", int(kv.sum()), "tau", tau, "gain", gain, flush=True)
    net.policy = (tau, delta)
    return net


def test_learned_run_matches_plain_within_tolerance(gpu, trained):
    pairs = noisy_pairs(48, seed=12)
    with Batch(pairs, ctx=gpu) as b:
        plain = b.run(rows_per_lane=2, spec_depth=1)
        learned = run_learned(b, trained, rows_per_lane=2, spec_depth=1)
    assert plain["iterations"].sum() > 0
    d0, d1 = plain["distance"], learned["distance"]
    assert np.all(np.abs(d1 - d0) <= 1e-9 * d0 + 1e-300)
    assert np.all(learned["lb"] <= learned["ub"])
    assert np.all(learned["ub"] - learned["lb"] <= 1e-9 * learned["ub"])
    # the two brackets overlap: both contain the distance
    assert np.all(learned["lb"] <= plain["ub"]) and np.all(plain["lb"] <= learned["ub"])
    open_pairs = plain["iterations"] > 0
    print("certified", learned["certified"][open_pairs].mean(), "sweeps plain",
          plain["iterations"].sum(), "learned", learned["iterations"].sum(), flush=True)


def test_trained_network_saves_decision_sweeps(gpu, trained):
    """What the network gains: fewer decision sweeps than plain bisection on
    fresh pairs, with at least some predictions certified."""
    pairs = noisy_pairs(400, seed=13, lo=60, hi=300)
    with Batch(pairs, ctx=gpu) as b:
        plain = b.run(rows_per_lane=2, spec_depth=1)
        learned = run_learned(b, trained, rows_per_lane=2, spec_depth=1)
    open_pairs = plain["iterations"] > 0
    print("certified", learned["certified"][open_pairs].mean(), "sweeps plain",
          plain["iterations"].sum(), "learned", learned["iterations"].sum(), flush=True)
    assert learned["certified"][open_pairs].any()
    assert learned["iterations"].sum() < plain["iterations"].sum()


def test_forced_end_certification_matches_oracle(gpu):
    """A network that always answers 1 certifies [ub(1 - eps/2), ub] for every
    open pair: certified exactly where the oracle's decision at that radius is
    unreachable, and then the result is within the stop rule of DFD."""
    net = BracketNet(seed=8, ctx=gpu)
    p = np.zeros(net.get_params().size, dtype=np.float32)
    p[-1] = 30.0
    net.set_params(p)
    net.policy = (0.01, 0.0)
    pairs = noisy_pairs(30, seed=16, lo=60, hi=250)
    with Batch(pairs, ctx=gpu) as b:
        plain = b.run(rows_per_lane=2, spec_depth=1)
        learned = run_learned(b, net, rows_per_lane=2)
    for k, (P, Q) in enumerate(pairs):
        if plain["iterations"][k] == 0:
            continue
        r = plain["dfd"][k] * (1 - 0.5e-9)
        assert learned["certified"][k] == (not O.decide(P, Q, r)), k
        if learned["certified"][k]:
            assert learned["lb"][k] == r and learned["ub"][k] == plain["dfd"][k]


def test_forced_fallback_still_exact(gpu):
    """A network that always answers 0.5 with a hair-thin bracket fails
    certification; bisection of the remaining bracket still lands within
    1e-9 of plain bisection."""
    net = BracketNet(seed=8, ctx=gpu)
    net.set_params(np.zeros(net.get_params().size, dtype=np.float32))
    net.policy = (0.0, 1e-7)
    pairs = noisy_pairs(24, seed=14)
    with Batch(pairs, ctx=gpu) as b:
        plain = b.run(rows_per_lane=2, spec_depth=1)
        learned = run_learned(b, net, rows_per_lane=2)
    open_pairs = plain["iterations"] > 0
    assert open_pairs.any()
    assert learned["certified"][open_pairs].mean() < 0.5
    assert np.all(np.abs(learned["distance"] - plain["distance"]) <= 1e-9 * plain["distance"])
