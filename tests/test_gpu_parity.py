"""GPU parity: the gfx950 kernels against the CPU oracle, through the C ABI.

Bars: DFD, VE-FD and the decision's terminal rightUB are bit-exact (same
formula, same FMA placement, correctly rounded div/sqrt on both sides); the
bisection trajectory is identical, so the final distance is identical.
"""
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O
from robotfd_amd import Context, Intersection, workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def walk(rng, k, d, sigma=1.0):
    return np.cumsum(rng.standard_normal((k, d)) * sigma, axis=0)


SHAPES = [(2, 2), (2, 3), (3, 2), (5, 7), (63, 40), (64, 64), (65, 66), (66, 65), (128, 70),
          (129, 300), (200, 129), (700, 50), (40, 700), (257, 100), (520, 77)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("dims", [2, 3])
def test_bounds_bit_exact(gpu, shape, dims):
    m, n = shape
    rng = np.random.default_rng(m * 1000 + n + dims)
    P, Q = walk(rng, m, dims), walk(rng, n, dims)
    say("bounds", shape, dims)
    with Intersection(P, Q, ctx=gpu) as s:
        lb = O.endpoint_lb(P, Q)
        dfd, ve = s.Bound_call_gpu(lb)
        assert dfd == O.dfd(P, Q)
        assert ve == O.vefd(P, Q, lb)
        _, ve32 = s.Bound_call_gpu(lb, fp32=True)
        assert ve32 == O.vefd(P, Q, lb, fp32=True)


@pytest.mark.parametrize("shape", [(1, 1), (1, 5), (7, 1)])
def test_dfd_single_point_curves(gpu, shape):
    rng = np.random.default_rng(3)
    P, Q = walk(rng, shape[0], 2), walk(rng, shape[1], 2)
    with Intersection(P, Q, ctx=gpu) as s:
        dfd, _ = s.Bound_call_gpu(0.0, run_vefd=False)
        assert dfd == O.dfd(P, Q)


@pytest.mark.parametrize("shape", [(2, 2), (9, 12), (65, 66), (130, 200), (300, 257)])
@pytest.mark.parametrize("dims", [2, 3])
def test_decision_terminal_ub_bit_exact(gpu, shape, dims):
    m, n = shape
    rng = np.random.default_rng(7 * m + n + dims)
    P, Q = walk(rng, m, dims), walk(rng, n, dims)
    lb, ub = O.endpoint_lb(P, Q), O.dfd(P, Q)
    say("decide", shape, dims)
    with Intersection(P, Q, ctx=gpu) as s:
        for r in np.linspace(lb, ub * 1.05 + 1e-12, 9):
            assert s.decide_ub(r) == O.decide_ub(P, Q, r)
            assert s.decide_ub(r, fp32=True) == O.decide_ub(P, Q, r, fp32=True)


def test_decision_on_noisy_copy_near_critical_value(gpu):
    P, Q = W.robot_pair(600, 700)
    res = O.distance(P, Q, stop_mode=1, eps=1e-12)
    fd = res["distance"]
    with Intersection(P, Q, ctx=gpu) as s:
        for f in (1 - 1e-6, 1 - 1e-9, 1 + 1e-9, 1 + 1e-6):
            assert s.decide_ub(fd * f) == O.decide_ub(P, Q, fd * f)


@pytest.mark.parametrize("stop,eps", [("abs", float(np.finfo(np.float32).eps)), ("rel", 1e-9)])
def test_distance_matches_oracle_small_2d(gpu, stop, eps):
    P, Q = W.small_2d()
    say("distance small_2d", stop)
    ref = O.distance(P, Q, stop_mode=0 if stop == "abs" else 1, eps=eps)
    from robotfd_amd import distance
    got = distance(P, Q, stop=stop, eps=eps, ctx=gpu)
    assert got["dfd"] == ref["dfd"]
    assert got["vefd"] == ref["vefd"]
    assert got["iterations"] == ref["iterations"]
    assert got["distance"] == ref["distance"]


@pytest.mark.parametrize("seed", range(4))
def test_distance_trajectory_noisy_pairs(gpu, seed):
    P, Q = W.robot_pair(500 + 37 * seed, 650, seed=seed)
    ref = O.distance(P, Q, stop_mode=1, eps=1e-9)
    say("distance robot", seed, ref["iterations"])
    with Intersection(P, Q, ctx=gpu) as s:
        got = s.distance(stop="rel", eps=1e-9)
        assert got["dfd"] == ref["dfd"] and got["vefd"] == ref["vefd"]
        assert got["iterations"] == ref["iterations"]
        assert got["distance"] == ref["distance"]
        # every split's decision equals the oracle's
        for split, dec in zip(ref["splits"], ref["decisions"]):
            assert s.intersection_interval_call_gpu(split) == bool(dec)


def test_band_queue_wraps_with_few_waves(gpu):
    """More bands than waves: each wave takes several bands in order."""
    os.environ["RFD_MAX_WAVES"] = "3"
    try:
        ctx = Context()
    finally:
        del os.environ["RFD_MAX_WAVES"]
    rng = np.random.default_rng(99)
    P, Q = walk(rng, 64 * 7 + 5, 3), walk(rng, 333, 3)
    lb = O.endpoint_lb(P, Q)
    with Intersection(P, Q, ctx=ctx) as s:
        dfd, ve = s.Bound_call_gpu(lb)
        assert dfd == O.dfd(P, Q) and ve == O.vefd(P, Q, lb)
        ub = O.dfd(P, Q)
        for r in np.linspace(lb, ub, 5):
            assert s.decide_ub(r) == O.decide_ub(P, Q, r)
    ctx.close()


def test_mixed_precision_result_within_tolerance(gpu):
    P, Q = W.robot_pair(900, 1000, seed=5)
    ref = O.distance(P, Q, stop_mode=1, eps=1e-9)
    from robotfd_amd import distance
    got = distance(P, Q, stop="rel", eps=1e-9, precision="mixed", ctx=gpu)
    assert abs(got["distance"] - ref["distance"]) <= 1e-9 * ref["distance"]
    assert got["lb"] <= ref["ub"] and ref["lb"] <= got["ub"]


@pytest.fixture(scope="module")
def pruning_ctx():
    """A context whose FP64 bounds run as separate sweeps with VE-FD pruning."""
    os.environ["RFD_PRUNE"] = "1"
    try:
        ctx = Context()
    finally:
        del os.environ["RFD_PRUNE"]
    yield ctx
    ctx.close()


@pytest.mark.parametrize("kind", ["walk2", "walk3", "robot"])
def test_pruned_and_skipped_sweeps_bit_exact(pruning_ctx, kind):
    """Sizes where VE-FD pruning (cells above DFD^2) and 16-step dead-block
    skipping both occur; results must not change."""
    gpu = pruning_ctx
    gpu.reset_stats()
    if kind == "robot":
        P, Q = W.robot_pair(1500, 1800, seed=3)
    else:
        rng = np.random.default_rng(17 if kind == "walk2" else 18)
        d = 2 if kind == "walk2" else 3
        P, Q = walk(rng, 1400, d), walk(rng, 1700, d)
    say("pruned", kind)
    lb = O.endpoint_lb(P, Q)
    dfd = O.dfd(P, Q)
    with Intersection(P, Q, ctx=gpu) as s:
        gd, gv = s.Bound_call_gpu(lb)
        assert gd == dfd
        assert gv == O.vefd(P, Q, lb)
        _, gv32 = s.Bound_call_gpu(lb, fp32=True)
        assert gv32 == O.vefd(P, Q, lb, fp32=True)
        st = gpu.stats()
        assert st["vefd_f64"]["evaluated"] < st["vefd_f64"]["cells"]
        for r in (lb, 0.5 * (lb + gv), gv * (1 - 1e-9), gv, 0.5 * (gv + dfd), dfd):
            assert s.decide_ub(r) == O.decide_ub(P, Q, r)


@pytest.mark.parametrize("rows", [1, 2, 3, 4])
def test_decision_dead_skips_across_bands(rows):
    """Decisions below and above the distance on pairs of several bands, at 1,
    2 and 4 rows per lane, where most cells are unreachable: whole 16-step
    blocks can be dead, skipped, and handed to the next band as dead columns.
    The terminal rightUB stays bit-exact; dead cells do occur (fewer cells
    evaluated than swept) and whole dead blocks are skipped (counted)."""
    # band_sweep's decision (RFD_DEC=0; the windowed sweep has its own tests).
    os.environ["RFD_ROWS"] = str(rows)
    os.environ["RFD_DEC"] = "0"
    try:
        ctx = Context()
    finally:
        del os.environ["RFD_ROWS"]
        del os.environ["RFD_DEC"]
    try:
        rng = np.random.default_rng(40 + rows)
        skipped = 0
        for d in (2, 3):
            P, Q = walk(rng, 1100, d), walk(rng, 2300, d)
            # A second pair whose endpoints coincide: its endpoint bound is 0,
            # and at radii well below DFD almost every cell is unreachable.
            ramp = np.linspace(0.0, 1.0, len(Q))[:, None]
            Qm = Q + (P[0] - Q[0]) * (1 - ramp) + (P[-1] - Q[-1]) * ramp
            for PP, QQ in ((P, Q), (P, Qm)):
                lb = O.endpoint_lb(PP, QQ)
                dfd = O.dfd(PP, QQ)
                say("dead skips", rows, d, lb)
                with Intersection(PP, QQ, ctx=ctx) as s:
                    s.data_move()
                    ctx.reset_stats()
                    for f in (0.0, 0.05, 0.2, 0.5, 0.8, 0.95, 1.0, 1.05):
                        r = lb + f * (dfd - lb)
                        assert s.decide_ub(r) == O.decide_ub(PP, QQ, r), (d, f)
                    st = ctx.stats()["decide_f64"]
                    assert st["evaluated"] < st["cells"]
                    skipped += st["skipped_blocks"]
        assert skipped > 0
    finally:
        ctx.close()
