"""The fused bound sweep (DFD and VE-FD in one pass, weights computed by
producer waves off the min/max chain) against the CPU oracle, bit-exact, for
every instantiated configuration, workgroups taking several bands each, and
the separate band sweeps it replaces."""
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O
from robotfd_amd import Context, Intersection, workloads as W

pytestmark = pytest.mark.gpu

SHAPES = [(2, 2), (2, 3), (3, 2), (5, 7), (63, 40), (64, 64), (65, 66), (128, 70), (129, 300),
          (200, 129), (700, 50), (40, 700), (257, 100), (520, 77), (1100, 90)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def walk(rng, k, d, sigma=1.0):
    return np.cumsum(rng.standard_normal((k, d)) * sigma, axis=0)


def make_ctx(**env):
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


@pytest.fixture(scope="module")
def contexts():
    made = {}

    def get(**env):
        key = tuple(sorted(env.items()))
        if key not in made:
            made[key] = make_ctx(**env)
        return made[key]
    yield get
    for c in made.values():
        c.close()


def check_pair(ctx, P, Q):
    lb = O.endpoint_lb(P, Q)
    with Intersection(P, Q, ctx=ctx) as s:
        dfd, ve = s.Bound_call_gpu(lb)
    assert dfd == O.dfd(P, Q)
    assert ve == O.vefd(P, Q, lb)


# The DFD-only form (one weight per cell, no VE-FD), which computes the DFD of
# the pruned bound path and of the FP32 one in configurations without the
# mixed form, and the mixed form (DFD in FP64, VE-FD in FP32: configurations
# 77, 107 and 1310), in every configuration they are built in.
@pytest.mark.parametrize("cfg", [42, 77, 107, 1310])
@pytest.mark.parametrize("dims", [2, 3])
def test_dfd_only_and_mixed_configs_bit_exact(contexts, cfg, dims):
    ctx = contexts(RFD_BOUNDS=cfg)
    for m, n in SHAPES:
        rng = np.random.default_rng(m * 1000 + n + dims + cfg + 7)
        P, Q = walk(rng, m, dims), walk(rng, n, dims)
        say("dfd only", cfg, dims, (m, n))
        lb = O.endpoint_lb(P, Q)
        with Intersection(P, Q, ctx=ctx) as s:
            dfd, _ = s.Bound_call_gpu(lb, run_vefd=False)
            assert dfd == O.dfd(P, Q)
            # The FP32 bound path: DFD in FP64, VE-FD in FP32.
            ctx.reset_stats()
            dfd32, ve32 = s.Bound_call_gpu(lb, fp32=True)
            assert dfd32 == dfd and ve32 == O.vefd(P, Q, lb, fp32=True)
            if m >= 2 and n >= 2 and cfg != 42:
                assert ctx.stats()["bounds_mixed"]["launches"] == 1


# 20000 + c: the same configuration at two columns per step (even n; odd n
# runs the one-column kernel).
@pytest.mark.parametrize("cfg", [21, 42, 62, 63, 72, 73, 77, 83, 84, 107, 136, 137, 1012, 1310, 1313,
                                 20077, 20107, 20137, 21310])
@pytest.mark.parametrize("dims", [2, 3])
def test_fused_configs_bit_exact(contexts, cfg, dims):
    ctx = contexts(RFD_BOUNDS=cfg)
    ctx.reset_stats()
    for m, n in SHAPES:
        rng = np.random.default_rng(m * 1000 + n + dims + cfg)
        P, Q = walk(rng, m, dims), walk(rng, n, dims)
        say("fused", cfg, dims, (m, n))
        check_pair(ctx, P, Q)
    st = ctx.stats()
    assert st["bounds_f64"]["launches"] == len(SHAPES) and "dfd" not in st


@pytest.mark.parametrize("cfg", [42, 77, 107, 1310, 21310])
def test_fused_random_shapes_many_bands(contexts, cfg):
    """Random shapes up to 6000 rows (up to eight bands of 13 rows per lane):
    the band hand-offs, the first band's row -1, and the last band stopping at
    the step where its terminal row reaches column n - 1, for terminal rows in
    any lane."""
    ctx = contexts(RFD_BOUNDS=cfg)
    rng = np.random.default_rng(cfg)
    for _ in range(12):
        m, n = int(rng.integers(2, 6000)), int(rng.integers(2, 2000))
        dims = int(rng.integers(2, 4))
        P, Q = walk(rng, m, dims), walk(rng, n, dims)
        say("random", cfg, dims, (m, n))
        check_pair(ctx, P, Q)


@pytest.mark.parametrize("cfg", [42, 84, 136, 137, 20137, 21310])
def test_fused_workgroups_take_several_bands(contexts, cfg):
    ctx = contexts(RFD_BOUNDS=cfg, RFD_MAX_WAVES=3)
    rng = np.random.default_rng(5)
    n = 334 if cfg >= 20000 else 333
    P, Q = walk(rng, 64 * 8 * 7 + 5, 3), walk(rng, n, 3)
    say("wrap", cfg)
    check_pair(ctx, P, Q)


@pytest.mark.parametrize("m,n", [(2, 2), (64, 130), (833, 2), (832 * 3 + 1, 4096), (5000, 1000)])
def test_two_column_sweep_even_shapes(contexts, m, n):
    """Two columns per step: the terminal column is the second of its step,
    the band fill is 63 steps of two columns, the hand-off rings carry two
    columns per step."""
    ctx = contexts(RFD_BOUNDS=21310)
    rng = np.random.default_rng(m + 7 * n)
    P, Q = walk(rng, m, 3), walk(rng, n, 3)
    say("two-column", (m, n))
    check_pair(ctx, P, Q)


def test_fused_noisy_copy_and_collapsed_bracket(contexts):
    ctx = contexts()
    for seed in range(3):
        P, Q = W.robot_pair(900 + 50 * seed, 1000, seed=seed)
        say("robot", seed)
        check_pair(ctx, P, Q)
    P, Q = W.walk_pair(3000, 2500, dims=3, seed=0)
    say("walk 3000x2500")
    check_pair(ctx, P, Q)


def test_fused_degenerate_segments(contexts):
    """Repeated points give zero-length segments: the projection divides 0 by 0
    and the clamp must take it to 0 as the oracle does."""
    ctx = contexts()
    rng = np.random.default_rng(8)
    P, Q = walk(rng, 150, 2), walk(rng, 170, 2)
    P[40:45] = P[40]
    Q[::7] = Q[3]
    check_pair(ctx, P, Q)


def test_separate_sweeps_still_bit_exact(contexts):
    ctx = contexts(RFD_FUSED=0)
    ctx.reset_stats()
    rng = np.random.default_rng(12)
    for m, n in [(65, 66), (257, 100), (40, 700)]:
        P, Q = walk(rng, m, 3), walk(rng, n, 3)
        check_pair(ctx, P, Q)
    st = ctx.stats()
    assert "bounds_f64" not in st and st["dfd"]["launches"] == 3


def test_fused_distance_matches_oracle_trajectory(contexts):
    ctx = contexts()
    P, Q = W.robot_pair(800, 700, seed=4)
    ref = O.distance(P, Q, stop_mode=1, eps=1e-9)
    with Intersection(P, Q, ctx=ctx) as s:
        got = s.distance(stop="rel", eps=1e-9)
    assert got["dfd"] == ref["dfd"] and got["vefd"] == ref["vefd"]
    assert got["iterations"] == ref["iterations"] and got["distance"] == ref["distance"]


DEC_SHAPES = [(2, 2), (2, 3), (3, 2), (5, 7), (65, 66), (130, 200), (300, 257), (700, 50), (40, 700),
              (1100, 90)]


@pytest.mark.parametrize("rows", [1, 2, 3, 4])
@pytest.mark.parametrize("dims", [2, 3])
def test_decision_rows_per_lane_bit_exact(contexts, rows, dims):
    """The decision sweep's terminal rightUB against the oracle's at every
    rows-per-lane choice, at radii from the endpoint bound to past DFD
    (band_sweep's decision, RFD_DEC=0)."""
    ctx = contexts(RFD_ROWS=rows, RFD_DEC=0)
    ctx.reset_stats()
    calls = 0
    for m, n in DEC_SHAPES:
        rng = np.random.default_rng(7 * m + n + dims + rows)
        P, Q = walk(rng, m, dims), walk(rng, n, dims)
        lb, ub = O.endpoint_lb(P, Q), O.dfd(P, Q)
        say("decision", rows, dims, (m, n))
        with Intersection(P, Q, ctx=ctx) as s:
            for r in np.linspace(lb, ub * 1.05 + 1e-12, 7):
                assert s.decide_ub(r) == O.decide_ub(P, Q, r), (m, n, r)
                calls += 1
    assert ctx.stats()["decide_f64"]["launches"] == calls


@pytest.mark.parametrize("rows", [1, 3])
def test_decision_near_critical_value_and_wrap(contexts, rows):
    """Waves taking several bands each (3 waves), at radii within 1e-9 and
    1e-6 of the distance (band_sweep's decision, RFD_DEC=0)."""
    ctx = contexts(RFD_ROWS=rows, RFD_MAX_WAVES=3, RFD_DEC=0)
    P, Q = W.robot_pair(64 * 13 * 4 + 7, 700, seed=2)
    res = O.distance(P, Q, stop_mode=1, eps=1e-12)
    fd = res["distance"]
    with Intersection(P, Q, ctx=ctx) as s:
        for f in (1 - 1e-6, 1 - 1e-9, 1 + 1e-9, 1 + 1e-6):
            assert s.decide_ub(fd * f) == O.decide_ub(P, Q, fd * f), f


def test_decision_without_fused_bounds_bit_exact(contexts):
    """With the separate bound sweeps (RFD_FUSED=0) decisions are unchanged."""
    ctx = contexts(RFD_FUSED=0)
    rng = np.random.default_rng(21)
    P, Q = walk(rng, 300, 3), walk(rng, 257, 3)
    lb, ub = O.endpoint_lb(P, Q), O.dfd(P, Q)
    with Intersection(P, Q, ctx=ctx) as s:
        for r in np.linspace(lb, ub, 5):
            assert s.decide_ub(r) == O.decide_ub(P, Q, r)
