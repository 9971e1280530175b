"""The windowed decision sweep (dec_sweep: producer waves compute the free
intervals, a chain wave applies the reachability rules, bands start at their
predecessor's first reachable column and stop once they and their inputs are
unreachable) against the oracle, bit-exact in the terminal rightUB, for every
configuration, in both decision forms, with bands that die early (radii just
above the endpoint bound), bands that never start, more bands than
workgroups, and against band_sweep.  Reference: reachable<D>, .cu:960-1256;
the two-phase form, .cu:674-958.
"""
import math
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O
from robotfd_amd import Context, Intersection, workloads as W

pytestmark = pytest.mark.gpu

CONFIGS = [12, 16, 23, 26, 36, 46]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def walk(rng, k, d, sigma=1.0):
    return np.cumsum(rng.standard_normal((k, d)) * sigma, axis=0)


def same(a, b):
    return (math.isnan(a) and math.isnan(b)) or a == b


def context(**env):
    """A context created with the given RFD_* settings."""
    old = {k: os.environ.get(k) for k in env}
    os.environ.update({k: str(v) for k, v in env.items()})
    try:
        return Context()
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


@pytest.fixture(scope="module", params=CONFIGS)
def dctx(request):
    c = context(RFD_DEC=request.param)
    yield request.param, c
    c.close()


def radii(P, Q, k=7):
    lb, ub = O.endpoint_lb(P, Q), O.dfd(P, Q)
    # Just above the endpoint bound (regions that die within a few bands),
    # across the bracket, and above DFD (everything reachable).
    return [lb * (1 + 1e-12), lb + 1e-3 * (ub - lb)] + list(np.linspace(lb, ub * 1.05 + 1e-12, k))


SHAPES = [(2, 2), (3, 70), (70, 3), (65, 66), (129, 200), (300, 257), (700, 90), (1500, 400)]


@pytest.mark.parametrize("dims", [2, 3])
def test_windowed_decision_bit_exact(dctx, dims):
    cfg, ctx = dctx
    rng = np.random.default_rng(1000 + cfg + dims)
    for m, n in SHAPES:
        P, Q = walk(rng, m, dims), walk(rng, n, dims)
        say("windowed decision", cfg, dims, (m, n))
        with Intersection(P, Q, ctx=ctx) as s:
            for r in radii(P, Q):
                assert s.decide_ub(r) == O.decide_ub(P, Q, r), (cfg, m, n, r)


@pytest.mark.parametrize("dims", [2, 3])
def test_windowed_decision_two_phase(dctx, dims):
    cfg, ctx = dctx
    rng = np.random.default_rng(2000 + cfg + dims)
    for m, n in [(2, 2), (65, 66), (300, 257), (40, 700)]:
        P, Q = walk(rng, m, dims), walk(rng, n, dims)
        say("windowed two-phase", cfg, dims, (m, n))
        with Intersection(P, Q, ctx=ctx, decide_form="two-phase") as s:
            for r in radii(P, Q, 5):
                assert same(s.decide_ub(r), O.decide_ub(P, Q, r, form=1)), (cfg, m, n, r)
    # Zero-length segments: NaN interval ends are not unreachable.
    P, Q = walk(rng, 150, dims), walk(rng, 170, dims)
    P[40:45] = P[40]
    Q[::7] = Q[3]
    with Intersection(P, Q, ctx=ctx, decide_form="two-phase") as s:
        for r in radii(P, Q, 5):
            assert same(s.decide_ub(r), O.decide_ub(P, Q, r, form=1)), (cfg, r)


def test_windowed_decision_noisy_copy_near_critical_value(dctx):
    cfg, ctx = dctx
    P, Q = W.robot_pair(3000, 3500, seed=2)
    fd = O.distance(P, Q, stop_mode=1, eps=1e-12)["distance"]
    say("windowed noisy copy", cfg)
    with Intersection(P, Q, ctx=ctx) as s:
        for f in (1 - 1e-6, 1 - 1e-9, 1 + 1e-9, 1 + 1e-6):
            assert s.decide_ub(fd * f) == O.decide_ub(P, Q, fd * f), (cfg, f)


@pytest.mark.parametrize("cfg", [16, 26, 46])
def test_windowed_decision_more_bands_than_workgroups(cfg):
    """Two workgroups take every band in turn (in-order queue)."""
    ctx = context(RFD_DEC=cfg, RFD_MAX_WAVES=2)
    try:
        rng = np.random.default_rng(cfg)
        P, Q = walk(rng, 2000, 3), walk(rng, 600, 3)
        say("windowed, two workgroups", cfg)
        with Intersection(P, Q, ctx=ctx) as s:
            for r in radii(P, Q, 5):
                assert s.decide_ub(r) == O.decide_ub(P, Q, r), (cfg, r)
    finally:
        ctx.close()


def test_windowed_decision_distance_trajectory(dctx):
    """The bisection through the windowed sweep lands on the oracle's
    distance in the oracle's number of steps."""
    cfg, ctx = dctx
    P, Q = W.robot_pair(900, 1000, seed=4)
    ref = O.distance(P, Q, stop_mode=1, eps=1e-9)
    with Intersection(P, Q, ctx=ctx) as s:
        got = s.distance(stop="rel", eps=1e-9)
    assert got["distance"] == ref["distance"] and got["iterations"] == ref["iterations"]


def test_windowed_matches_band_sweep_and_skips_work():
    """The windowed sweep and band_sweep agree, and on a radius just above
    the endpoint bound the windowed sweep runs a fraction of the steps."""
    win, old = Context(), context(RFD_DEC=0)
    try:
        rng = np.random.default_rng(11)
        P, Q = walk(rng, 4000, 3), walk(rng, 3000, 3)
        rs = radii(P, Q, 5)
        with Intersection(P, Q, ctx=win) as a, Intersection(P, Q, ctx=old) as b:
            for r in rs:
                assert a.decide_ub(r) == b.decide_ub(r) == O.decide_ub(P, Q, r), r
        st = win.stats()["decide_f64"]
        say("windowed stats", st)
        assert st["skipped_blocks"] > 0
    finally:
        win.close()
        old.close()
