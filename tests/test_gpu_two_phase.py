"""The two-phase decision form (the reference's interval_calculation<D> +
reachability, .cu:674-958) on the GPU against the oracle's restatement,
bit-exact: terminal rightUB at radii where the two interval forms decide
differently, zero-length segments (NaN intervals), the distance trajectory,
the reference's choice of form, and batches."""
import math
import sys

import numpy as np
import pytest

from oracle import oracle as O
from robotfd_amd import Context, Curve, Intersection, compute_distance_parallel
from robotfd_amd.batch import Batch
from tests import independent as I

pytestmark = pytest.mark.gpu

FLT_EPSILON = float(np.finfo(np.float32).eps)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def walk(rng, k, d):
    return np.cumsum(rng.standard_normal((k, d)), axis=0)


def same(a, b):
    return (math.isnan(a) and math.isnan(b)) or a == b


@pytest.fixture(scope="module")
def ctx():
    c = Context()
    yield c
    c.close()


def diverging_cases(count, seed):
    """(P, Q, r) at radii a few ulps from a critical value where the oracle's
    two forms give different terminal rightUB."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(400):
        m, n = rng.integers(3, 9, 2)
        P, Q = walk(rng, m, 2 + len(out) % 2), walk(rng, n, 2 + len(out) % 2)
        lb = O.endpoint_lb(P, Q)
        for c in I.critical_values(P, Q):
            if c <= lb:
                continue
            for k in range(-3, 4):
                r = float(c) * (1.0 + k * 2.0 ** -52)
                if not same(O.decide_ub(P, Q, r), O.decide_ub(P, Q, r, form=1)):
                    out.append((P, Q, r))
                    break
            if len(out) >= count:
                return out
    return out


def test_two_phase_decision_where_forms_differ(ctx):
    cases = diverging_cases(12, 5)
    assert len(cases) >= 4, "no radius where the two forms decide differently"
    for P, Q, r in cases:
        with Intersection(P, Q, ctx=ctx, decide_form="two-phase") as s:
            assert s.effective_form == "two-phase"
            say("two-phase", P.shape, Q.shape, r)
            assert same(s.decide_ub(r), O.decide_ub(P, Q, r, form=1))
        with Intersection(P, Q, ctx=ctx, decide_form="fused") as s:
            assert same(s.decide_ub(r), O.decide_ub(P, Q, r))


@pytest.mark.parametrize("dims", [2, 3])
def test_two_phase_decision_shapes(ctx, dims):
    rng = np.random.default_rng(40 + dims)
    for m, n in [(2, 2), (3, 7), (65, 66), (130, 200), (300, 257), (700, 50), (40, 700)]:
        P, Q = walk(rng, m, dims), walk(rng, n, dims)
        lb, ub = O.endpoint_lb(P, Q), O.dfd(P, Q)
        say("two-phase shapes", dims, (m, n))
        with Intersection(P, Q, ctx=ctx, decide_form="two-phase") as s:
            for r in np.linspace(lb, ub * 1.05 + 1e-12, 7):
                assert same(s.decide_ub(r), O.decide_ub(P, Q, r, form=1)), (m, n, r)


def test_two_phase_zero_length_segments(ctx):
    """Repeated points: the two-phase intervals of a zero-length segment are
    NaN, and the reachability rules carry them on as the reference's do."""
    rng = np.random.default_rng(8)
    P, Q = walk(rng, 150, 2), walk(rng, 170, 2)
    P[40:45] = P[40]
    Q[::7] = Q[3]
    lb, ub = O.endpoint_lb(P, Q), O.dfd(P, Q)
    with Intersection(P, Q, ctx=ctx, decide_form="two-phase") as s:
        for r in np.linspace(lb, ub * 1.05, 9):
            assert same(s.decide_ub(r), O.decide_ub(P, Q, r, form=1)), r


def test_two_phase_distance_trajectory(ctx):
    for seed in range(3):
        rng = np.random.default_rng(60 + seed)
        P, Q = walk(rng, 300 + 50 * seed, 2), walk(rng, 280, 2)
        ref = O.distance(P, Q, stop_mode=1, eps=1e-9, form=1)
        with Intersection(P, Q, ctx=ctx) as s:
            got = s.distance(stop="rel", eps=1e-9, decide_form="two-phase")
        assert got["iterations"] == ref["iterations"] and got["distance"] == ref["distance"]


def test_reference_rule_and_force_oversize(ctx):
    """The reference takes the two-phase path when its interval arrays fit in
    free device memory and forceOversize is not set (.cu:87-96)."""
    rng = np.random.default_rng(3)
    P, Q = walk(rng, 500, 2), walk(rng, 600, 2)
    with Intersection(P, Q, ctx=ctx, decide_form="reference") as s:
        assert s.effective_form == "two-phase"
    with Intersection(P, Q, forceOversize=True, ctx=ctx, decide_form="reference") as s:
        assert s.effective_form == "fused"
    ref = O.distance(P, Q, stop_mode=0, eps=FLT_EPSILON, form=1)
    got = compute_distance_parallel(Curve(P), Curve(Q))
    assert got == ref["distance"]
    ref_f = O.distance(P, Q, stop_mode=0, eps=FLT_EPSILON)
    assert compute_distance_parallel(Curve(P), Curve(Q), True) == ref_f["distance"]


@pytest.mark.parametrize("form", ["two-phase", "reference"])
def test_two_phase_batch(ctx, form):
    rng = np.random.default_rng(77)
    pairs = [(walk(rng, int(rng.integers(50, 400)), d), walk(rng, int(rng.integers(50, 400)), d))
             for d in (2, 3, 2, 3, 2, 3)]
    b = Batch(pairs, ctx=ctx)
    out = b.run(stop="rel", eps=1e-9, decide_form=form)
    for k, (P, Q) in enumerate(pairs):
        ref = O.distance(P, Q, stop_mode=1, eps=1e-9, form=1)
        assert out["distance"][k] == ref["distance"], k
    b.close()
