"""Bounded hand-off waits and input validation.

Every hand-off wait of the sweeps gives up after a bound on its polls
(RFD_SPIN_LIMIT), raises the launch's abort word, and lets every wave run its
remaining steps without waiting, so the kernels end normally and the call
returns RFD_ERR_STATE.  Here the bound is set to one poll from the host, so
the first wait that finds its input not yet ready gives up; nothing faults or
hangs.  Curves holding NaN or infinite coordinates are refused at creation.
"""
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O
from robotfd_amd import Context, Intersection
from robotfd_amd._lib import RFDError
from robotfd_amd.batch import Batch

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def walk(rng, k, d):
    return np.cumsum(rng.standard_normal((k, d)), axis=0)


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


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_non_finite_coordinates_refused(gpu, bad):
    rng = np.random.default_rng(3)
    P, Q = walk(rng, 300, 3), walk(rng, 200, 3)
    for which in ("P", "Q"):
        P2, Q2 = P.copy(), Q.copy()
        (P2 if which == "P" else Q2)[17, 1] = bad
        say("non-finite", bad, which)
        with pytest.raises(RFDError, match="non-finite"):
            with Intersection(P2, Q2, ctx=gpu):
                pass
        with pytest.raises(RFDError, match="pair 1: curve " + which):
            Batch([(P, Q), (P2, Q2)], ctx=gpu)
    # The context still works.
    with Intersection(P, Q, ctx=gpu) as s:
        assert s.Bound_call_gpu(O.endpoint_lb(P, Q))[0] == O.dfd(P, Q)


def test_spin_limit_gives_up_and_reports():
    rng = np.random.default_rng(11)
    # Several bands in every sweep: band 1 waits for band 0 from its first step.
    P, Q = walk(rng, 2300, 3), walk(rng, 900, 3)
    lb, dfd = O.endpoint_lb(P, Q), O.dfd(P, Q)
    r = 0.5 * (lb + dfd)
    tiny = make_ctx(RFD_SPIN_LIMIT=1)
    tiny_sep = make_ctx(RFD_SPIN_LIMIT=1, RFD_FUSED=0)
    try:
        with Intersection(P, Q, ctx=tiny) as s:
            say("fused bound sweep, spin limit 1")
            with pytest.raises(RFDError, match="gave up"):
                s.Bound_call_gpu(lb)
            say("decision sweep, spin limit 1")
            with pytest.raises(RFDError, match="gave up"):
                s.decide_ub(r)
        with Intersection(P, Q, ctx=tiny_sep) as s:
            say("separate DFD sweep, spin limit 1")
            with pytest.raises(RFDError, match="gave up"):
                s.Bound_call_gpu(lb)
        say("batch, spin limit 1")
        with Batch([(P, Q), (Q, P)], ctx=tiny) as b:
            with pytest.raises(RFDError, match="gave up"):
                b.run(rows_per_lane=1)
    finally:
        tiny.close()
        tiny_sep.close()
    # Every kernel ended: a context with the default bound computes exactly.
    ok = make_ctx()
    try:
        with Intersection(P, Q, ctx=ok) as s:
            dfd_g, ve_g = s.Bound_call_gpu(lb)
            assert dfd_g == dfd and ve_g == O.vefd(P, Q, lb)
            assert s.decide_ub(r) == O.decide_ub(P, Q, r)
    finally:
        ok.close()
