"""The tighter upper bound (DFD of both curves refined at segment midpoints)
against the oracle: bit-exact, between FD and DFD, and the bisection that
starts from it reproduces the oracle's trajectory from the same bracket."""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from robotfd_amd import Intersection, workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def walk(rng, k, d):
    return np.cumsum(rng.standard_normal((k, d)), axis=0)


@pytest.mark.parametrize("levels", [1, 2])
@pytest.mark.parametrize("dims", [2, 3])
def test_refined_ub_bit_exact_and_between_fd_and_dfd(gpu, levels, dims):
    rng = np.random.default_rng(11 * levels + dims)
    for m, n in ((2, 2), (5, 7), (130, 90), (700, 260)):
        P, Q = walk(rng, m, dims), walk(rng, n, dims)
        say("refined ub", levels, dims, (m, n))
        with Intersection(P, Q, ctx=gpu) as s:
            ub = s.refined_ub(levels)
        assert ub == O.dfd_refined(P, Q, levels), (m, n)
        fd = O.distance(P, Q, stop_mode=1, eps=1e-12)["distance"]
        assert fd * (1 - 1e-12) <= ub <= O.dfd(P, Q)


def oracle_bisection(P, Q, lb, ub, eps):
    """frechet.cpp:41-63 from a given bracket, rel stop, oracle decisions."""
    it = 0
    while ub - lb > eps * ub:
        split = (ub + lb) / 2.0
        if O.decide(P, Q, split):
            ub = split
        else:
            lb = split
        it += 1
    return (ub + lb) / 2.0, it


@pytest.mark.parametrize("seed", range(3))
def test_distance_from_the_refined_bracket(gpu, seed):
    P, Q = W.robot_pair(700 + 50 * seed, 500, seed=seed)
    plain = None
    with Intersection(P, Q, ctx=gpu) as s:
        plain = s.distance(stop="rel", eps=1e-9)
        for levels in (1, 2):
            got = s.distance(stop="rel", eps=1e-9, refine_ub=levels)
            lb = O.vefd(P, Q, O.endpoint_lb(P, Q))
            dfd = O.dfd(P, Q)
            if dfd - lb > 1e-9 * dfd:
                ub0 = max(lb, min(dfd, O.dfd_refined(P, Q, levels)))
            else:
                ub0 = dfd
            want, it = oracle_bisection(P, Q, lb, ub0, 1e-9)
            say("seed", seed, "levels", levels, "steps", it, "plain steps", plain["iterations"])
            assert got["distance"] == want and got["iterations"] == it
            assert got["iterations"] <= plain["iterations"]
            assert abs(got["distance"] - plain["distance"]) <= 1e-9 * plain["distance"]
