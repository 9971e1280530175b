"""The committed full-size reference values (tests/golden/fullsize.json,
batch4096.npz, written by make_fullsize_golden.py) against the oracle where
that is cheap, and against the relations the bisection fixes everywhere:

* the smallest batch pairs recomputed from scratch (or_distance, single
  thread) give the fixture's distance, bounds, bracket and steps bit for bit;
* every entry satisfies frechet.cpp:18-75: the bracket starts at
  [VE-FD, DFD], the distance is its midpoint, steps are zero exactly when the
  bracket starts closed, and the final bracket is closed under the rel 1e-9
  stop;
* the long pairs' inputs regenerate to the recorded digests, and their
  decisions at DFD (1 -/+ 0.5e-9) bracket DFD as the distance requires.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from robotfd_amd import workloads as W

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REL = 1e-9


@pytest.fixture(scope="module")
def batch():
    return np.load(os.path.join(GOLDEN, "batch4096.npz"), allow_pickle=False)


def test_batch_fixture_matches_oracle_on_smallest_pairs(batch):
    pairs = W.batch_pairs(4096, seed=0)
    sizes = np.array([(len(P), len(Q), P.shape[1]) for P, Q in pairs])
    assert (sizes == batch["sizes"]).all()
    order = np.argsort(sizes[:, 0] * sizes[:, 1], kind="stable")
    # The six smallest were the generator's own self-check; take others.
    for k in order[6:14]:
        P, Q = pairs[k]
        ref = O.distance(P, Q, stop_mode=1, eps=REL, max_trace=1)
        assert ref["distance"] == batch["distance"][k], k
        assert ref["iterations"] == batch["iterations"][k], k
        assert O.dfd(P, Q) == batch["dfd"][k]
        assert O.endpoint_lb(P, Q) == batch["endpoint_lb"][k]


def test_batch_fixture_obeys_the_bisection(batch):
    lb, ub, d = batch["lb"], batch["ub"], batch["distance"]
    assert (d == (ub + lb) / 2.0).all()
    # Steps are zero exactly when [VE-FD, DFD] already meets the stop rule.
    start_open = batch["dfd"] - batch["vefd"] > REL * batch["dfd"]
    assert ((batch["iterations"] > 0) == start_open).all()
    # A closed start keeps its bracket; an open one ends closed inside it.
    assert (lb[~start_open] == batch["vefd"][~start_open]).all()
    assert (ub[~start_open] == batch["dfd"][~start_open]).all()
    assert (ub - lb <= REL * ub)[start_open].all()
    assert (lb >= batch["vefd"])[start_open].all() and (ub <= batch["dfd"])[start_open].all()
    # Each step halves the bracket (up to the midpoints' rounding, a few ulps
    # of the distance against a gap of 1e-9 of it), and the loop stops at the
    # first step that meets the rule: one step fewer would not.
    gap0 = (batch["dfd"] - batch["vefd"])[start_open]
    gap = (ub - lb)[start_open]
    k = batch["iterations"][start_open]
    assert np.allclose(2.0 ** k * gap, gap0, rtol=1e-6, atol=0)
    assert (2.0 * gap > REL * ub[start_open] * (1 - 1e-6)).all()
    # The endpoint bound rounds at most an ulp above DFD.  It does so on 36
    # pairs, whose DFD is attained at an end pair: the bound follows
    # Point::dist_sqr's operation order (geometry_basics.hpp:72-81), DFD its
    # FMA chain.  VE-FD starts from the bound, so it ends above DFD there too.
    over = batch["endpoint_lb"] > batch["dfd"]
    assert (batch["endpoint_lb"] <= batch["dfd"] * (1 + 4.5e-16)).all()
    assert over.sum() == 36
    assert (batch["vefd"][over] > batch["dfd"][over]).all()
    assert (batch["vefd"] > batch["dfd"]).sum() == 36


@pytest.mark.parametrize("key", ["c2", "c3"])
def test_long_pair_fixture(key):
    g = json.load(open(os.path.join(GOLDEN, "fullsize.json")))[key]
    P, Q = W.robot_pair(100_000, 120_000, seed=0) if key == "c2" else W.long_3d(200_000, 200_000, seed=0)
    assert [W.sha256(P), W.sha256(Q)] == g["sha256"]
    assert g["endpoint_lb"] == O.endpoint_lb(P, Q)
    assert g["endpoint_lb"] <= g["vefd"] <= g["dfd"]
    assert g["distance"] == (g["lb"] + g["ub"]) / 2.0
    assert (g["iterations"] == 0) == (g["dfd"] - g["vefd"] <= REL * g["dfd"])
    (r_lo, ub_lo), (r_hi, ub_hi) = g["decisions"]
    assert r_lo < g["dfd"] < r_hi
    # Above DFD the end is reachable (rightUB 1); the distance is DFD here,
    # so just below it the end is not (rightUB -2).
    assert ub_hi == 1.0
    assert ub_lo == -2.0 and g["distance"] == g["dfd"]
