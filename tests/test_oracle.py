"""Pins the CPU oracle (the checker) against independent definitions.

The reference ships no tests, fixtures or golden vectors for this path, so
each oracle function is checked against a textbook restatement, brute force,
closed forms, or the reference's own header-only free-interval routine.
"""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests import independent as I

RNG = np.random.default_rng(12345)


def rand_pair(m, n, d=2, scale=1.0, rng=RNG):
    return rng.standard_normal((m, d)) * scale, rng.standard_normal((n, d)) * scale


def walk(rng, k, d):
    return np.cumsum(rng.standard_normal((k, d)), axis=0)


# ---------------- DFD ----------------

@pytest.mark.parametrize("seed", range(12))
def test_dfd_matches_eiter_mannila(seed):
    rng = np.random.default_rng(seed)
    m, n = rng.integers(1, 30, 2)
    d = 2 + seed % 2
    P, Q = walk(rng, m, d), walk(rng, n, d)
    assert O.dfd(P, Q) == pytest.approx(I.dfd_eiter_mannila(P, Q), rel=4e-16, abs=0)


@pytest.mark.parametrize("seed", range(6))
def test_dfd_matches_coupling_brute_force(seed):
    rng = np.random.default_rng(100 + seed)
    m, n = rng.integers(1, 6, 2)
    P, Q = walk(rng, m, 2), walk(rng, n, 2)
    assert O.dfd(P, Q) == pytest.approx(I.dfd_brute(P, Q), rel=4e-16, abs=0)


# ---------------- VE-FD ----------------

def test_ve_weights_match_projection_formula():
    rng = np.random.default_rng(7)
    for _ in range(500):
        d = int(rng.integers(2, 4))
        P, Q = rng.standard_normal((4, d)), rng.standard_normal((4, d))
        i, j = rng.integers(0, 3, 2)
        wt, wr = O.ve_weights(P, Q, i, j)
        assert math.sqrt(wt) == pytest.approx(I.point_seg_dist(Q[j + 1], P[i], P[i + 1]), rel=1e-12)
        assert math.sqrt(wr) == pytest.approx(I.point_seg_dist(P[i + 1], Q[j], Q[j + 1]), rel=1e-12)


def test_ve_weight_top_and_right_are_not_swapped():
    # cell (0,0): top = |q1 - seg(p0,p1)|^2, right = |p1 - seg(q0,q1)|^2
    P = np.array([[0.0, 0.0], [2.0, 0.0], [3.0, 1.0]])
    Q = np.array([[0.0, 1.0], [1.0, 1.0], [3.0, 3.0]])
    assert O.ve_weights(P, Q, 0, 0) == (1.0, 2.0)


@pytest.mark.parametrize("seed", range(10))
def test_vefd_matches_brute_force_paths(seed):
    rng = np.random.default_rng(200 + seed)
    m, n = rng.integers(2, 7, 2)
    d = 2 + seed % 2
    P, Q = walk(rng, m, d), walk(rng, n, d)
    init = O.endpoint_lb(P, Q)
    brute = I.vefd_brute(m, n, lambda i, j: O.ve_weights(P, Q, i, j), init)
    assert O.vefd(P, Q, init) == math.sqrt(brute)  # selection of one weight: bit-exact


@pytest.mark.parametrize("seed", range(6))
def test_vefd_matches_readme_recurrence(seed):
    rng = np.random.default_rng(300 + seed)
    m, n = rng.integers(2, 25, 2)
    P, Q = walk(rng, m, 3), walk(rng, n, 3)
    init = O.endpoint_lb(P, Q)
    rec = I.vefd_readme(m, n, lambda i, j: O.ve_weights(P, Q, i, j), init)
    assert O.vefd(P, Q, init) == math.sqrt(rec)


# ---------------- endpoint bound ----------------

def test_endpoint_lb_matches_reference_dist_sqr():
    R = O.ref_lib()
    if R is None:
        pytest.skip("reference free-interval library not built (no /root/reference)")
    import ctypes
    dp = ctypes.POINTER(ctypes.c_double)
    rng = np.random.default_rng(5)
    for _ in range(200):
        d = int(rng.integers(2, 4))
        P, Q = rng.standard_normal((3, d)) * 50, rng.standard_normal((4, d)) * 50
        a = R.ref_dist_sqr(np.ascontiguousarray(P[0]).ctypes.data_as(dp),
                           np.ascontiguousarray(Q[0]).ctypes.data_as(dp), d)
        b = R.ref_dist_sqr(np.ascontiguousarray(P[-1]).ctypes.data_as(dp),
                           np.ascontiguousarray(Q[-1]).ctypes.data_as(dp), d)
        assert O.endpoint_lb(P, Q) == math.sqrt(max(a, b))


# ---------------- free intervals ----------------

def test_free_interval_matches_quadratic_roots():
    rng = np.random.default_rng(11)
    for _ in range(2000):
        d = int(rng.integers(2, 4))
        P, Q = rng.standard_normal((3, d)), rng.standard_normal((3, d))
        r = float(rng.uniform(0.05, 2.5))
        which = int(rng.integers(0, 2))
        lb, ub = O.free_interval(P, Q, 0, 0, r, which)
        c, s0, s1 = (P[1], Q[0], Q[1]) if which == 0 else (Q[1], P[0], P[1])
        ref = I.free_interval(c, s0, s1, r)
        if ref is None:
            # empty, or empty up to rounding at a tangency
            assert (lb, ub) == (-1.0, -2.0) or ub - lb < 1e-6
        else:
            assert lb == pytest.approx(ref[0], abs=1e-9)
            assert ub == pytest.approx(ref[1], abs=1e-9)


def test_free_interval_within_reference_intersection_interval():
    """IntersectionAlgorithm::intersection_interval (geometry_basics.hpp:161-298)
    returns an inner interval of the free interval within eps = 2.5e-4."""
    R = O.ref_lib()
    if R is None:
        pytest.skip("reference free-interval library not built (no /root/reference)")
    import ctypes
    dp = ctypes.POINTER(ctypes.c_double)
    rng = np.random.default_rng(13)
    eps = 0.001 / 4
    checked = 0
    for _ in range(3000):
        d = int(rng.integers(2, 4))
        P, Q = rng.standard_normal((3, d)), rng.standard_normal((3, d))
        r = float(rng.uniform(0.05, 2.5))
        lb, ub = O.free_interval(P, Q, 0, 0, r, 0)
        c, s0, s1 = (np.ascontiguousarray(x) for x in (P[1], Q[0], Q[1]))
        out = np.zeros(2)
        R.ref_intersection_interval(c.ctypes.data_as(dp), s0.ctypes.data_as(dp),
                                    s1.ctypes.data_as(dp), d, r, out.ctypes.data_as(dp))
        ref_empty = out[0] > out[1]
        ours_empty = ub < lb
        if ref_empty:
            assert ours_empty or ub - lb <= 2 * eps
            continue
        assert not ours_empty
        assert lb <= out[0] + 1e-12 and out[1] <= ub + 1e-12
        assert out[0] - lb <= eps + 1e-12 and ub - out[1] <= eps + 1e-12
        checked += 1
    assert checked > 1000


# ---------------- decision ----------------

@pytest.mark.parametrize("seed", range(8))
def test_decision_matches_alt_godau(seed):
    rng = np.random.default_rng(400 + seed)
    for _ in range(25):
        m, n = rng.integers(2, 12, 2)
        d = int(rng.integers(2, 4))
        P, Q = walk(rng, m, d), walk(rng, n, d)
        lb, ub = O.endpoint_lb(P, Q), O.dfd(P, Q)
        r = float(rng.uniform(lb, ub * 1.05 + 1e-9))
        assert O.decide(P, Q, r) == I.decide_altgodau(P, Q, r)


def test_decision_is_monotone_in_radius():
    rng = np.random.default_rng(17)
    P, Q = walk(rng, 40, 2), walk(rng, 50, 2)
    rs = np.linspace(O.endpoint_lb(P, Q), O.dfd(P, Q) * 1.01, 60)
    decs = [O.decide(P, Q, r) for r in rs]
    first = decs.index(True)
    assert all(decs[first:]) and not any(decs[:first])


@pytest.mark.parametrize("seed", range(6))
def test_distance_equals_a_critical_value(seed):
    rng = np.random.default_rng(500 + seed)
    m, n = rng.integers(3, 7, 2)
    P, Q = walk(rng, m, 2), walk(rng, n, 2)
    res = O.distance(P, Q, stop_mode=1, eps=1e-12)
    # The decision is defined for r >= the endpoint bound only: the reference
    # never tests the start corner (row 0 enters with botLB = 0, .cu:1026).
    lb = O.endpoint_lb(P, Q)
    cand = [c for c in I.critical_values(P, Q) if c >= lb]
    ok = [c for c in cand if O.decide(P, Q, c * (1 + 1e-11))]
    fd = min(ok)
    assert res["distance"] == pytest.approx(fd, rel=1e-9)
    if fd * (1 - 1e-9) >= lb:
        assert not O.decide(P, Q, fd * (1 - 1e-9))


# ---------------- bounds and closed forms ----------------

@pytest.mark.parametrize("seed", range(10))
def test_bound_order(seed):
    rng = np.random.default_rng(600 + seed)
    m, n = rng.integers(2, 60, 2)
    P, Q = walk(rng, m, 2 + seed % 2), walk(rng, n, 2 + seed % 2)
    res = O.distance(P, Q, stop_mode=1, eps=1e-10, vefd_gate=False)
    assert res["endpoint_lb"] <= res["vefd"] * (1 + 1e-15)
    assert res["vefd"] <= res["distance"] * (1 + 1e-9)
    assert res["distance"] <= res["dfd"] * (1 + 1e-15)


def test_translated_copy_distance_is_offset():
    rng = np.random.default_rng(21)
    P = walk(rng, 30, 3)
    v = np.array([0.3, -0.4, 1.2])
    res = O.distance(P, P + v, stop_mode=1, eps=1e-12, vefd_gate=False)
    assert res["distance"] == pytest.approx(np.linalg.norm(v), rel=1e-9)


def test_single_segments_distance_is_endpoint_max():
    P = np.array([[0.0, 0.0], [4.0, 1.0]])
    Q = np.array([[1.0, 2.0], [3.0, -1.0]])
    res = O.distance(P, Q, stop_mode=1, eps=1e-12, vefd_gate=False)
    expect = max(np.linalg.norm(P[0] - Q[0]), np.linalg.norm(P[1] - Q[1]))
    assert res["distance"] == pytest.approx(expect, rel=1e-9)


def test_identical_curves_distance_zero():
    rng = np.random.default_rng(22)
    P = walk(rng, 50, 2)
    res = O.distance(P, P.copy(), stop_mode=1, eps=1e-9, vefd_gate=False)
    assert res["distance"] == 0.0 and res["iterations"] == 0


def test_2d_embedded_in_3d_is_bit_identical():
    rng = np.random.default_rng(23)
    P, Q = walk(rng, 70, 2), walk(rng, 90, 2)
    P3 = np.hstack([P, np.zeros((70, 1))])
    Q3 = np.hstack([Q, np.zeros((90, 1))])
    assert O.dfd(P, Q) == O.dfd(P3, Q3)
    assert O.vefd(P, Q) == O.vefd(P3, Q3)
    r = 0.5 * (O.vefd(P, Q) + O.dfd(P, Q))
    assert O.decide_ub(P, Q, r) == O.decide_ub(P3, Q3, r)


@pytest.mark.parametrize("seed", range(4))
def test_symmetries(seed):
    rng = np.random.default_rng(700 + seed)
    P, Q = walk(rng, 25 + seed, 2), walk(rng, 31, 2)
    a = O.distance(P, Q, stop_mode=1, eps=1e-11, vefd_gate=False)["distance"]
    b = O.distance(Q, P, stop_mode=1, eps=1e-11, vefd_gate=False)["distance"]
    c = O.distance(P[::-1], Q[::-1], stop_mode=1, eps=1e-11, vefd_gate=False)["distance"]
    assert a == pytest.approx(b, rel=1e-9)
    assert a == pytest.approx(c, rel=1e-9)


def test_collinear_midpoints_leave_distance_unchanged():
    rng = np.random.default_rng(24)
    P, Q = walk(rng, 12, 2), walk(rng, 15, 2)
    mids = (P[:-1] + P[1:]) / 2
    P2 = np.empty((2 * len(P) - 1, 2))
    P2[0::2], P2[1::2] = P, mids
    a = O.distance(P, Q, stop_mode=1, eps=1e-11, vefd_gate=False)["distance"]
    b = O.distance(P2, Q, stop_mode=1, eps=1e-11, vefd_gate=False)["distance"]
    assert a == pytest.approx(b, rel=1e-8)


# ---------------- bisection driver ----------------

def test_bisection_trajectory_restates_frechet_cpp():
    rng = np.random.default_rng(31)
    P, Q = walk(rng, 60, 2), walk(rng, 80, 2)
    res = O.distance(P, Q, stop_mode=0, eps=float(np.float32(np.finfo(np.float32).eps)))
    lb, ub = res["vefd"], res["dfd"]
    for split, dec in zip(res["splits"], res["decisions"]):
        assert split == (ub + lb) / 2.0
        assert dec == O.decide(P, Q, split)
        if dec:
            ub = split
        else:
            lb = split
    assert not (ub - lb > np.finfo(np.float32).eps)
    assert res["distance"] == (ub + lb) / 2.0


def test_gate_skips_vefd_on_small_grids():
    rng = np.random.default_rng(32)
    P, Q = walk(rng, 50, 2), walk(rng, 60, 2)  # 3000 <= 10000 (.cu:1343)
    res = O.distance(P, Q)
    assert res["vefd"] == res["endpoint_lb"]


def test_degenerate_single_point_is_dfd():
    P = np.array([[0.0, 0.0]])
    Q = np.array([[1.0, 0.0], [0.0, 3.0], [2.0, 2.0]])
    res = O.distance(P, Q)
    assert res["distance"] == pytest.approx(3.0)


# ---------------- FP32 forms ----------------

def test_fp32_forms_track_fp64():
    rng = np.random.default_rng(41)
    P, Q = walk(rng, 300, 3), walk(rng, 280, 3)
    v64, v32 = O.vefd(P, Q), O.vefd(P, Q, fp32=True)
    assert v32 == pytest.approx(v64, rel=1e-5)
    lb, ub = O.endpoint_lb(P, Q), O.dfd(P, Q)
    for r in np.linspace(lb, ub * 1.02, 15):
        if abs(r - ub) / ub < 1e-4:
            continue
        assert O.decide(P, Q, r, fp32=True) == O.decide(P, Q, r)


# ---------------- timing sweeps (CPU baseline) ----------------

@pytest.mark.parametrize("seed", range(3))
def test_sweep_rows_pinned(seed):
    """The timing sweeps compute the oracle's recurrences: the DFD kind returns
    DFD of the sampled rows against Q, the VE-FD kind the sum of the last
    sampled row's right-edge values of the README recurrence."""
    rng = np.random.default_rng(200 + seed)
    m, n = 23 + seed, 31 - seed
    P, Q = walk(rng, m, 3), walk(rng, n, 3)
    assert O.sweep_rows(P, Q, 2, 0.0, 0, m) == pytest.approx(I.dfd_eiter_mannila(P, Q), rel=4e-16)
    assert O.sweep_rows(P, Q, 2, 0.0, 5, 17) == pytest.approx(
        I.dfd_eiter_mannila(P[5:17], Q), rel=4e-16)
    init = O.endpoint_lb(P, Q) * 1.3
    _, Rr = I.vefd_readme_tables(m, n, lambda i, j: O.ve_weights(P, Q, i, j), init)
    for rows in (1, 7, m - 1):
        assert O.sweep_rows(P, Q, 0, init, 0, rows) == pytest.approx(Rr[rows - 1].sum(), rel=1e-15)


@pytest.mark.parametrize("kind", [0, 2])
@pytest.mark.parametrize("threads", [1, 2, 3, 8, 40])
def test_sweep_rows_mt_matches_single_thread(kind, threads):
    """The multi-core timing sweep is bit-identical to the single-thread one:
    ragged row blocks, more threads than rows, column counts that are not a
    multiple of the 256-column chunk, samples that do not start at row 0."""
    rng = np.random.default_rng(300 + threads)
    P, Q = walk(rng, 61, 3), walk(rng, 700, 3)
    r = O.endpoint_lb(P, Q) * 2.0
    rmax = len(P) - 1 if kind == 0 else len(P)
    for r0, r1 in ((0, rmax), (0, 1), (0, 17), (9, 40), (rmax - 3, rmax)):
        st = O.sweep_rows(P, Q, kind, r, r0, r1)
        mt = O.sweep_rows_mt(P, Q, kind, r, r0, r1, threads)
        assert mt == st, (r0, r1)


@pytest.mark.parametrize("threads", [1, 2, 3, 8, 40])
def test_sweep_mt_terminal_values_equal_single_thread(threads):
    """The row-pipelined whole sweeps that produce the reference values of
    long pairs are bit-identical to or_vefd, or_dfd and or_decide_ub (whose
    decision computes every cell's intervals; the pipelined one skips cells
    whose inputs are both unreachable): ragged row blocks, more threads than
    rows, columns not a multiple of the 256-column chunk, 2D and 3D, radii from
    the endpoint bound to past DFD and within 1e-9 of the distance."""
    rng = np.random.default_rng(900 + threads)
    shapes = [(2, 2, 2), (3, 2, 3), (2, 300, 2), (61, 700, 3), (300, 257, 2), (40, 513, 3)]
    for m, n, d in shapes:
        P, Q = walk(rng, m, d), walk(rng, n, d)
        lb, dfd = O.endpoint_lb(P, Q), O.dfd(P, Q)
        same = lambda a, b: a == b and math.copysign(1.0, a) == math.copysign(1.0, b)
        assert same(O.sweep_mt(P, Q, 2, 0.0, threads), dfd), (m, n)
        assert same(O.sweep_mt(P, Q, 0, lb, threads), O.vefd(P, Q, lb)), (m, n)
        fd = O.distance(P, Q, stop_mode=1, eps=1e-12)["distance"]
        radii = list(np.linspace(lb, dfd * 1.05 + 1e-12, 6)) + [fd * (1 - 1e-9), fd * (1 + 1e-9)]
        for r in radii:
            assert same(O.sweep_mt(P, Q, 1, r, threads), O.decide_ub(P, Q, r)), (m, n, r)


# ---------------- two-phase decision form ----------------

@pytest.mark.parametrize("seed", range(6))
def test_two_phase_decision_away_from_critical_values(seed):
    """interval_calculation + reachability (.cu:674-958) decides like the
    fused reachable<D> and like textbook Alt-Godau at every radius that is not
    within 1e-9 (relative) of a critical value, where rounding may differ."""
    rng = np.random.default_rng(500 + seed)
    checked = 0
    for _ in range(25):
        m, n = rng.integers(2, 8, 2)
        d = 2 + seed % 2
        P, Q = walk(rng, m, d), walk(rng, n, d)
        crit = I.critical_values(P, Q)
        # From the endpoint bound up: the sweep takes cell (0,0) as reachable
        # (.cu:1026) and the bisection never asks below that bound.
        for r in np.linspace(O.endpoint_lb(P, Q), O.dfd(P, Q) * 1.1, 13):
            if np.min(np.abs(crit - r)) <= 1e-9 * r:
                continue
            two = O.decide(P, Q, r, form=1)
            assert two == O.decide(P, Q, r)
            assert two == I.decide_altgodau(P, Q, r)
            checked += 1
    assert checked > 150


def test_two_phase_intervals_differ_from_fused_in_rounding_only():
    """The C term is (sum mi^2) - r^2 (.cu:733-742) instead of a sum started at
    -r^2 (.cu:1176-1183): some interval ends round differently, by a few ulps,
    never by more."""
    rng = np.random.default_rng(77)
    differ = 0
    for _ in range(60):
        P, Q = walk(rng, 6, 3), walk(rng, 6, 3)
        for i in range(5):
            for j in range(5):
                for which in (0, 1):
                    r = rng.uniform(0.3, 3.0)
                    a = O.free_interval(P, Q, i, j, r, which)
                    b = O.free_interval(P, Q, i, j, r, which, form=1)
                    if a != b:
                        differ += 1
                        assert (a[0] < 0) == (b[0] < 0)
                        assert a == pytest.approx(b, rel=1e-12, abs=1e-12)
    assert differ > 0


def test_two_phase_zero_length_segment_keeps_nan():
    """A zero-length segment makes 0.5/|ab|^2 infinite and t = 0 * inf NaN.
    The two-phase clamps are comparisons (.cu:751-752, 778-779), so the
    interval stays NaN; reachable<D>'s max / min give [0, 1]."""
    P = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 0.0], [2.0, 0.0]])
    Q = np.array([[0.0, 1.0], [2.0, 1.0]])
    for r in (0.5, 1.5):
        lb, ub = O.free_interval(P, Q, 1, 0, r, 1, form=1)
        assert math.isnan(lb) and math.isnan(ub)
        assert O.free_interval(P, Q, 1, 0, r, 1) == (0.0, 1.0)


@pytest.mark.parametrize("seed", range(4))
def test_two_phase_distance_within_tolerance(seed):
    rng = np.random.default_rng(900 + seed)
    P, Q = walk(rng, 40, 2), walk(rng, 55, 2)
    a = O.distance(P, Q, stop_mode=1, eps=1e-9)
    b = O.distance(P, Q, stop_mode=1, eps=1e-9, form=1)
    assert b["dfd"] == a["dfd"] and b["vefd"] == a["vefd"]
    assert b["distance"] == pytest.approx(a["distance"], rel=2e-9)


# ---------------- refined upper bound (DFD of midpoint-refined curves) ----------------

def numpy_refine(P, levels):
    for _ in range(levels):
        mid = P[:-1] + 0.5 * (P[1:] - P[:-1])
        out = np.empty((2 * len(P) - 1, P.shape[1]))
        out[0::2] = P
        out[1::2] = mid
        P = out
    return P


@pytest.mark.parametrize("levels", [0, 1, 2, 3])
def test_refine_is_the_midpoint_rule(levels):
    rng = np.random.default_rng(70 + levels)
    for d in (2, 3):
        P = walk(rng, 9, d)
        got = O.refine(P, levels)
        assert got.shape == ((9 - 1) * 2 ** levels + 1, d)
        assert np.array_equal(got, numpy_refine(P, levels))
    one = walk(rng, 1, 2)
    assert np.array_equal(O.refine(one, levels), one)  # a single point has no segment to split


@pytest.mark.parametrize("seed", range(8))
def test_refined_dfd_is_eiter_mannila_of_refined_curves(seed):
    rng = np.random.default_rng(80 + seed)
    m, n = rng.integers(2, 9, 2)
    P, Q = walk(rng, m, 2 + seed % 2), walk(rng, n, 2 + seed % 2)
    for levels in (1, 2):
        want = I.dfd_eiter_mannila(numpy_refine(P, levels), numpy_refine(Q, levels))
        assert O.dfd_refined(P, Q, levels) == pytest.approx(want, rel=1e-15)
    assert O.dfd_refined(P, Q, 0) == O.dfd(P, Q)


@pytest.mark.parametrize("seed", range(8))
def test_refined_dfd_lies_between_fd_and_dfd(seed):
    """FD <= DFD(refined) <= DFD, with FD the least Alt-Godau critical value
    (endpoint, vertex-edge, monotonicity) that the independent decision
    accepts; each level of refinement is at most the one before."""
    rng = np.random.default_rng(90 + seed)
    m, n = rng.integers(3, 7, 2)
    P, Q = walk(rng, m, 2), walk(rng, n, 2)
    lb = O.endpoint_lb(P, Q)
    cand = [c for c in I.critical_values(P, Q) if c >= lb]
    fd = min(c for c in cand if I.decide_altgodau(P, Q, c * (1 + 1e-11)))
    prev = O.dfd(P, Q)
    for levels in (1, 2, 3, 4):
        ub = O.dfd_refined(P, Q, levels)
        assert fd * (1 - 1e-12) <= ub <= prev * (1 + 1e-15)
        prev = ub
    assert prev - fd <= O.dfd(P, Q) - fd
