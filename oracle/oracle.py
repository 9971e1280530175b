"""ctypes view of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker the GPU library is compared against.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_dp = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.c_int64


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        for name in ("or_dfd",):
            getattr(L, name).argtypes = [_dp, _i64, _dp, _i64, ctypes.c_int]
            getattr(L, name).restype = ctypes.c_double
        L.or_endpoint_lb.argtypes = [_dp, _i64, _dp, _i64, ctypes.c_int]
        L.or_endpoint_lb.restype = ctypes.c_double
        for name in ("or_vefd", "or_vefd_f32", "or_decide_ub", "or_decide_ub_f32", "or_decide_ub_2ph"):
            getattr(L, name).argtypes = [_dp, _i64, _dp, _i64, ctypes.c_int, ctypes.c_double]
            getattr(L, name).restype = ctypes.c_double
        L.or_decide.argtypes = [_dp, _i64, _dp, _i64, ctypes.c_int, ctypes.c_double]
        L.or_decide.restype = ctypes.c_int
        L.or_free_interval.argtypes = [_dp, _dp, ctypes.c_int, _i64, _i64, ctypes.c_double,
                                       ctypes.c_int, _dp]
        L.or_free_interval.restype = None
        L.or_free_interval_2ph.argtypes = L.or_free_interval.argtypes
        L.or_free_interval_2ph.restype = None
        L.or_ve_weights.argtypes = [_dp, _dp, ctypes.c_int, _i64, _i64, _dp]
        L.or_ve_weights.restype = None
        L.or_distance.argtypes = [_dp, _i64, _dp, _i64, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_double, ctypes.c_int, _dp, _dp,
                                  ctypes.POINTER(ctypes.c_int32), _i64]
        L.or_distance.restype = _i64
        L.or_distance_form.argtypes = [_dp, _i64, _dp, _i64, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                       ctypes.POINTER(ctypes.c_int32), _i64]
        L.or_distance_form.restype = _i64
        L.or_sweep_rows.argtypes = [_dp, _i64, _dp, _i64, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_double, _i64, _i64]
        L.or_sweep_rows.restype = ctypes.c_double
        L.or_sweep_rows_mt.argtypes = [_dp, _i64, _dp, _i64, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_double, _i64, _i64, ctypes.c_int]
        L.or_sweep_rows_mt.restype = ctypes.c_double
        L.or_sweep_mt.argtypes = [_dp, _i64, _dp, _i64, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_int]
        L.or_sweep_mt.restype = ctypes.c_double
        L.or_refine.argtypes = [_dp, _i64, ctypes.c_int, ctypes.c_int, _dp]
        L.or_refine.restype = _i64
        L.or_dfd_refined.argtypes = [_dp, _i64, _dp, _i64, ctypes.c_int, ctypes.c_int]
        L.or_dfd_refined.restype = ctypes.c_double
        _LIB = L
    return _LIB


def ref_lib():
    """The reference's own free-interval routine built into _ref/, or None."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libref_intervals.so")
        if not os.path.exists(path):
            return None
        R = ctypes.CDLL(path)
        R.ref_intersection_interval.argtypes = [_dp, _dp, _dp, ctypes.c_int, ctypes.c_double, _dp]
        R.ref_intersection_interval.restype = None
        R.ref_dist_sqr.argtypes = [_dp, _dp, ctypes.c_int]
        R.ref_dist_sqr.restype = ctypes.c_double
        _REF = R
    return _REF


def _c(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def endpoint_lb(P, Q) -> float:
    P, Q = _c(P), _c(Q)
    return lib().or_endpoint_lb(_p(P), len(P), _p(Q), len(Q), P.shape[1])


def dfd(P, Q) -> float:
    P, Q = _c(P), _c(Q)
    return lib().or_dfd(_p(P), len(P), _p(Q), len(Q), P.shape[1])


def vefd(P, Q, init_lb: float | None = None, fp32: bool = False) -> float:
    P, Q = _c(P), _c(Q)
    if init_lb is None:
        init_lb = endpoint_lb(P, Q)
    f = lib().or_vefd_f32 if fp32 else lib().or_vefd
    return f(_p(P), len(P), _p(Q), len(Q), P.shape[1], init_lb)


def decide_ub(P, Q, r: float, fp32: bool = False, form: int = 0) -> float:
    """Terminal rightUB of the decision at r.  form 0: reachable<D>'s
    intervals; form 1: the two-phase path's (interval_calculation<D>)."""
    P, Q = _c(P), _c(Q)
    if form == 1:
        if fp32:
            raise ValueError("the two-phase form is FP64")
        f = lib().or_decide_ub_2ph
    else:
        f = lib().or_decide_ub_f32 if fp32 else lib().or_decide_ub
    return f(_p(P), len(P), _p(Q), len(Q), P.shape[1], r)


def decide(P, Q, r: float, fp32: bool = False, form: int = 0) -> bool:
    return decide_ub(P, Q, r, fp32, form) == 1.0


def free_interval(P, Q, i: int, j: int, r: float, which: int, form: int = 0):
    P, Q = _c(P), _c(Q)
    out = np.zeros(2)
    f = lib().or_free_interval_2ph if form == 1 else lib().or_free_interval
    f(_p(P), _p(Q), P.shape[1], i, j, r, which, _p(out))
    return float(out[0]), float(out[1])


def ve_weights(P, Q, i: int, j: int):
    P, Q = _c(P), _c(Q)
    out = np.zeros(2)
    lib().or_ve_weights(_p(P), _p(Q), P.shape[1], i, j, _p(out))
    return float(out[0]), float(out[1])


def distance(P, Q, stop_mode: int = 1, eps: float = 1e-9, vefd_gate: bool = True,
             max_trace: int = 256, form: int = 0):
    """Returns dict(distance, endpoint_lb, dfd, vefd, lb, ub, iterations, splits, decisions).
    form selects the decisions' interval form (0 fused, 1 two-phase)."""
    P, Q = _c(P), _c(Q)
    out = np.zeros(6)
    splits = np.zeros(max_trace)
    decs = np.zeros(max_trace, dtype=np.int32)
    it = lib().or_distance_form(_p(P), len(P), _p(Q), len(Q), P.shape[1], stop_mode, eps,
                                1 if vefd_gate else 0, form, _p(out), _p(splits),
                                decs.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), max_trace)
    k = min(it, max_trace)
    return dict(distance=out[0], endpoint_lb=out[1], dfd=out[2], vefd=out[3], lb=out[4],
                ub=out[5], iterations=int(it), splits=splits[:k].copy(),
                decisions=decs[:k].astype(bool).copy())


def sweep_rows(P, Q, kind: int, r: float, r0: int, r1: int) -> float:
    P, Q = _c(P), _c(Q)
    return lib().or_sweep_rows(_p(P), len(P), _p(Q), len(Q), P.shape[1], kind, r, r0, r1)


def sweep_rows_mt(P, Q, kind: int, r: float, r0: int, r1: int, threads: int) -> float:
    """sweep_rows on `threads` CPU threads (row blocks pipelined over column
    chunks); kinds 0 (VE-FD) and 2 (DFD), same result as sweep_rows."""
    P, Q = _c(P), _c(Q)
    return lib().or_sweep_rows_mt(_p(P), len(P), _p(Q), len(Q), P.shape[1], kind, r, r0, r1,
                                  threads)


def sweep_mt(P, Q, kind: int, r: float, threads: int) -> float:
    """Terminal value of a whole sweep on `threads` CPU threads: kind 0 VE-FD
    (r = endpoint bound; same as vefd), 1 decision at r (terminal rightUB;
    same as decide_ub), 2 DFD (same as dfd)."""
    P, Q = _c(P), _c(Q)
    return lib().or_sweep_mt(_p(P), len(P), _p(Q), len(Q), P.shape[1], kind, float(r), int(threads))


def refine(P, levels: int) -> np.ndarray:
    """The curve with every segment split at its midpoint, `levels` times."""
    P = _c(P)
    n = lib().or_refine(_p(P), len(P), P.shape[1], levels, None)
    out = np.zeros((n, P.shape[1]))
    lib().or_refine(_p(P), len(P), P.shape[1], levels, _p(out))
    return out


def dfd_refined(P, Q, levels: int) -> float:
    """DFD of the midpoint-refined curves: an upper bound between FD and DFD."""
    P, Q = _c(P), _c(Q)
    return lib().or_dfd_refined(_p(P), len(P), _p(Q), len(Q), P.shape[1], levels)
