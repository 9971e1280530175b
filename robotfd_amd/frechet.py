"""Python mirror of RobotFD's continuous Frechet distance interface.

Names and argument meanings follow the reference:

* ``Point``, ``Curve``                    include/geometry_basics.hpp:11-99, include/curve.hpp:9-40
* ``compute_distance_parallel``           include/wrapper.hpp:4, src/compute_distance_wrapper.cpp:8-12
* ``distance_cuda``                       include/frechet.hpp:14, src/frechet.cpp:18-75
* ``Intersection`` (``Cuda_intersection``) include/intersection_algorithm_in_parallel.hpp:11-61

Every call goes through the C ABI of librfd.so (include/rfd.h) to the gfx950
kernels.  ``forceOversize`` selects the decision path as in the reference:
without it the reference picks its two-phase path when the interval arrays fit
in free device memory (.cu:87-96), with it always the fused one
(.cu:960-1256); both are reproduced bit for bit.
"""
from __future__ import annotations

import ctypes
import math
import sys
import threading

import numpy as np

from . import _lib
from ._lib import check, dptr, lib

FLT_EPSILON = float(np.finfo(np.float32).eps)  # compute_distance_wrapper.cpp:6


class Point:
    """A point with `coordinates` (geometry_basics.hpp:11-99)."""

    __slots__ = ("coordinates",)

    def __init__(self, coordinates=()):
        self.coordinates = [float(c) for c in coordinates]

    def dist_sqr(self, other: "Point") -> float:
        # geometry_basics.hpp:72-81: each product rounded before the sum.
        result = 0.0
        for a, b in zip(self.coordinates, other.coordinates):
            d = a - b
            result += d * d
        return result

    def dist(self, other: "Point") -> float:
        return math.sqrt(self.dist_sqr(other))

    def __getitem__(self, i):
        return self.coordinates[i]

    def __len__(self):
        return len(self.coordinates)


class Curve:
    """A trajectory of points of a fixed dimension (curve.hpp:9-40)."""

    def __init__(self, points, dimensions: int | None = None):
        arr = np.asarray([p.coordinates if isinstance(p, Point) else p for p in points],
                         dtype=np.float64) if not isinstance(points, np.ndarray) else points
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        if arr.size == 0:
            print("warning: constructed empty curve", file=sys.stderr)  # curve.cpp:5-8
            arr = np.zeros((0, dimensions or 2))
        self._pts = arr
        self._dims = int(dimensions if dimensions is not None else arr.shape[1])

    def size(self) -> int:
        return int(self._pts.shape[0])

    def empty(self) -> bool:
        return self.size() == 0

    def dimensions(self) -> int:
        return self._dims

    def __len__(self):
        return self.size()

    def __getitem__(self, i) -> Point:
        return Point(self._pts[i])

    def front(self) -> Point:
        return Point(self._pts[0])

    def back(self) -> Point:
        return Point(self._pts[-1])

    def push_back(self, point) -> None:
        p = np.asarray(point.coordinates if isinstance(point, Point) else point, dtype=np.float64)
        self._pts = np.ascontiguousarray(np.vstack([self._pts, p[None, :]]))

    def array(self) -> np.ndarray:
        """The AoS FP64 array (points, dims) handed to the library."""
        return self._pts


class Context:
    """One device + stream (rfd_ctx).  `Context.default()` is a per-thread context."""

    _tls = threading.local()

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        check(lib().rfd_create(ctypes.byref(h), device))
        self.handle = h

    @classmethod
    def default(cls) -> "Context":
        ctx = getattr(cls._tls, "ctx", None)
        if ctx is None:
            ctx = cls()
            cls._tls.ctx = ctx
        return ctx

    def stats(self) -> dict:
        st = _lib.Stats()
        check(lib().rfd_get_stats(self.handle, ctypes.byref(st)), self.handle)
        return st.as_dict()

    def reset_stats(self) -> None:
        check(lib().rfd_reset_stats(self.handle), self.handle)

    def close(self) -> None:
        if self.handle:
            lib().rfd_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _opts(stop: str = "abs", eps: float = FLT_EPSILON, precision: str = "fp64",
          vefd_gate: bool = True, verbose: bool = False, mixed_switch: float = 1e-4,
          decide_form: str = "fused", refine_ub: int = 0) -> _lib.Opts:
    """decide_form: "fused" (reachable<D>'s intervals), "two-phase"
    (interval_calculation<D>'s) or "reference" (the reference's choice from
    free device memory, per pair)."""
    o = _lib.Opts()
    lib().rfd_default_opts(ctypes.byref(o))
    o.decide_form = _lib.DECIDE_FORMS[decide_form]
    o.stop_mode = {"abs": _lib.STOP_ABS, "rel": _lib.STOP_REL}[stop]
    o.precision = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED}[precision]
    o.eps = float(eps)
    o.vefd_gate = 1 if vefd_gate else 0
    o.verbose = 1 if verbose else 0
    o.mixed_switch = float(mixed_switch)
    o.refine_ub = int(refine_ub)
    return o


class Intersection:
    """Device-resident pair of curves (the reference's Cuda_intersection).

    Constructor + ``data_move`` upload both curves (.cu:23-160);
    ``Bound_call_gpu`` returns (dfd, vefd) (.cu:1258-1395);
    ``intersection_interval_call_gpu(r)`` is the decision at radius r
    (.cu:1398-1597); ``free_memory`` releases the device buffers (.cu:162-175).
    """

    def __init__(self, curve1, curve2, forceOversize: bool = False, ctx: Context | None = None,
                 decide_form: str = "fused"):
        """decide_form fixes the session's decision path as the reference's
        constructor does (.cu:87-96): "fused", "two-phase", or "reference"
        (two-phase when the reference's interval arrays would fit in free
        device memory and forceOversize is not set)."""
        self.ctx = ctx or Context.default()
        self.decide_form = "fused" if (forceOversize and decide_form == "reference") else decide_form
        if self.decide_form not in _lib.DECIDE_FORMS:
            raise ValueError(f"decide_form must be one of {sorted(_lib.DECIDE_FORMS)}")
        self.effective_form = None
        self.P = _lib.as_curve(curve1.array() if isinstance(curve1, Curve) else curve1)
        self.Q = _lib.as_curve(curve2.array() if isinstance(curve2, Curve) else curve2)
        if self.P.shape[1] != self.Q.shape[1]:
            raise ValueError("curves must have the same dimension")
        self.handle = None

    def data_move(self) -> None:
        if self.handle is not None:
            return
        h = ctypes.c_void_p()
        check(lib().rfd_session_create(self.ctx.handle, dptr(self.P), len(self.P), dptr(self.Q),
                                       len(self.Q), self.P.shape[1], ctypes.byref(h)),
              self.ctx.handle)
        self.handle = h
        self.set_decide_form(self.decide_form)

    def set_decide_form(self, form: str) -> str:
        """Decision path of intersection_interval_call_gpu / decide_ub; returns
        the form in use ("fused" or "two-phase")."""
        self._need()
        eff = ctypes.c_int(0)
        check(lib().rfd_session_set_decide_form(self.handle, _lib.DECIDE_FORMS[form], ctypes.byref(eff)),
              self.ctx.handle)
        self.effective_form = "two-phase" if eff.value == _lib.DECIDE_TWO_PHASE else "fused"
        return self.effective_form

    def _need(self):
        if self.handle is None:
            self.data_move()

    def Bound_call_gpu(self, lb: float, run_vefd: bool = True, fp32: bool = False):
        """Returns (dfd, vefd); `lb` is the endpoint bound VE-FD starts from."""
        self._need()
        ub = ctypes.c_double(0.0)
        ve = ctypes.c_double(lb)
        check(lib().rfd_session_bounds(self.handle, ctypes.byref(ub), ctypes.byref(ve),
                                       1 if run_vefd else 0, 1 if fp32 else 0), self.ctx.handle)
        return ub.value, ve.value

    def refined_ub(self, levels: int = 1) -> float:
        """DFD of both curves refined `levels` times at segment midpoints: an
        upper bound of the distance at most DFD (rfd_session_refined_ub)."""
        self._need()
        ub = ctypes.c_double(0.0)
        check(lib().rfd_session_refined_ub(self.handle, int(levels), ctypes.byref(ub)), self.ctx.handle)
        return ub.value

    def intersection_interval_call_gpu(self, radius: float, fp32: bool = False) -> bool:
        return self.decide_ub(radius, fp32) == 1.0

    def decide_ub(self, radius: float, fp32: bool = False) -> float:
        """Terminal rightUB of the decision sweep (1.0 means reachable)."""
        self._need()
        r = ctypes.c_int(0)
        ub = ctypes.c_double(0.0)
        check(lib().rfd_session_decide(self.handle, float(radius), 1 if fp32 else 0,
                                       ctypes.byref(r), ctypes.byref(ub)), self.ctx.handle)
        return ub.value

    def distance(self, **kw) -> dict:
        self._need()
        res = _lib.Result()
        o = _opts(**kw)
        check(lib().rfd_session_distance(self.handle, ctypes.byref(o), ctypes.byref(res)),
              self.ctx.handle)
        return res.as_dict()

    def free_memory(self) -> None:
        if self.handle is not None:
            lib().rfd_session_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        self.data_move()
        return self

    def __exit__(self, *exc):
        self.free_memory()

    def __del__(self):
        try:
            self.free_memory()
        except Exception:
            pass


def distance(P, Q, *, stop: str = "rel", eps: float = 1e-9, precision: str = "fp64",
             vefd_gate: bool = True, verbose: bool = False, ctx: Context | None = None,
             decide_form: str = "fused") -> dict:
    """Full distance computation on two AoS arrays; returns the rfd_result fields."""
    ctx = ctx or Context.default()
    P = _lib.as_curve(P)
    Q = _lib.as_curve(Q)
    res = _lib.Result()
    o = _opts(stop=stop, eps=eps, precision=precision, vefd_gate=vefd_gate, verbose=verbose,
              decide_form=decide_form)
    check(lib().rfd_distance(ctx.handle, dptr(P), len(P), dptr(Q), len(Q), P.shape[1],
                             ctypes.byref(o), ctypes.byref(res)), ctx.handle)
    return res.as_dict()


def distance_cuda(curve1: Curve, curve2: Curve, eps: float, forceOversize: bool = False,
                  verbose: bool = True) -> float:
    """frechet.cpp:18-75: bracket [VE-FD, DFD], bisection while ub - lb > eps,
    with the reference's decision path (fused when forceOversize is set or its
    interval arrays would not fit, else two-phase; .cu:87-96)."""
    return distance(curve1.array(), curve2.array(), stop="abs", eps=eps, verbose=verbose,
                    decide_form="fused" if forceOversize else "reference")["distance"]


def compute_distance_parallel(curve1: Curve, curve2: Curve, forceOversize: bool = False) -> float:
    """compute_distance_wrapper.cpp:8-12: distance_cuda with eps = FLT_EPSILON.

    On failure it prints the error and returns NaN, as the C++ drop-in does."""
    try:
        return distance_cuda(curve1, curve2, FLT_EPSILON, forceOversize, verbose=False)
    except (_lib.RFDError, ValueError) as e:
        print(f"compute_distance_parallel: {e}", file=sys.stderr)
        return float("nan")
