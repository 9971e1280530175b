"""Batches of pairs on the device (rfd_batch_* in include/rfd.h).

Every round of every pair (DFD, VE-FD, the bisection's decisions) runs as a
persistent kernel over all active pairs; the rounds are captured once in a
hipGraph and replayed.  No counterpart in the reference, which computes one
pair per call.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib
from .frechet import Context, _opts


class Batch:
    def __init__(self, pairs, ctx: Context | None = None):
        """pairs: sequence of (P, Q) arrays of shape (points, 2|3)."""
        self.ctx = ctx or Context.default()
        self._keep = []
        n = len(pairs)
        Ps = (_lib._dp * n)()
        Qs = (_lib._dp * n)()
        ms = (ctypes.c_int64 * n)()
        ns = (ctypes.c_int64 * n)()
        ds = (ctypes.c_int32 * n)()
        for k, (P, Q) in enumerate(pairs):
            P = _lib.as_curve(P)
            Q = _lib.as_curve(Q)
            if P.shape[1] != Q.shape[1]:
                raise ValueError(f"pair {k}: curves of different dimension")
            self._keep += [P, Q]
            Ps[k] = _lib.dptr(P)
            Qs[k] = _lib.dptr(Q)
            ms[k], ns[k], ds[k] = len(P), len(Q), P.shape[1]
        h = ctypes.c_void_p()
        check(lib().rfd_batch_create(self.ctx.handle, n, Ps, ms, Qs, ns, ds, ctypes.byref(h)),
              self.ctx.handle)
        self.handle = h
        self.npairs = n
        self.sizes = np.array([(len(P), len(Q)) for P, Q in pairs], dtype=np.int64)
        self._keep = []  # the library copied the points

    def tune(self, **kw):
        """Times rows-per-lane 1..4 on this batch; returns (best, [ms1, ms2, ms3, ms4])."""
        o = _opts(**kw)
        r = ctypes.c_int32(0)
        ms = (ctypes.c_double * 4)()
        check(lib().rfd_batch_tune(self.handle, ctypes.byref(o), ctypes.byref(r), ms),
              self.ctx.handle)
        return r.value, list(ms)

    def run(self, *, stop: str = "rel", eps: float = 1e-9, vefd_gate: bool = True,
            rows_per_lane: int = 0, use_graph: bool = True, prune: int = -1,
            spec_depth: int = 0, decide_form: str = "fused") -> dict:
        """Returns arrays: distance, dfd, vefd, lb, ub, iterations, latency_ms, ...
        decide_form: "fused", "two-phase" or "reference" (per pair)."""
        o = _opts(stop=stop, eps=eps, vefd_gate=vefd_gate, decide_form=decide_form)
        o.rows_per_lane = rows_per_lane
        o.use_graph = 1 if use_graph else 0
        o.prune = prune
        o.spec_depth = spec_depth
        res = (_lib.Result * self.npairs)()
        lat = (ctypes.c_double * self.npairs)()
        check(lib().rfd_batch_run(self.handle, ctypes.byref(o), res, lat), self.ctx.handle)
        out = {name: np.array([getattr(r, name) for r in res]) for name, _ in _lib.Result._fields_}
        out["latency_ms"] = np.array(lat[:])
        return out

    def close(self):
        if getattr(self, "handle", None):
            lib().rfd_batch_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
