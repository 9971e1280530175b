"""rfd_distance_noalloc (curves already in device memory, every working buffer
in the library's fixed per-GPU array, nothing allocated) against rfd_distance
on the same curves: the same results bit for bit, on two calls in a row, the
second on smaller curves, and a clear error past the array's size limit."""
import ctypes
import sys

import numpy as np
import pytest

from robotfd_amd import Context, _lib
from robotfd_amd._lib import RFDError, check, lib
from robotfd_amd.frechet import _opts, distance

pytestmark = pytest.mark.gpu

FIELDS = ("distance", "endpoint_lb", "dfd", "vefd", "lb", "ub", "err_bound", "iterations")
MAX_POINTS = 1 << 19  # RFD_NOALLOC_MAX_POINTS


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class DeviceArray:
    """A device copy of a host FP64 array, through the HIP runtime."""

    def __init__(self, hip, host: np.ndarray):
        host = np.ascontiguousarray(host, dtype=np.float64)
        self.hip = hip
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(host.nbytes)) == 0
        assert hip.hipMemcpy(self.ptr, host.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(host.nbytes),
                             1) == 0  # hipMemcpyHostToDevice

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = None


@pytest.fixture(scope="module")
def hip():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


def noalloc(ctx, dP, m, dQ, n, dims, **kw):
    res = _lib.Result()
    o = _opts(**kw)
    check(lib().rfd_distance_noalloc(ctx.handle, dP.ptr, m, dQ.ptr, n, dims, ctypes.byref(o),
                                     ctypes.byref(res)), ctx.handle)
    return res.as_dict()


def walk(rng, k, d):
    return np.cumsum(rng.standard_normal((k, d)), axis=0)


def test_noalloc_equals_rfd_distance_twice_then_smaller(hip):
    ctx = Context()
    rng = np.random.default_rng(2024)
    cases = [(walk(rng, 3000, 3), walk(rng, 2600, 3), dict(stop="rel", eps=1e-9)),
             (walk(rng, 700, 2), walk(rng, 450, 2), dict(stop="abs", eps=float(np.finfo(np.float32).eps)))]
    # Noisy copies give open brackets (decisions run) as well as collapsed ones.
    P3 = walk(rng, 1500, 3)
    Q3 = P3[::2] + rng.standard_normal((750, 3)) * 0.4
    cases.insert(1, (P3, Q3, dict(stop="rel", eps=1e-9)))
    try:
        for k, (P, Q, kw) in enumerate(cases):
            dims = P.shape[1]
            say("case", k, P.shape, Q.shape, kw)
            ref = distance(P, Q, ctx=ctx, **kw)
            dP, dQ = DeviceArray(hip, P), DeviceArray(hip, Q)
            try:
                got = noalloc(ctx, dP, len(P), dQ, len(Q), dims, **kw)
            finally:
                dP.free()
                dQ.free()
            for f in FIELDS:
                assert got[f] == ref[f], (k, f, got[f], ref[f])
        # Two-phase decisions (no extra buffers either).
        P, Q, kw = cases[1]
        ref = distance(P, Q, ctx=ctx, decide_form="two-phase", **kw)
        dP, dQ = DeviceArray(hip, P), DeviceArray(hip, Q)
        try:
            got = noalloc(ctx, dP, len(P), dQ, len(Q), 3, decide_form="two-phase", **kw)
        finally:
            dP.free()
            dQ.free()
        for f in FIELDS:
            assert got[f] == ref[f], ("two-phase", f, got[f], ref[f])
    finally:
        ctx.close()


def test_noalloc_rejects_oversize_mixed_and_non_finite(hip):
    ctx = Context()
    rng = np.random.default_rng(5)
    P, Q = walk(rng, 40, 2), walk(rng, 30, 2)
    dP, dQ = DeviceArray(hip, P), DeviceArray(hip, Q)
    try:
        with pytest.raises(RFDError, match="fixed working array"):
            noalloc(ctx, dP, MAX_POINTS + 1, dQ, len(Q), 2)  # refused before any access
        with pytest.raises(RFDError, match="FP64 only"):
            noalloc(ctx, dP, len(P), dQ, len(Q), 2, precision="mixed")
    finally:
        dP.free()
        dQ.free()
    P[7, 0] = np.nan
    dP, dQ = DeviceArray(hip, P), DeviceArray(hip, Q)
    try:
        with pytest.raises(RFDError, match="non-finite"):
            noalloc(ctx, dP, len(P), dQ, len(Q), 2)
    finally:
        dP.free()
        dQ.free()
    ctx.close()
