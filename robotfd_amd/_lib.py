"""ctypes binding of librfd.so (include/rfd.h).

The library is loaded from the package directory only; a missing library is
an error (there is no CPU fallback behind this interface).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
# RFD_LIB names another build of the same library (A/B timing of two builds).
LIB_PATH = os.environ.get("RFD_LIB") or os.path.join(PKG, "librfd.so")

RFD_OK = 0
STATUS_NAMES = {0: "ok", 1: "invalid argument", 2: "HIP runtime error", 3: "no GPU device",
                4: "inconsistent sweep result", 5: "device allocation failed"}

STOP_ABS, STOP_REL = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
# rfd_decide_form
DECIDE_FUSED, DECIDE_TWO_PHASE, DECIDE_REFERENCE = 0, 1, 2
DECIDE_FORMS = {"fused": DECIDE_FUSED, "two-phase": DECIDE_TWO_PHASE, "reference": DECIDE_REFERENCE}

K_DFD, K_VEFD64, K_DEC64, K_VEFD32, K_DEC32, K_BATCH, K_MLP, K_OTHER, K_BOUNDS64, K_BOUNDS_MIXED = range(10)
KERNEL_NAMES = ["dfd", "vefd_f64", "decide_f64", "vefd_f32", "decide_f32", "batch", "mlp", "other",
                "bounds_f64", "bounds_mixed", "kind10", "kind11"]
N_KINDS = 12


class Opts(ctypes.Structure):
    _fields_ = [("stop_mode", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("eps", ctypes.c_double), ("vefd_gate", ctypes.c_int32),
                ("verbose", ctypes.c_int32), ("mixed_switch", ctypes.c_double),
                ("rows_per_lane", ctypes.c_int32), ("use_graph", ctypes.c_int32),
                ("prune", ctypes.c_int32), ("spec_depth", ctypes.c_int32),
                ("decide_form", ctypes.c_int32), ("refine_ub", ctypes.c_int32)]


class Result(ctypes.Structure):
    _fields_ = [("distance", ctypes.c_double), ("endpoint_lb", ctypes.c_double),
                ("dfd", ctypes.c_double), ("vefd", ctypes.c_double),
                ("lb", ctypes.c_double), ("ub", ctypes.c_double),
                ("err_bound", ctypes.c_double), ("iterations", ctypes.c_int64),
                ("iterations_fp32", ctypes.c_int64), ("certify_steps", ctypes.c_int64),
                ("ms_total", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Stats(ctypes.Structure):
    _fields_ = [("launches", ctypes.c_int64 * N_KINDS), ("ms", ctypes.c_double * N_KINDS),
                ("cells", ctypes.c_double * N_KINDS), ("evaluated", ctypes.c_double * N_KINDS),
                ("skipped_blocks", ctypes.c_double * N_KINDS)]

    def as_dict(self) -> dict:
        return {KERNEL_NAMES[k]: dict(launches=int(self.launches[k]), ms=float(self.ms[k]),
                                      cells=float(self.cells[k]),
                                      evaluated=float(self.evaluated[k]),
                                      skipped_blocks=float(self.skipped_blocks[k]))
                for k in range(N_KINDS) if self.launches[k]}


# Every symbol include/rfd.h declares, with its ctypes signature.
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i64 = ctypes.c_int64
_fp = ctypes.POINTER(ctypes.c_float)
SIGNATURES = {
    "rfd_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int]),
    "rfd_destroy": (ctypes.c_int, [_vp]),
    "rfd_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "rfd_last_error": (ctypes.c_char_p, [_vp]),
    "rfd_default_opts": (None, [ctypes.POINTER(Opts)]),
    "rfd_get_stats": (ctypes.c_int, [_vp, ctypes.POINTER(Stats)]),
    "rfd_reset_stats": (ctypes.c_int, [_vp]),
    "rfd_session_create": (ctypes.c_int, [_vp, _dp, _i64, _dp, _i64, ctypes.c_int,
                                          ctypes.POINTER(_vp)]),
    "rfd_session_destroy": (ctypes.c_int, [_vp]),
    "rfd_session_bounds": (ctypes.c_int, [_vp, _dp, _dp, ctypes.c_int, ctypes.c_int]),
    "rfd_session_decide": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, _ip, _dp]),
    "rfd_session_set_decide_form": (ctypes.c_int, [_vp, ctypes.c_int32, _ip]),
    "rfd_session_refined_ub": (ctypes.c_int, [_vp, ctypes.c_int32, _dp]),
    "rfd_session_distance": (ctypes.c_int, [_vp, ctypes.POINTER(Opts), ctypes.POINTER(Result)]),
    "rfd_distance": (ctypes.c_int, [_vp, _dp, _i64, _dp, _i64, ctypes.c_int,
                                    ctypes.POINTER(Opts), ctypes.POINTER(Result)]),
    "rfd_distance_noalloc": (ctypes.c_int, [_vp, _vp, _i64, _vp, _i64, ctypes.c_int,
                                            ctypes.POINTER(Opts), ctypes.POINTER(Result)]),
    "rfd_batch_create": (ctypes.c_int, [_vp, _i64, ctypes.POINTER(_dp), ctypes.POINTER(_i64),
                                        ctypes.POINTER(_dp), ctypes.POINTER(_i64),
                                        ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(_vp)]),
    "rfd_batch_run": (ctypes.c_int, [_vp, ctypes.POINTER(Opts), ctypes.POINTER(Result), _dp]),
    "rfd_batch_tune": (ctypes.c_int, [_vp, ctypes.POINTER(Opts), ctypes.POINTER(ctypes.c_int32), _dp]),
    "rfd_batch_destroy": (ctypes.c_int, [_vp]),
    "rfd_mlp_param_count": (ctypes.c_int64, []),
    "rfd_mlp_create": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    "rfd_mlp_destroy": (ctypes.c_int, [_vp]),
    "rfd_mlp_get_params": (ctypes.c_int, [_vp, _fp]),
    "rfd_mlp_set_params": (ctypes.c_int, [_vp, _fp]),
    "rfd_mlp_save": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "rfd_mlp_load": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "rfd_mlp_set_policy": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double]),
    "rfd_mlp_get_policy": (ctypes.c_int, [_vp, _dp, _dp]),
    "rfd_mlp_forward": (ctypes.c_int, [_vp, _fp, _i64, _fp]),
    "rfd_mlp_gradients": (ctypes.c_int, [_vp, _fp, _fp, _i64, _fp, _fp]),
    "rfd_mlp_train": (ctypes.c_int, [_vp, _fp, _fp, _i64, ctypes.c_int32, ctypes.c_int32,
                                     ctypes.c_float, ctypes.c_uint64, _fp]),
    "rfd_batch_features": (ctypes.c_int, [_vp, _fp]),
    "rfd_batch_run_learned": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(Opts), ctypes.POINTER(Result),
                                             _dp, ctypes.POINTER(ctypes.c_int32)]),
}

_LIB = None


class RFDError(RuntimeError):
    pass


def lib():
    """The loaded library; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RFDError(f"{LIB_PATH} is missing: build it with `python -m robotfd_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(status: int, ctx=None) -> None:
    if status != RFD_OK:
        detail = ""
        if ctx is not None:
            detail = lib().rfd_last_error(ctx).decode(errors="replace")
        raise RFDError(f"rfd status {status} ({STATUS_NAMES.get(status, '?')}): {detail}")


def as_curve(a) -> np.ndarray:
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if arr.ndim != 2 or arr.shape[1] not in (2, 3):
        raise ValueError(f"curve must have shape (points, 2|3), got {arr.shape}")
    return arr


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_fp)
