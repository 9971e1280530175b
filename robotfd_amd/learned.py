"""Learned bracket: the BF16 MFMA network that predicts where the distance lies
in [VE-FD, DFD], its training on records the exact method labels, and the
certified batch run (rfd_mlp_* and rfd_batch_run_learned in include/rfd.h).

Feature row (386 FP32): both curves resampled to 64 points by index, centred
on the mean of the 128 samples, divided by DFD (2 x 64 x 3), then VE-FD/DFD and
endpoint-bound/DFD.  Label: (distance - VE-FD) / (DFD - VE-FD), for pairs whose
bracket is not already closed.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib
from ._lib import check, fptr, lib
from .batch import Batch
from .frechet import Context, _opts

N_FEATURES = 386
PKG = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(PKG, "data", "bracket_net.bin")


def param_count() -> int:
    return int(lib().rfd_mlp_param_count())


class BracketNet:
    def __init__(self, seed: int = 0, ctx: Context | None = None, path: str | None = None):
        self.ctx = ctx or Context.default()
        h = ctypes.c_void_p()
        check(lib().rfd_mlp_create(self.ctx.handle, seed, ctypes.byref(h)), self.ctx.handle)
        self.handle = h
        if path:
            self.load(path)

    # -- parameters --------------------------------------------------------
    def get_params(self) -> np.ndarray:
        out = np.zeros(param_count(), dtype=np.float32)
        check(lib().rfd_mlp_get_params(self.handle, fptr(out)), self.ctx.handle)
        return out

    def set_params(self, params: np.ndarray) -> None:
        params = np.ascontiguousarray(params, dtype=np.float32)
        assert params.size == param_count()
        check(lib().rfd_mlp_set_params(self.handle, fptr(params)), self.ctx.handle)

    def save(self, path: str) -> None:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        check(lib().rfd_mlp_save(self.handle, path.encode()), self.ctx.handle)

    def load(self, path: str) -> None:
        check(lib().rfd_mlp_load(self.handle, path.encode()), self.ctx.handle)

    @property
    def policy(self):
        """(tau, delta) of the certification policy."""
        t, d = ctypes.c_double(0), ctypes.c_double(0)
        check(lib().rfd_mlp_get_policy(self.handle, ctypes.byref(t), ctypes.byref(d)), self.ctx.handle)
        return float(t.value), float(d.value)

    @policy.setter
    def policy(self, v) -> None:
        tau, delta = v
        check(lib().rfd_mlp_set_policy(self.handle, float(tau), float(delta)), self.ctx.handle)

    # -- compute -----------------------------------------------------------
    def forward(self, features: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(features, dtype=np.float32)
        out = np.zeros(len(X), dtype=np.float32)
        check(lib().rfd_mlp_forward(self.handle, fptr(X), len(X), fptr(out)), self.ctx.handle)
        return out

    def gradients(self, features: np.ndarray, labels: np.ndarray):
        X = np.ascontiguousarray(features, dtype=np.float32)
        y = np.ascontiguousarray(labels, dtype=np.float32)
        g = np.zeros(param_count(), dtype=np.float32)
        loss = ctypes.c_float(0)
        check(lib().rfd_mlp_gradients(self.handle, fptr(X), fptr(y), len(X), fptr(g),
                                      ctypes.byref(loss)), self.ctx.handle)
        return g, float(loss.value)

    def train(self, features: np.ndarray, labels: np.ndarray, epochs: int = 20, batch: int = 256,
              lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        X = np.ascontiguousarray(features, dtype=np.float32)
        y = np.ascontiguousarray(labels, dtype=np.float32)
        losses = np.zeros(epochs, dtype=np.float32)
        check(lib().rfd_mlp_train(self.handle, fptr(X), fptr(y), len(X), epochs, batch, lr, seed,
                                  fptr(losses)), self.ctx.handle)
        return losses

    def close(self):
        if getattr(self, "handle", None):
            lib().rfd_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_features(b: Batch) -> np.ndarray:
    """Features of every pair of `b` from the bounds of its last run."""
    out = np.zeros((b.npairs, N_FEATURES), dtype=np.float32)
    check(lib().rfd_batch_features(b.handle, fptr(out)), b.ctx.handle)
    return out


def run_learned(b: Batch, net: BracketNet, *, stop: str = "rel", eps: float = 1e-9,
                rows_per_lane: int = 0, use_graph: bool = True, spec_depth: int = 0) -> dict:
    o = _opts(stop=stop, eps=eps)
    o.rows_per_lane = rows_per_lane
    o.use_graph = 1 if use_graph else 0
    o.spec_depth = spec_depth
    res = (_lib.Result * b.npairs)()
    lat = (ctypes.c_double * b.npairs)()
    cert = (ctypes.c_int32 * b.npairs)()
    check(lib().rfd_batch_run_learned(b.handle, net.handle, ctypes.byref(o), res, lat, cert),
          b.ctx.handle)
    out = {name: np.array([getattr(r, name) for r in res]) for name, _ in _lib.Result._fields_}
    out["latency_ms"] = np.array(lat[:])
    out["certified"] = np.array(cert[:], dtype=bool)
    return out


def calibrate_policy(pred: np.ndarray, label: np.ndarray, steps: np.ndarray,
                     at_low: np.ndarray, at_high: np.ndarray):
    """Chooses tau (end-of-bracket threshold) on held-out records to maximise
    the decision sweeps saved: a certified end saves the record's bisection
    steps minus the certifying decision, a miss costs that decision.
    at_low / at_high: whether the distance lies within half the stop tolerance
    of the bracket's lower / upper end (when the one-decision bracket holds).
    Middle predictions are not certified (delta = 0): a bracket of useful width
    saves less than its two decisions cost on these workloads."""
    best_tau, best_gain = 0.0, 0.0
    for tau in np.concatenate([[0.0], np.quantile(np.minimum(pred, 1 - pred), np.linspace(0.02, 0.98, 49))]):
        lo = pred < tau
        hi = pred > 1 - tau
        gain = (np.where(at_low[lo], steps[lo] - 1, -1).sum() +
                np.where(at_high[hi], steps[hi] - 1, -1).sum())
        if gain > best_gain and tau <= 0.5:
            best_tau, best_gain = float(tau), float(gain)
    return best_tau, 0.0, best_gain


def end_flags(b: Batch, *, stop: str = "rel", eps: float = 1e-9, rows_per_lane: int = 0):
    """Whether the one-decision certification at each end of the bracket holds
    for every pair of `b`: runs the batch twice with a constant network that
    always answers 0 (low-end certification) and always answers 1 (high end),
    and reads which pairs certified."""
    net = BracketNet(seed=0, ctx=b.ctx)
    try:
        p = np.zeros(param_count(), dtype=np.float32)
        flags = []
        for bias in (-30.0, 30.0):
            p[-1] = bias  # head bias: sigmoid(-30) ~ 0, sigmoid(30) ~ 1
            net.set_params(p)
            net.policy = (0.01, 0.0)
            r = run_learned(b, net, stop=stop, eps=eps, rows_per_lane=rows_per_lane)
            flags.append(r["certified"])
        return flags[0], flags[1]
    finally:
        net.close()


def records_from_run(b: Batch, out: dict):
    """Training records (features, label, mask) from a completed plain run:
    label = (distance - VE-FD) / (DFD - VE-FD) where the bracket was open."""
    X = batch_features(b)
    w = out["dfd"] - out["vefd"]
    keep = w > 0
    y = np.zeros(b.npairs, dtype=np.float32)
    y[keep] = np.clip((out["distance"][keep] - out["vefd"][keep]) / w[keep], 0.0, 1.0)
    return X, y, keep


def save_records(path: str, X, y, vefd, dfd, distance) -> None:
    """Training-record file: features, label, and the (VE-FD, DFD, distance) it came from."""
    np.savez_compressed(path, features=X.astype(np.float32), label=y.astype(np.float32),
                        vefd=np.asarray(vefd, dtype=np.float64), dfd=np.asarray(dfd, dtype=np.float64),
                        distance=np.asarray(distance, dtype=np.float64))


def load_records(path: str):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}
