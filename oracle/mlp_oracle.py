"""CPU restatement of the bracket-position network and its features.

TEST INFRASTRUCTURE ONLY: the checker for the GPU network kernels.

Same parameter layout, BF16 rounding points and formulas as the device code:
hidden layers a = bf16(relu(bf16(x) @ bf16(W) + b)) accumulated in FP32, head
p = sigmoid(a @ w5 + b5), binary cross-entropy averaged over rows; backward
with dZ rounded to BF16 where the device rounds it.  The FP64 variant drops
every rounding and is what the finite-difference check pins.
"""
from __future__ import annotations

import numpy as np

IN, FEAT, H, LAYERS = 400, 386, 256, 4


def w_off(l: int) -> int:
    return 0 if l == 0 else IN * H + H + (l - 1) * (H * H + H)


def b_off(l: int) -> int:
    return w_off(l) + (IN if l == 0 else H) * H


W5 = w_off(3) + H * H + H
B5 = W5 + H
NPARAMS = B5 + 1


def bf16(x: np.ndarray) -> np.ndarray:
    """Round FP32 values to BF16 (nearest, ties to even) and return them as FP32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(params: np.ndarray, dtype=np.float32):
    p = np.asarray(params, dtype=dtype)
    Ws, bs = [], []
    for l in range(LAYERS):
        K = IN if l == 0 else H
        Ws.append(p[w_off(l):w_off(l) + K * H].reshape(K, H))
        bs.append(p[b_off(l):b_off(l) + H])
    return Ws, bs, p[W5:W5 + H], p[B5]


def pad(X: np.ndarray, dtype=np.float32) -> np.ndarray:
    X = np.asarray(X, dtype=dtype)
    out = np.zeros((len(X), IN), dtype=dtype)
    out[:, :FEAT] = X
    return out


def forward(params, X, exact: bool = False):
    """Returns (p, cache).  exact=True: FP64, no BF16 rounding."""
    dt = np.float64 if exact else np.float32
    r = (lambda a: a) if exact else bf16
    Ws, bs, w5, b5 = unpack(params, dt)
    A = [r(pad(X, dt))]
    for l in range(LAYERS):
        Z = A[-1] @ r(Ws[l]) + bs[l]
        A.append(r(np.maximum(Z, 0)))
    z = A[-1] @ w5 + b5
    p = 1.0 / (1.0 + np.exp(-z))
    return p.astype(dt), (A, z)


def loss(params, X, y, exact: bool = False) -> float:
    p, (A, z) = forward(params, X, exact)
    y = np.asarray(y, dtype=np.float64)
    z = z.astype(np.float64)
    return float(np.mean(np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z)))))


def gradients(params, X, y, exact: bool = False) -> np.ndarray:
    dt = np.float64 if exact else np.float32
    r = (lambda a: a) if exact else bf16
    Ws, bs, w5, b5 = unpack(params, dt)
    p, (A, z) = forward(params, X, exact)
    rows = len(X)
    g = np.zeros(NPARAMS, dtype=dt)
    dz = (p - np.asarray(y, dtype=dt)) / dt(rows)
    g[W5:W5 + H] = A[4].T @ dz
    g[B5] = dz.sum()
    dZ = r(dz[:, None] * w5[None, :]) * (A[4] > 0)
    for l in range(LAYERS - 1, -1, -1):
        K = IN if l == 0 else H
        g[w_off(l):w_off(l) + K * H] = (A[l].T @ dZ).reshape(-1)
        g[b_off(l):b_off(l) + H] = dZ.sum(axis=0)
        if l > 0:
            dZ = r(dZ @ r(Ws[l]).T) * (A[l] > 0)
    return g


def features(P: np.ndarray, Q: np.ndarray, vefd: float, dfd: float, endpoint_lb: float) -> np.ndarray:
    """Feature row of one pair (386 FP32), as the device's feature kernel forms it."""
    def resample(C):
        C3 = np.zeros((len(C), 3))
        C3[:, :C.shape[1]] = C
        if len(C) == 1:
            return np.repeat(C3, 64, axis=0)
        s = np.arange(64, dtype=np.float64)
        pos = s * float(len(C) - 1) / 63.0
        i0 = np.minimum(pos.astype(np.int64), len(C) - 2)
        f = pos - i0
        return C3[i0] + f[:, None] * (C3[i0 + 1] - C3[i0])
    a, b = resample(P), resample(Q)
    mean = np.concatenate([a, b]).sum(axis=0) / 128.0
    scale = dfd if dfd > 0 else 1.0
    out = np.zeros(FEAT, dtype=np.float32)
    out[:192] = ((a - mean) / scale).astype(np.float32).reshape(-1)
    out[192:384] = ((b - mean) / scale).astype(np.float32).reshape(-1)
    out[384] = np.float32(vefd / scale)
    out[385] = np.float32(endpoint_lb / scale)
    return out
