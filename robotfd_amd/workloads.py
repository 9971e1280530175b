"""Synthetic trajectory generators for the benchmark and parity workloads.

All curves are FP64 arrays of shape (points, dims), C-contiguous (the AoS
layout the library consumes, point i at [i*dims + k]).  The random stream is
numpy's PCG64 seeded with the given seed; curve 1 is drawn in full before
curve 2.  A random walk starts at the origin and adds i.i.d. N(0, sigma) steps
per axis (sigma is the standard deviation).
"""
from __future__ import annotations

import hashlib

import numpy as np


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def random_walk(rng: np.random.Generator, points: int, dims: int, sigma: float = 1.0) -> np.ndarray:
    """Walk of `points` points: x_0 = 0, x_k = x_{k-1} + N(0, sigma) per axis."""
    steps = rng.standard_normal((points - 1, dims)) * sigma
    out = np.zeros((points, dims), dtype=np.float64)
    np.cumsum(steps, axis=0, out=out[1:])
    return out


def walk_pair(m: int, n: int, dims: int, seed: int = 0, sigma: float = 1.0):
    """Two independent random walks (m and n points) from one seeded stream."""
    rng = _rng(seed)
    P = random_walk(rng, m, dims, sigma)
    Q = random_walk(rng, n, dims, sigma)
    return P, Q


def small_2d(seed: int = 0):
    """2D random walks of 2,000 and 2,500 points, N(0,1) steps (CPU-checkable case)."""
    return walk_pair(2000, 2500, 2, seed)


def robot_pair(m: int = 100_000, n: int = 120_000, seed: int = 0):
    """3D robot-like pair.

    P: random walk with N(0, 0.05) steps over m + 15 points, smoothed by a
    16-point moving average (valid mode) down to m points.  Q: P plus
    N(0, 0.2) noise per axis, then resampled linearly by index to n points.
    """
    rng = _rng(seed)
    raw = random_walk(rng, m + 15, 3, 0.05)
    c = np.cumsum(np.vstack([np.zeros((1, 3)), raw]), axis=0)
    P = (c[16:] - c[:-16]) / 16.0
    noisy = P + rng.standard_normal(P.shape) * 0.2
    src = np.linspace(0.0, m - 1.0, n)
    Q = np.empty((n, 3), dtype=np.float64)
    for k in range(3):
        Q[:, k] = np.interp(src, np.arange(m, dtype=np.float64), noisy[:, k])
    return np.ascontiguousarray(P), Q


def long_3d(m: int = 200_000, n: int = 200_000, seed: int = 0):
    """3D random walks, N(0,1) steps (the 200k x 200k precision workload)."""
    return walk_pair(m, n, 3, seed)


def batch_pairs(count: int = 4096, lo: int = 500, hi: int = 20_000, seed: int = 0):
    """`count` pairs; each curve's length log-uniform in [lo, hi], drawn
    independently; dims 2 or 3 with probability 1/2 per pair; N(0,1) walks.
    Returns a list of (P, Q) arrays."""
    rng = _rng(seed)
    lens = np.exp(rng.uniform(np.log(lo), np.log(hi), size=(count, 2))).astype(np.int64)
    dims = np.where(rng.uniform(size=count) < 0.5, 2, 3)
    pairs = []
    for c in range(count):
        P = random_walk(rng, int(lens[c, 0]), int(dims[c]))
        Q = random_walk(rng, int(lens[c, 1]), int(dims[c]))
        pairs.append((P, Q))
    return pairs


def sha256(a: np.ndarray) -> str:
    """SHA-256 of the array's raw little-endian FP64 bytes."""
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()
