"""Independent pure-Python/numpy restatements used to pin the oracle.

These are written from the textbook definitions, not from the oracle's code:
Eiter-Mannila discrete Frechet, brute force over couplings and VE paths,
Alt-Godau free-space reachability with full reachable intervals, and the
Alt-Godau critical values.  Sizes are tiny; nothing here is fast.
"""
from __future__ import annotations

import itertools
import math

import numpy as np


def dfd_eiter_mannila(P: np.ndarray, Q: np.ndarray) -> float:
    m, n = len(P), len(Q)
    d = np.sqrt(((P[:, None, :] - Q[None, :, :]) ** 2).sum(-1))
    c = np.full((m, n), np.inf)
    for i in range(m):
        for j in range(n):
            if i == 0 and j == 0:
                c[i, j] = d[0, 0]
                continue
            prev = min(c[i - 1, j] if i else np.inf, c[i, j - 1] if j else np.inf,
                       c[i - 1, j - 1] if i and j else np.inf)
            c[i, j] = max(d[i, j], prev)
    return float(c[-1, -1])


def dfd_brute(P: np.ndarray, Q: np.ndarray) -> float:
    """Minimum over all monotone couplings (steps (1,0), (0,1), (1,1))."""
    m, n = len(P), len(Q)
    d = np.sqrt(((P[:, None, :] - Q[None, :, :]) ** 2).sum(-1))
    best = math.inf

    def walk(i, j, cur):
        nonlocal best
        cur = max(cur, d[i, j])
        if cur >= best:
            return
        if i == m - 1 and j == n - 1:
            best = cur
            return
        if i + 1 < m:
            walk(i + 1, j, cur)
        if j + 1 < n:
            walk(i, j + 1, cur)
        if i + 1 < m and j + 1 < n:
            walk(i + 1, j + 1, cur)

    walk(0, 0, 0.0)
    return best


def vefd_brute(m: int, n: int, weights, init: float) -> float:
    """VE-FD^2 by enumerating every monotone cell path (README.md:73-95).

    weights(i, j) -> (w_top, w_right) squared; a path from cell (0,0) to
    (m-2, n-2) leaves each cell through its right edge (i+1) or top edge (j+1),
    and the terminal cell through either.  Cost = max(init^2, weights)."""
    R, C = m - 1, n - 1
    W = {(i, j): weights(i, j) for i in range(R) for j in range(C)}
    best = math.inf
    moves = R - 1 + C - 1
    for rights in itertools.combinations(range(moves), R - 1):
        rs = set(rights)
        i = j = 0
        cost = init * init
        for k in range(moves):
            wt, wr = W[(i, j)]
            if k in rs:
                cost = max(cost, wr)
                i += 1
            else:
                cost = max(cost, wt)
                j += 1
        wt, wr = W[(i, j)]
        cost = max(cost, min(wt, wr))
        best = min(best, cost)
    return best


def vefd_readme(m: int, n: int, weights, init: float) -> float:
    """T/R recurrence of README.md:97-126 over full matrices, squared values."""
    T, Rr = vefd_readme_tables(m, n, weights, init)
    return float(min(T[-1, -1], Rr[-1, -1]))


def vefd_readme_tables(m: int, n: int, weights, init: float):
    """The full T and R tables of the README.md:97-126 recurrence."""
    R, C = m - 1, n - 1
    T = np.full((R, C), np.inf)
    Rr = np.full((R, C), np.inf)
    for i in range(R):
        for j in range(C):
            below = T[i, j - 1] if j > 0 else (np.inf)
            left = Rr[i - 1, j] if i > 0 else (init * init if j == 0 else np.inf)
            inn = min(below, left)
            wt, wr = weights(i, j)
            T[i, j] = max(wt, inn)
            Rr[i, j] = max(wr, inn)
    return T, Rr


def point_seg_dist(x, a, b) -> float:
    ab = b - a
    t = float(np.clip(np.dot(x - a, ab) / np.dot(ab, ab), 0.0, 1.0))
    return float(np.linalg.norm(x - (a + t * ab)))


def free_interval(c, s0, s1, r):
    """{t in [0,1] : |s0 + t (s1 - s0) - c| <= r} as (lo, hi) or None."""
    d = s1 - s0
    w = s0 - c
    a = float(d @ d)
    b = 2.0 * float(w @ d)
    cc = float(w @ w) - r * r
    disc = b * b - 4 * a * cc
    if disc < 0:
        return None
    sq = math.sqrt(disc)
    lo, hi = (-b - sq) / (2 * a), (-b + sq) / (2 * a)
    lo, hi = max(lo, 0.0), min(hi, 1.0)
    if lo > hi:
        return None
    return (lo, hi)


def decide_altgodau(P: np.ndarray, Q: np.ndarray, r: float) -> bool:
    """Alt-Godau decision with full reachable intervals on every cell boundary,
    boundary lines included."""
    m, n = len(P), len(Q)
    if np.linalg.norm(P[0] - Q[0]) > r or np.linalg.norm(P[-1] - Q[-1]) > r:
        return False
    # LF[i][j]: vertical line x = i (vertex p_i) over Q segment j.
    # BF[i][j]: horizontal line y = j (vertex q_j) over P segment i.
    LF = [[free_interval(P[i], Q[j], Q[j + 1], r) for j in range(n - 1)] for i in range(m)]
    BF = [[free_interval(Q[j], P[i], P[i + 1], r) for j in range(n)] for i in range(m - 1)]
    LR = [[None] * (n - 1) for _ in range(m)]
    BR = [[None] * n for _ in range(m - 1)]
    # Boundary x = 0: reachable while the free space is contiguous from (0,0).
    ok = True
    for j in range(n - 1):
        f = LF[0][j]
        if ok and f is not None and f[0] == 0.0:
            LR[0][j] = f
            ok = f[1] == 1.0
        else:
            ok = False
    ok = True
    for i in range(m - 1):
        f = BF[i][0]
        if ok and f is not None and f[0] == 0.0:
            BR[i][0] = f
            ok = f[1] == 1.0
        else:
            ok = False
    for i in range(m - 1):
        for j in range(n - 1):
            left, bot = LR[i][j], BR[i][j]
            fr, ft = LF[i + 1][j], BF[i][j + 1]
            if fr is not None:
                if bot is not None:
                    LR[i + 1][j] = fr
                elif left is not None:
                    lo = max(fr[0], left[0])
                    LR[i + 1][j] = (lo, fr[1]) if lo <= fr[1] else None
            if ft is not None:
                if left is not None:
                    BR[i][j + 1] = ft
                elif bot is not None:
                    lo = max(ft[0], bot[0])
                    BR[i][j + 1] = (lo, ft[1]) if lo <= ft[1] else None
    a = LR[m - 1][n - 2]
    b = BR[m - 2][n - 1]
    return (a is not None and a[1] == 1.0) or (b is not None and b[1] == 1.0)


def critical_values(P: np.ndarray, Q: np.ndarray) -> np.ndarray:
    """Alt-Godau candidate values: endpoints, vertex-edge distances, and the
    equidistance points of two vertices of one curve on a segment of the other."""
    out = [max(np.linalg.norm(P[0] - Q[0]), np.linalg.norm(P[-1] - Q[-1]))]
    for A, B in ((P, Q), (Q, P)):
        for x in A:
            for j in range(len(B) - 1):
                out.append(point_seg_dist(x, B[j], B[j + 1]))
        for k in range(len(A)):
            for l in range(k + 1, len(A)):
                for j in range(len(B) - 1):
                    s0, s1 = B[j], B[j + 1]
                    d = s1 - s0
                    # |s0 + t d - A_k|^2 = |s0 + t d - A_l|^2 is linear in t.
                    den = 2.0 * float(d @ (A[l] - A[k]))
                    if den == 0.0:
                        continue
                    t = (float(A[l] @ A[l] - A[k] @ A[k]) - 2.0 * float(s0 @ (A[l] - A[k]))) / den
                    if 0.0 <= t <= 1.0:
                        out.append(float(np.linalg.norm(s0 + t * d - A[k])))
    return np.unique(np.array(out))
