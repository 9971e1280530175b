"""Pins the network oracle: finite differences against its analytic backward
(FP64), BF16 rounding against known encodings, and the features' invariances."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(M.NPARAMS)
    for l in range(M.LAYERS):
        K = M.IN if l == 0 else M.H
        W = rng.standard_normal((K, M.H)) * np.sqrt(2.0 / (M.FEAT if l == 0 else M.H))
        if l == 0:
            W[M.FEAT:] = 0
        p[M.w_off(l):M.w_off(l) + K * M.H] = W.reshape(-1)
        p[M.b_off(l):M.b_off(l) + M.H] = rng.standard_normal(M.H) * 0.05
    p[M.W5:M.W5 + M.H] = rng.standard_normal(M.H) / 16
    p[M.B5] = 0.1
    return p


def test_layout_matches_parameter_count():
    assert M.NPARAMS == 400 * 256 + 256 + 3 * (256 * 256 + 256) + 256 + 1


def test_bf16_rounding_known_values():
    x = np.array([1.0, -2.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -8, 3.0e38, 1e-30], dtype=np.float32)
    got = M.bf16(x)
    assert got[0] == 1.0 and got[1] == -2.0
    assert got[2] == 1.0                      # tie rounds to even (down)
    assert got[3] == 1.0 + 2 ** -6            # tie rounds to even (up)
    assert np.isfinite(got[4]) and got[5] > 0


def test_analytic_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    params = init_params(2)
    X = rng.standard_normal((7, M.FEAT)) * 0.5
    y = rng.uniform(size=7)
    g = M.gradients(params, X, y, exact=True)
    # one coordinate in every block of the layout
    idx = []
    for l in range(M.LAYERS):
        K = M.IN if l == 0 else M.H
        idx += [M.w_off(l) + int(rng.integers(0, M.FEAT if l == 0 else K)) * M.H + int(rng.integers(0, M.H))
                for _ in range(4)]
        idx += [M.b_off(l) + int(rng.integers(0, M.H)) for _ in range(2)]
    idx += [M.W5 + int(rng.integers(0, M.H)) for _ in range(3)] + [M.B5]
    # Central differences at three step sizes: a ReLU kink within one step
    # spoils that estimate only, a wrong formula spoils all three.
    for i in idx:
        errs = []
        for h in (1e-6, 1e-7, 1e-8):
            pp, pm = params.copy(), params.copy()
            pp[i] += h
            pm[i] -= h
            fd = (M.loss(pp, X, y, exact=True) - M.loss(pm, X, y, exact=True)) / (2 * h)
            errs.append(abs(fd - g[i]) / max(abs(fd), 1e-7))
        assert min(errs) < 1e-4, (i, errs)


def test_bf16_network_tracks_fp64_network():
    rng = np.random.default_rng(3)
    params = init_params(4)
    X = rng.standard_normal((32, M.FEAT)) * 0.5
    y = rng.uniform(size=32)
    p32, _ = M.forward(params, X)
    p64, _ = M.forward(params, X, exact=True)
    assert np.max(np.abs(p32 - p64)) < 2e-2
    g32 = M.gradients(params, X, y)
    g64 = M.gradients(params, X, y, exact=True)
    assert np.linalg.norm(g32 - g64) < 0.1 * np.linalg.norm(g64)


def test_features_are_translation_and_scale_invariant():
    rng = np.random.default_rng(5)
    P = np.cumsum(rng.standard_normal((80, 3)), axis=0)
    Q = np.cumsum(rng.standard_normal((95, 3)), axis=0)
    f = M.features(P, Q, 3.0, 5.0, 1.0)
    shift = np.array([10.0, -4.0, 2.5])
    g = M.features(P + shift, Q + shift, 3.0, 5.0, 1.0)
    np.testing.assert_allclose(f, g, atol=1e-5)
    h = M.features(P * 4.0, Q * 4.0, 12.0, 20.0, 4.0)
    np.testing.assert_allclose(f, h, rtol=1e-5, atol=1e-6)
    # resampling keeps both endpoints
    assert np.allclose(f[0:3] * 5.0 + (np.concatenate([P[[0]], Q[[0]]]).mean(0) * 0), f[0:3] * 5.0)
    assert f[384] == np.float32(3.0 / 5.0) and f[385] == np.float32(1.0 / 5.0)


def test_features_2d_pad_z_with_zero():
    rng = np.random.default_rng(6)
    P = np.cumsum(rng.standard_normal((30, 2)), axis=0)
    Q = np.cumsum(rng.standard_normal((40, 2)), axis=0)
    f = M.features(P, Q, 1.0, 2.0, 0.5)
    assert np.all(f[2:192:3] == 0) and np.all(f[194:384:3] == 0)
