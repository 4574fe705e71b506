"""Pinning the FP64 predictor oracle (oracle/mlp_oracle.py) by central finite
differences of its loss -- CPU only."""

import numpy as np

import mlp_oracle as mo


def _setup(seed=0, m=16):
    rng = np.random.default_rng(seed)
    p = {}
    for n, s in mo.SHAPES:
        fan = s[1] if len(s) == 2 else 1
        p[n] = rng.standard_normal(s) * (np.sqrt(2.0 / fan) if len(s) == 2 else 0.1)
    X = rng.standard_normal((m, 192)) * 0.5
    T = rng.standard_normal((m, 24)) * 2.0
    return p, X, T


def test_gradients_match_finite_differences():
    p, X, T = _setup()
    g, _ = mo.gradients(p, X, T, bf=False)
    rng = np.random.default_rng(1)
    h = 1e-6
    for name, shape in mo.SHAPES:
        for _ in range(6):
            idx = tuple(rng.integers(0, s) for s in shape)
            q = {k: v.copy() for k, v in p.items()}
            q[name][idx] += h
            lp = mo.loss(q, X, T)
            q[name][idx] -= 2 * h
            lm = mo.loss(q, X, T)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[name][idx]) <= 1e-6 + 1e-5 * abs(fd), (name, idx, fd, g[name][idx])


def test_gelu_grad():
    x = np.linspace(-6, 6, 101)
    h = 1e-6
    fd = (mo.gelu(x + h) - mo.gelu(x - h)) / (2 * h)
    assert np.max(np.abs(fd - mo.gelu_grad(x))) < 1e-8


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.14159265])
    r = mo.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0 + 2 ** -7
    assert r[3] == -2.5 and abs(r[4] - 3.140625) < 1e-12
