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


def _bf_case(seed=2, m=256):
    rng = np.random.default_rng(seed)
    p = {}
    for n, s in mo.SHAPES:
        fan = s[1] if len(s) == 2 else 1
        p[n] = (rng.standard_normal(s) * (np.sqrt(2.0 / fan) if len(s) == 2 else 0.1))
        p[n] = p[n].astype(np.float32).astype(np.float64)
    X = mo.bf16(rng.standard_normal((m, 192)) * 0.6)
    T = rng.standard_normal((m, 24)) * 3.0 + 2.0
    return p, X, T


def test_gradient_bar_catches_a_dropped_gelu_derivative_term(monkeypatch):
    # the bar the device backward pass is held to: FP32 instead of FP64
    # accumulation passes it, while dropping the x * phi(x) term of the GELU
    # derivative on one row in eight fails it
    p, X, T = _bf_case()
    gr, _ = mo.gradients(p, X, T, bf=True)
    g32 = {k: v.astype(np.float32).astype(np.float64) for k, v in gr.items()}
    assert mo.gradients_agree(g32, gr)
    good = mo.gelu_grad

    def planted(x):
        g = good(x)
        half = 0.5 * (1.0 + mo.erf(x / np.sqrt(2.0)))
        rows = (np.arange(x.shape[0]) % 8 == 3)[:, None]
        return np.where(rows, half, g)

    monkeypatch.setattr(mo, "gelu_grad", planted)
    gbad, _ = mo.gradients(p, X, T, bf=True)
    err = mo.gradient_errors(gbad, gr)
    assert not mo.gradients_agree(gbad, gr)
    # the layers below the planted derivative see it; w3 / b3 do not
    assert err["w3"][0] == 0.0 and err["w1"][0] > mo.GRAD_REL_TOL
