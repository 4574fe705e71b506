"""FP64 restatement of the predictor network (TEST INFRASTRUCTURE ONLY).

Checks the device forward / backward kernels of csrc/mlp.hip.  The network
is y = W3 g(W2 g(W1 x + b1) + b2) + b3 with g the erf GELU; the loss is the
mean squared error over rows and the 24 outputs.  With ``bf16=True`` the
values the device keeps in BF16 (weights, inputs, hidden activations, the
output gradient and the pre-activation gradients) are rounded to BF16 at the
same points, so the comparison isolates accumulation-order differences.
Pinned by central finite differences of the plain FP64 network
(tests/test_mlp_oracle.py).  There is no reference counterpart.
"""

from __future__ import annotations

import numpy as np
from scipy.special import erf

SHAPES = [("w1", (256, 192)), ("b1", (256,)), ("w2", (256, 256)), ("b2", (256,)),
          ("w3", (24, 256)), ("b3", (24,))]


def unflatten(flat):
    out, o = {}, 0
    for n, s in SHAPES:
        k = int(np.prod(s))
        out[n] = np.asarray(flat[o:o + k], np.float64).reshape(s)
        o += k
    return out


def flatten(d):
    return np.concatenate([np.asarray(d[n], np.float64).ravel() for n, _ in SHAPES])


def bf16(x):
    """Round to BF16 (nearest even), returned as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2 * np.pi)


def forward(p, X, bf=False):
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    W1, W2, W3 = r(p["w1"]), r(p["w2"]), r(p["w3"])
    X = r(X)
    Z1 = X @ W1.T + p["b1"]
    H1 = r(gelu(Z1))
    Z2 = H1 @ W2.T + p["b2"]
    H2 = r(gelu(Z2))
    Y = H2 @ W3.T + p["b3"]
    return dict(X=X, Z1=Z1, H1=H1, Z2=Z2, H2=H2, Y=Y, W1=W1, W2=W2, W3=W3)


def loss(p, X, T, bf=False):
    Y = forward(p, X, bf)["Y"]
    return float(np.mean((Y - T) ** 2))


def gradients(p, X, T, bf=False):
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    f = forward(p, X, bf)
    M = X.shape[0]
    dY = r(2.0 * (f["Y"] - T) / (M * T.shape[1]))
    g = {}
    g["w3"] = dY.T @ f["H2"]
    g["b3"] = dY.sum(0)
    dZ2 = r((dY @ f["W3"]) * gelu_grad(f["Z2"]))
    g["w2"] = dZ2.T @ f["H1"]
    g["b2"] = dZ2.sum(0)
    dZ1 = r((dZ2 @ f["W2"]) * gelu_grad(f["Z1"]))
    g["w1"] = dZ1.T @ f["X"]
    g["b1"] = dZ1.sum(0)
    return g, float(np.mean((f["Y"] - T) ** 2))


# Agreement bars of a device backward pass with gradients(..., bf=True):
# per tensor, the error norm relative to the reference norm and the largest
# element error relative to the largest reference element.
GRAD_REL_TOL = 2e-3
GRAD_MAX_TOL = 1e-2


def gradient_errors(g, gr):
    """{name: (|g - gr| / |gr|, max|g - gr| / max|gr|)} per parameter tensor."""
    out = {}
    for name, _ in SHAPES:
        a = np.asarray(g[name], np.float64)
        b = np.asarray(gr[name], np.float64)
        nb = max(float(np.linalg.norm(b)), 1e-30)
        mb = max(float(np.abs(b).max()), 1e-30)
        out[name] = (float(np.linalg.norm(a - b)) / nb, float(np.abs(a - b).max()) / mb)
    return out


def gradients_agree(g, gr):
    return all(r <= GRAD_REL_TOL and m <= GRAD_MAX_TOL for r, m in gradient_errors(g, gr).values())
