"""ORACLE — test infrastructure only (see oracle/oracle.py for the rules).

FP64 numpy restatement of the learned selector's network, loss and optimiser,
used to check the GPU kernels' forward outputs, gradients and Adam steps. The
network has no counterpart in the reference; this file is pinned by central
finite differences of the loss (tests/test_oracle_mlp.py) and by hand-computed
values on a one-unit network.

Parameter layout (one flat vector): W1[64][4], b1[64], W2[64][64], b2[64],
W3[nout][64], b3[nout].
"""
from __future__ import annotations

import numpy as np

IN = 4
H = 64


def num_params(nout: int) -> int:
    return H * IN + H + H * H + H + nout * H + nout


def unpack(p: np.ndarray, nout: int):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + H * IN].reshape(H, IN); o += H * IN
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + nout * H].reshape(nout, H); o += nout * H
    b3 = p[o:o + nout]; o += nout
    assert o == p.size
    return W1, b1, W2, b2, W3, b3


def forward(p, X, nout):
    W1, b1, W2, b2, W3, b3 = unpack(p, nout)
    X = np.asarray(X, np.float64)
    H1 = np.maximum(X @ W1.T + b1, 0.0)
    H2 = np.maximum(H1 @ W2.T + b2, 0.0)
    Y = H2 @ W3.T + b3
    return H1, H2, Y


def loss(p, X, T, Mk, nout) -> float:
    _, _, Y = forward(p, X, nout)
    cnt = float(np.sum(Mk))
    return float(np.sum(Mk * (Y - T) ** 2) / cnt) if cnt > 0 else 0.0


def backward(p, X, T, Mk, nout) -> tuple[float, np.ndarray]:
    """Masked mean-squared error and its gradient in the flat layout."""
    W1, b1, W2, b2, W3, b3 = unpack(p, nout)
    X = np.asarray(X, np.float64)
    H1, H2, Y = forward(p, X, nout)
    cnt = float(np.sum(Mk))
    dY = 2.0 * Mk * (Y - T) / cnt if cnt > 0 else np.zeros_like(Y)
    dW3 = dY.T @ H2
    db3 = dY.sum(0)
    dZ2 = (dY @ W3) * (H2 > 0)
    dW2 = dZ2.T @ H1
    db2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2) * (H1 > 0)
    dW1 = dZ1.T @ X
    db1 = dZ1.sum(0)
    g = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3.ravel(), db3])
    L = float(np.sum(Mk * (Y - T) ** 2) / cnt) if cnt > 0 else 0.0
    return L, g


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(nout: int, seed: int) -> np.ndarray:
    """He-style initialisation with a seeded generator (used by tests)."""
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0, np.sqrt(2 / IN), H * IN), np.zeros(H),
             rng.normal(0, np.sqrt(2 / H), H * H), np.zeros(H),
             rng.normal(0, np.sqrt(1 / H), nout * H), np.zeros(nout)]
    return np.concatenate(parts).astype(np.float32)
