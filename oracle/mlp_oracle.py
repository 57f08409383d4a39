"""CPU oracle of the selector MLP (test infrastructure only): float64 numpy forward / backward of
hidden = relu(X W1^T + b1), y = hidden W2^T + b2, loss = sum mask (y - t)^2 / sum mask -- the
definition the HIP kernels implement (hacs_amd/csrc/hx_mlp.hip).  Pinned by finite differences
of the loss in tests/test_oracle.py."""
import numpy as np


def forward(X, W1, b1, W2, b2):
    h = np.maximum(X @ W1.T + b1, 0.0)
    return h, h @ W2.T + b2


def loss(X, W1, b1, W2, b2, T, mask=None):
    _, y = forward(X, W1, b1, W2, b2)
    mk = np.ones_like(y) if mask is None else mask
    return float((mk * (y - T) ** 2).sum() / max(mk.sum(), 1.0))


def backward(X, W1, b1, W2, b2, T, mask=None):
    h, y = forward(X, W1, b1, W2, b2)
    mk = np.ones_like(y) if mask is None else mask
    cnt = max(mk.sum(), 1.0)
    dy = 2.0 * mk * (y - T) / cnt
    dW2 = dy.T @ h
    db2 = dy.sum(0)
    dh = (dy @ W2) * (h > 0)
    dW1 = dh.T @ X
    db1 = dh.sum(0)
    return float((mk * (y - T) ** 2).sum() / cnt), dW1, db1, dW2, db2
