"""ORACLE — test infrastructure only; nothing in the product imports this.

Float64 numpy restatement of the selector perceptron that the library trains
and evaluates on the GPU (12 -> 64 -> 64 -> 2, ReLU, softmax cross-entropy),
with the same parameter layout: W1 (12 x 64 row-major), b1 (64), W2 (64 x 64),
b2 (64), W3 (64 x 2), b3 (2); inputs standardized by (x - mu) / sig.  The
network has no counterpart in the reference; this restatement is pinned by
the central finite differences of tests/test_mlp_oracle.py.
"""
from __future__ import annotations

import numpy as np

IN, HID, OUT = 12, 64, 2
SIZES = [IN * HID, HID, HID * HID, HID, HID * OUT, OUT]
NPARAMS = int(sum(SIZES))


def unpack(p):
    p = np.asarray(p, np.float64)
    parts, o = [], 0
    for sz in SIZES:
        parts.append(p[o:o + sz])
        o += sz
    W1, b1, W2, b2, W3, b3 = parts
    return W1.reshape(IN, HID), b1, W2.reshape(HID, HID), b2, W3.reshape(HID, OUT), b3


def init(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAMS)
    o = 0
    for (fi, fo), sz in zip([(IN, HID), (1, 0), (HID, HID), (1, 0), (HID, OUT), (1, 0)], SIZES):
        if fo:
            a = np.sqrt(6.0 / (fi + fo))
            p[o:o + sz] = rng.uniform(-a, a, sz)
        o += sz
    return p


def forward(p, mu, sig, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = (np.asarray(X, np.float64) - mu) / sig
    z1 = x @ W1 + b1
    a1 = np.maximum(z1, 0.0)
    z2 = a1 @ W2 + b2
    a2 = np.maximum(z2, 0.0)
    z3 = a2 @ W3 + b3
    zm = z3.max(1, keepdims=True)
    e = np.exp(z3 - zm)
    prob = e / e.sum(1, keepdims=True)
    return dict(x=x, z1=z1, a1=a1, z2=z2, a2=a2, z3=z3, prob=prob)


def loss(p, mu, sig, X, y):
    f = forward(p, mu, sig, X)
    z3 = f["z3"]
    zm = z3.max(1)
    lse = zm + np.log(np.exp(z3 - zm[:, None]).sum(1))
    return float(np.mean(lse - z3[np.arange(len(y)), y]))


def loss_grad(p, mu, sig, X, y):
    """Mean cross-entropy and its gradient (float64)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    f = forward(p, mu, sig, X)
    n = len(y)
    Y = np.zeros((n, OUT))
    Y[np.arange(n), y] = 1.0
    d3 = (f["prob"] - Y) / n
    gW3 = f["a2"].T @ d3
    gb3 = d3.sum(0)
    d2 = (d3 @ W3.T) * (f["z2"] > 0)
    gW2 = f["a1"].T @ d2
    gb2 = d2.sum(0)
    d1 = (d2 @ W2.T) * (f["z1"] > 0)
    gW1 = f["x"].T @ d1
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss(p, mu, sig, X, y), g
