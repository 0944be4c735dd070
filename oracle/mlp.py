"""CPU reference of the selector MLP (TEST INFRASTRUCTURE ONLY).

n_in -> hidden (ReLU) -> hidden (ReLU) -> n_out, parameters flat in the
device layout: W1 [hidden][n_in], b1, W2 [hidden][hidden], b2, W3 [n_out][hidden],
b3.  Loss = sum over valid outputs of (y - t)^2 / count.  The gradients are
written out by the chain rule; tests pin them with central finite differences
in float64 (not by restating the same formulas).  There is no counterpart in
the reference (the selector is this build's addition).
"""
from __future__ import annotations

import numpy as np


def layout(n_in, hidden, n_out):
    dims = [n_in, hidden, hidden, n_out]
    out, o = [], 0
    for l in range(3):
        w = (o, o + dims[l + 1] * dims[l], (dims[l + 1], dims[l]))
        o = w[1]
        b = (o, o + dims[l + 1])
        o = b[1]
        out.append((w, b))
    return out, o


def unpack(p, n_in, hidden, n_out):
    lay, _ = layout(n_in, hidden, n_out)
    return [(p[w[0]:w[1]].reshape(w[2]), p[b[0]:b[1]]) for w, b in lay]


def forward(p, X, n_in, hidden, n_out):
    acts = [X]
    h = X
    for l, (W, b) in enumerate(unpack(p, n_in, hidden, n_out)):
        z = h @ W.T + b
        h = np.maximum(z, 0) if l < 2 else z
        acts.append(h)
    return h, acts


def loss(p, X, T, M, n_in, hidden, n_out):
    Y, _ = forward(p, X, n_in, hidden, n_out)
    cnt = float((M > 0).sum())
    d = np.where(M > 0, Y - T, 0.0)
    return float((d * d).sum() / cnt)


def grad(p, X, T, M, n_in, hidden, n_out):
    Y, acts = forward(p, X, n_in, hidden, n_out)
    cnt = float((M > 0).sum())
    dY = 2.0 * np.where(M > 0, Y - T, 0.0) / cnt
    params = unpack(p, n_in, hidden, n_out)
    lay, n = layout(n_in, hidden, n_out)
    g = np.zeros(n, dtype=p.dtype)
    d = dY
    for l in (2, 1, 0):
        W, b = params[l]
        (w0, w1, shp), (b0, b1) = lay[l]
        g[w0:w1] = (d.T @ acts[l]).ravel()
        g[b0:b1] = d.sum(axis=0)
        if l > 0:
            d = (d @ W) * (acts[l] > 0)
    return g


def finite_difference(p, X, T, M, n_in, hidden, n_out, idx, h=1e-6):
    """Central differences of the loss for the parameters in idx (float64)."""
    out = []
    for i in idx:
        q = p.astype(np.float64).copy()
        q[i] += h
        lp = loss(q, X, T, M, n_in, hidden, n_out)
        q[i] -= 2 * h
        lm = loss(q, X, T, M, n_in, hidden, n_out)
        out.append((lp - lm) / (2 * h))
    return np.array(out)
