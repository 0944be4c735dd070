"""Pins the CPU reference of the selector MLP (oracle/mlp.py): its chain-rule
gradient must match central finite differences of its own loss, and planted
mistakes in the backward pass must be caught by the same check."""
import numpy as np

from oracle import mlp


def _problem(seed, n_in=6, hidden=8, n_out=5, B=11):
    rng = np.random.default_rng(seed)
    _, n = mlp.layout(n_in, hidden, n_out)
    p = rng.standard_normal(n) * 0.5
    X = rng.standard_normal((B, n_in))
    T = rng.standard_normal((B, n_out))
    M = (rng.random((B, n_out)) < 0.7).astype(float)
    M[0, 0] = 1.0
    return p, X, T, M, (n_in, hidden, n_out)


def test_gradient_matches_finite_differences():
    for seed in range(3):
        p, X, T, M, dims = _problem(seed)
        g = mlp.grad(p, X, T, M, *dims)
        idx = np.arange(p.size)
        fd = mlp.finite_difference(p, X, T, M, *dims, idx)
        np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-7)


def test_masked_outputs_do_not_contribute():
    p, X, T, M, dims = _problem(4)
    T2 = T.copy()
    T2[M == 0] += 100.0
    assert mlp.loss(p, X, T, M, *dims) == mlp.loss(p, X, T2, M, *dims)
    np.testing.assert_array_equal(mlp.grad(p, X, T, M, *dims), mlp.grad(p, X, T2, M, *dims))


def test_planted_missing_relu_mask_is_caught():
    p, X, T, M, dims = _problem(5)

    def bad_grad(p, X, T, M, n_in, hidden, n_out):
        Y, acts = mlp.forward(p, X, n_in, hidden, n_out)
        cnt = float((M > 0).sum())
        d = 2.0 * np.where(M > 0, Y - T, 0.0) / cnt
        params = mlp.unpack(p, n_in, hidden, n_out)
        lay, n = mlp.layout(n_in, hidden, n_out)
        g = np.zeros(n)
        for l in (2, 1, 0):
            W, b = params[l]
            (w0, w1, _), (b0, b1) = lay[l]
            g[w0:w1] = (d.T @ acts[l]).ravel()
            g[b0:b1] = d.sum(axis=0)
            if l > 0:
                d = d @ W  # ReLU derivative forgotten
        return g

    fd = mlp.finite_difference(p, X, T, M, *dims, np.arange(p.size))
    assert not np.allclose(bad_grad(p, X, T, M, *dims), fd, rtol=1e-5, atol=1e-7)


def test_forward_known_values():
    # one hidden unit path: W1 = I-like, everything else identity scalars
    n_in, hidden, n_out = 1, 1, 1
    _, n = mlp.layout(n_in, hidden, n_out)
    p = np.array([2.0, -1.0, 3.0, 0.5, -1.0, 4.0])  # W1 b1 W2 b2 W3 b3
    assert n == 6
    y, _ = mlp.forward(p, np.array([[1.0], [0.25]]), n_in, hidden, n_out)
    # x=1: relu(2-1)=1 -> relu(3*1+0.5)=3.5 -> -3.5+4 = 0.5 ; x=0.25: relu(-0.5)=0 -> relu(0.5)=0.5 -> 3.5
    np.testing.assert_allclose(y.ravel(), [0.5, 3.5])
