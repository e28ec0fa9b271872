"""The float64 restatement of the selector perceptron (oracle/mlp_oracle.py)
pinned by central finite differences: a wrong transpose, a dropped ReLU mask
or a wrong softmax sign changes the gradient by O(1) relative."""
import numpy as np

from oracle import mlp_oracle as M


def data(n=64, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, M.IN)) * 2 + 0.5
    y = (X[:, 0] + 0.3 * X[:, 3] > 0.5).astype(np.int64)
    mu, sig = X.mean(0), X.std(0) + 0.1
    return X, y, mu, sig


def test_gradient_matches_central_differences():
    X, y, mu, sig = data()
    p = M.init(3)
    _, g = M.loss_grad(p, mu, sig, X, y)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(M.NPARAMS, 200, replace=False), np.arange(M.NPARAMS - 130, M.NPARAMS)])
    h = 1e-6
    for i in idx:
        e = np.zeros(M.NPARAMS)
        e[i] = h
        fd = (M.loss(p + e, mu, sig, X, y) - M.loss(p - e, mu, sig, X, y)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_layout_and_probabilities():
    X, y, mu, sig = data(8)
    p = M.init(0)
    f = M.forward(p, mu, sig, X)
    assert M.NPARAMS == 12 * 64 + 64 + 64 * 64 + 64 + 64 * 2 + 2
    assert np.allclose(f["prob"].sum(1), 1.0)
    # a positive bias on class 1 raises its probability everywhere
    q = p.copy()
    q[-1] += 5.0
    assert np.all(M.forward(q, mu, sig, X)["prob"][:, 1] > f["prob"][:, 1])
