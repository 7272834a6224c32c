"""The CPU reference of the learned preconditioner's network: its analytic
gradient against central finite differences of its loss (float64), and the
loss's per-G minimiser K = B/A."""
import numpy as np

import mlp_ref as M


def sample_set(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, size=(n, M.NF))
    A = rng.uniform(0.1, 1.0, size=n)
    B = A * np.exp(rng.uniform(-1, 2, size=n))
    return X, A, B


def test_gradient_matches_finite_differences():
    X, A, B = sample_set(200, 1)
    p = M.init_params(3).astype(np.float64)
    p += np.random.default_rng(4).normal(0, 0.1, size=p.size)    # biases off zero too
    L, g = M.loss_grad(X, A, B, p)
    assert L == M.loss(X, A, B, p)
    h = 1e-6
    idx = np.random.default_rng(5).choice(M.NP, 60, replace=False)
    idx = np.union1d(idx, [0, M.NF * M.NH, M.NP - M.NH - 1, M.NP - 1])   # every block
    for q in idx:
        e = np.zeros(M.NP); e[q] = h
        fd = (M.loss(X, A, B, p + e) - M.loss(X, A, B, p - e)) / (2 * h)
        assert abs(fd - g[q]) < 1e-6 * max(1.0, abs(fd)), (q, fd, g[q])


def test_loss_minimum_is_b_over_a():
    """with y free per sample (only b3 varies, one sample), the loss is
    minimal at K = B/A, where it equals -B^2/A"""
    X, A, B = sample_set(1, 2)
    p = M.init_params(6).astype(np.float64)
    p[:-1] = 0.0
    ks = np.linspace(0.01, 20, 4001)
    Ls = [M.loss(X, A, B, np.concatenate([p[:-1], [np.log(k)]])) for k in ks]
    kbest = ks[int(np.argmin(Ls))]
    assert abs(kbest - B[0] / A[0]) < 0.01
    assert min(Ls) == np.float64(min(Ls))
    assert abs(min(Ls) + B[0] ** 2 / A[0]) < 1e-6 * B[0] ** 2 / A[0]
