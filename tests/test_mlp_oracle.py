"""The FP64 network oracle against finite differences and a closed form (CPU)."""
import numpy as np

from oracle import mlp_ref


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    widths = [5, 7, 6, 1]
    count = sum(a * b + b for a, b in zip(widths[:-1], widths[1:]))
    p = rng.standard_normal(count) * 0.5
    X = rng.standard_normal((11, 5))
    t = rng.standard_normal(11)
    loss, g = mlp_ref.loss_and_grad(p, widths, X, t)
    h = 1e-6
    for k in rng.choice(count, 25, replace=False):
        e = np.zeros(count)
        e[k] = h
        num = (mlp_ref.loss_and_grad(p + e, widths, X, t)[0] - mlp_ref.loss_and_grad(p - e, widths, X, t)[0]) / (2 * h)
        assert abs(num - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, num, g[k])


def test_linear_model_closed_form():
    """One linear layer: gradient of mean((Xw + b - t)^2) is 2/n X^T r, 2/n sum r."""
    rng = np.random.default_rng(1)
    X = rng.standard_normal((9, 3))
    w = rng.standard_normal(3)
    b = 0.3
    t = rng.standard_normal(9)
    p = np.concatenate([w, [b]])
    loss, g = mlp_ref.loss_and_grad(p, [3, 1], X, t)
    r = X @ w + b - t
    assert np.isclose(loss, np.mean(r * r))
    assert np.allclose(g[:3], 2.0 / 9 * X.T @ r)
    assert np.isclose(g[3], 2.0 / 9 * r.sum())
