"""The NumPy MLP reference (checker of the HIP training kernels) against finite differences."""
import numpy as np

import mlp_ref


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = mlp_ref.init(7).astype(np.float64)
    p[mlp_ref.NPARAM - 130:] += rng.normal(0, 0.1, 130)      # non-zero output biases / weights
    X = rng.normal(size=(37, 9))
    T = rng.normal(size=(37, 2))
    loss, g = mlp_ref.loss_grad(p, X, T, 1.0, 0.3)
    idx = rng.choice(mlp_ref.NPARAM, size=60, replace=False)
    for i in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h; pm[i] -= h
        fd = (mlp_ref.loss_only(pp, X, T, 1.0, 0.3) - mlp_ref.loss_only(pm, X, T, 1.0, 0.3)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    assert loss > 0


def test_parameter_count():
    assert mlp_ref.NPARAM == 9 * 64 + 64 + 2 * (64 * 64 + 64) + 64 * 2 + 2
