"""Pins the numpy f64 MLP oracle: its gradient against central finite
differences of its own loss, GELU's closed forms, and the blob size the
library reports."""
import numpy as np

import mlp
import nebula_hip as nh


def test_blob_size_matches_library():
    assert mlp.COUNT == nh.mlp_param_count() == 9310


def test_gelu_closed_forms():
    assert mlp.gelu(0.0) == 0.0
    assert abs(mlp.gelu_grad(0.0) - 0.5) < 1e-15
    x = np.linspace(-4, 4, 17)
    # gelu(x) - gelu(-x) = x
    np.testing.assert_allclose(mlp.gelu(x) - mlp.gelu(-x), x, rtol=1e-14, atol=1e-14)
    h = 1e-6
    np.testing.assert_allclose((mlp.gelu(x + h) - mlp.gelu(x - h)) / (2 * h), mlp.gelu_grad(x),
                               rtol=1e-8, atol=1e-9)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = mlp.random_params(5)
    f = rng.normal(size=(9, 12))
    t = rng.normal(size=(9, 2))
    loss, g = mlp.loss_grad(p, f, t)
    base = 2 * mlp.IN
    idx = rng.choice(mlp.NET, size=60, replace=False)
    h = 1e-6
    for i in idx:
        pp = p.copy()
        pm = p.copy()
        pp[base + i] += h
        pm[base + i] -= h
        fd = (mlp.loss_grad(pp, f, t)[0] - mlp.loss_grad(pm, f, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])
