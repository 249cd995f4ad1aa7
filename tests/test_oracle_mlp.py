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


def test_forward_and_gradient_match_torch():
    """The oracle's forward pass, Huber loss and gradient against torch's own
    layers in float64 (nn.functional.linear, gelu(approximate='tanh'),
    huber_loss with delta 1 and autograd): a wrong GELU constant, a transposed
    weight or a dropped bias fails here, which finite differences of the
    oracle's own loss cannot show."""
    import torch
    rng = np.random.default_rng(11)
    p = mlp.random_params(7)
    f = rng.normal(size=(13, 12))
    t = rng.normal(size=(13, 2))
    q = mlp.unpack(p)
    tq = {k: torch.tensor(v, dtype=torch.float64, requires_grad=k in dict(mlp.SHAPES))
          for k, v in q.items()}
    x = (torch.tensor(f) - tq["fm"]) / tq["fs"]
    a = x
    for i in (1, 2, 3):
        a = torch.nn.functional.gelu(torch.nn.functional.linear(a, tq[f"W{i}"], tq[f"b{i}"]),
                                     approximate="tanh")
    out = torch.nn.functional.linear(a, tq["W4"], tq["b4"])
    want_out, _ = mlp.forward(p, f)
    np.testing.assert_allclose(out.detach().numpy(), want_out, rtol=1e-13, atol=1e-13)
    target = (torch.tensor(t) - torch.tensor(f[:, mlp.BASE:mlp.BASE + mlp.OUT]) - tq["tm"]) / tq["ts"]
    loss = torch.nn.functional.huber_loss(out, target, reduction="sum", delta=1.0) / len(f)
    loss.backward()
    want_loss, want_grad = mlp.loss_grad(p, f, t)
    np.testing.assert_allclose(loss.item(), want_loss, rtol=1e-13)
    got = np.concatenate([tq[name].grad.numpy().ravel() for name, _ in mlp.SHAPES])
    np.testing.assert_allclose(got, want_grad, rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(mlp.predict(p, f), out.detach().numpy() * q["ts"] + q["tm"]
                               + f[:, mlp.BASE:mlp.BASE + mlp.OUT], rtol=1e-13, atol=1e-13)
