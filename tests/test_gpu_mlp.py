"""The warm-start MLP's device kernels against the numpy f64 oracle
(oracle/mlp.py), through the C ABI: FP32 training gradients, BF16-MFMA
inference, and training itself.  Marked gpu."""
import numpy as np
import pytest

import mlp
import nebula_hip as nh

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 37, 256])
def test_training_gradient_matches_f64(n):
    """Loss and every gradient entry of the FP32 forward/backward kernels
    agree with the f64 oracle to FP32 accumulation error."""
    rng = np.random.default_rng(n)
    p = mlp.random_params(11 + n)
    f = rng.normal(size=(n, 12)) * p[12:24] + p[:12]
    t = rng.normal(size=(n, 2)) * 2.0
    loss, g = nh.mlp_loss_grad(p, f, t)
    p32 = p.astype(np.float32).astype(np.float64)
    want_loss, want = mlp.loss_grad(p32, f.astype(np.float32), t.astype(np.float32))
    assert abs(loss - want_loss) <= 1e-5 * max(1.0, abs(want_loss))
    scale = np.abs(want).max()
    err = np.abs(g - want)
    assert err.max() <= 1e-4 * scale + 1e-7, (err.max(), scale, int(np.argmax(err)))


@pytest.mark.parametrize("n", [5, 64, 100, 1000])
def test_mfma_inference_matches_bf16_reference(n):
    """BF16 weights and layer inputs, FP32 accumulation: equal to the oracle's
    forward with the same roundings up to accumulation order, and to the f64
    forward within BF16 resolution."""
    rng = np.random.default_rng(100 + n)
    p = mlp.random_params(7)
    f = rng.normal(size=(n, 12)) * p[12:24] + p[:12]
    got = nh.mlp_predict(p, f)
    p32 = p.astype(np.float32).astype(np.float64)
    f32 = f.astype(np.float32).astype(np.float64)
    want_bf16 = mlp.predict(p32, f32, rnd=mlp.bf16)
    want_f64 = mlp.predict(p32, f32)
    scale = p32[-2:]  # target sd: outputs are unstandardised
    assert np.all(np.isfinite(got))
    np.testing.assert_allclose(got, want_bf16, rtol=0, atol=float(3e-2 * scale.max()))
    np.testing.assert_allclose(got, want_f64, rtol=0, atol=float(0.15 * scale.max()))
    # the BF16 path is much closer to its own emulation than to f64
    assert np.abs(got - want_bf16).mean() < np.abs(got - want_f64).mean() + 1e-6


def test_training_fits_a_smooth_map():
    """Adam on the device drives the loss down on a smooth 12 -> 2 map, and
    the trained blob predicts the same on the matrix cores as in f64."""
    rng = np.random.default_rng(0)
    n = 4096
    f = rng.normal(size=(n, 12))
    t = np.c_[np.sin(f[:, 0]) + 0.5 * f[:, 1] * f[:, 2], np.tanh(f[:, 3] - f[:, 4]) + 0.1 * f[:, 5]]
    params, loss = nh.mlp_train(f, t, epochs=40, batch=128, lr=3e-3, seed=1)
    q = mlp.unpack(params)
    # standardisation taken from the data; the network fits the correction to
    # features 10 and 11
    r = t - f[:, 10:12]
    np.testing.assert_allclose(q["fm"], f.mean(0), atol=1e-5)
    np.testing.assert_allclose(q["ts"], r.std(0), rtol=1e-4)
    e = (r - r.mean(0)) / r.std(0)
    untrained_loss = np.sum(np.where(np.abs(e) <= 1, 0.5 * e * e, np.abs(e) - 0.5)) / n
    assert loss < 0.05 * untrained_loss, (loss, untrained_loss)
    loss_f64, _ = mlp.loss_grad(params.astype(np.float64), f, t)
    assert abs(loss_f64 - loss) <= 1e-3 * untrained_loss
    pred = nh.mlp_predict(params, f[:500])
    np.testing.assert_allclose(pred, mlp.predict(params.astype(np.float64), f[:500]), atol=0.05)
    assert np.abs(pred - t[:500]).mean() < 0.1 * np.abs(t[:500] - t.mean(0)).mean()
