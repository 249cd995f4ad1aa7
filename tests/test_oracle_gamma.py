"""Pins the oracle's lgamma/digamma/trigamma against scipy.special and the
closed forms trigamma(1) = pi^2/6, trigamma(1/2) = pi^2/2, psi(1) = -euler."""
import numpy as np
from scipy import special

import oracle

X = np.concatenate([np.geomspace(1e-6, 1e6, 97), [0.5, 1.0, 2.0, 19.999, 20.0, 20.001, 21.0]])


def test_against_scipy():
    np.testing.assert_allclose(oracle.special("lgamma", X), special.gammaln(X), rtol=1e-13,
                               atol=1e-14)
    np.testing.assert_allclose(oracle.special("digamma", X), special.digamma(X), rtol=1e-13,
                               atol=1e-14)
    np.testing.assert_allclose(oracle.special("trigamma", X), special.polygamma(1, X),
                               rtol=1e-13)


def test_closed_forms():
    assert abs(oracle.special("trigamma", 1.0)[0] - np.pi ** 2 / 6) < 1e-14
    assert abs(oracle.special("trigamma", 0.5)[0] - np.pi ** 2 / 2) < 1e-13
    assert abs(oracle.special("digamma", 1.0)[0] + np.euler_gamma) < 1e-15
    x = np.geomspace(1e-3, 1e3, 50)
    np.testing.assert_allclose(oracle.special("digamma", x + 1) - oracle.special("digamma", x),
                               1 / x, rtol=1e-12)
    assert np.isnan(oracle.special("digamma", -1.0)[0])
