"""Pins the oracle's L-BFGS-B (oracle/lbfgsb.hpp, a restatement of
src/numeric/lbfgsb.rs: compact representation, generalised Cauchy point,
subspace minimisation, strong-Wolfe line search) at the level the reference
pins its own: the known-answer problems of src/numeric/lbfgsb.rs:1040-1296
(Rosenbrock free and boxed, a separable quadratic with and without an active
bound, a degenerate box, an infeasible start, NEBULA's stage-one shape
against scipy's values, the badly scaled box whose corner the Cauchy point
must reach to 1e-12) at the reference's tolerances; and against scipy's
L-BFGS-B (the Fortran code of Byrd, Lu, Nocedal and Zhu) on random boxed
quadratics, where the optimum is unique.  Iterates are not compared with
scipy's: its line search is More and Thuente's, the reference's a
strong-Wolfe bisection."""
import numpy as np
import pytest
from scipy.optimize import minimize

import oracle

INF = np.inf


def rosenbrock(x):
    a, b = x
    f = 100.0 * (b - a * a) ** 2 + (1.0 - a) ** 2
    return f, np.array([-400.0 * a * (b - a * a) - 2.0 * (1.0 - a), 200.0 * (b - a * a)])


def quadratic(x):
    d = x - np.array([1.0, 2.0, 3.0])
    return float(d @ d), 2.0 * d


def nebula_shaped(p):
    # src/numeric/lbfgsb.rs:1192-1226: NB negative log-likelihood with a log
    # link and a quadratic penalty on sigma, 60 cells
    n = 60
    counts = np.array([[0.0, 3.0, 1.0, 2.0, 0.0, 5.0, 2.0, 1.0, 4.0, 2.0][i % 10]
                       for i in range(n)])
    off = np.log(800.0)
    b0, b1, sigma, phi = p
    r = 1.0 / phi
    xi = -1.0 + 2.0 * np.arange(n) / (n - 1)
    mu = np.exp(off + b0 + b1 * xi)
    f = 0.5 * sigma * sigma + np.sum((counts + r) * np.log(mu + r) - counts * np.log(mu)
                                     - r * np.log(r))
    d_mu = (counts + r) / (mu + r) - counts / mu
    dr = np.sum(np.log(mu + r) + (counts + r) / (mu + r) - np.log(r) - 1.0)
    return f, np.array([np.sum(d_mu * mu), np.sum(xi * d_mu * mu), sigma, -dr / phi ** 2])


def decaying(x):
    rate = np.array([1.0, 30.0, 300.0, 3000.0])
    e = np.exp(-rate * x)
    return float(e.sum()), -rate * e


def test_unbounded_rosenbrock_finds_the_optimum():
    x, f, ok = oracle.lbfgsb(rosenbrock, [-1.2, 1.0], [-INF, -INF], [INF, INF], 1e-14, 1e-10,
                             500)
    assert ok
    np.testing.assert_allclose(x, [1.0, 1.0], rtol=1e-5)
    assert f < 1e-10


def test_quadratic_converges():
    x, _, ok = oracle.lbfgsb(quadratic, [0.0, 0.0, 0.0], [-INF] * 3, [INF] * 3, 1e-14, 1e-12)
    assert ok
    np.testing.assert_allclose(x, [1.0, 2.0, 3.0], rtol=1e-8)


def test_active_bound_is_respected():
    x, _, _ = oracle.lbfgsb(quadratic, [5.0, 0.0, 0.0], [2.0, -INF, -INF], [INF] * 3, 1e-14,
                            1e-12)
    np.testing.assert_allclose(x[0], 2.0, rtol=1e-10)
    np.testing.assert_allclose(x[1:], [2.0, 3.0], rtol=1e-8)


def test_degenerate_box_returns_the_only_feasible_point():
    x, _, _ = oracle.lbfgsb(quadratic, [0.0, 0.0, 0.0], [1.5] * 3, [1.5] * 3)
    np.testing.assert_allclose(x, [1.5] * 3, rtol=1e-12)


def test_bounded_rosenbrock_stops_on_the_boundary():
    x, _, _ = oracle.lbfgsb(rosenbrock, [-1.2, 1.0], [-INF, -INF], [0.5, INF], 1e-14, 1e-10,
                            500)
    np.testing.assert_allclose(x[0], 0.5, rtol=1e-6)
    np.testing.assert_allclose(x[1], 0.25, rtol=1e-4)


def test_infeasible_start_is_clamped():
    lo, hi = np.zeros(3), np.full(3, 10.0)
    x, _, _ = oracle.lbfgsb(quadratic, [-50.0, 100.0, 3.0], lo, hi)
    assert np.all(x >= lo - 1e-12) and np.all(x <= hi + 1e-12)
    np.testing.assert_allclose(x, [1.0, 2.0, 3.0], rtol=1e-6)


def test_nebula_shaped_problem_matches_scipy():
    # values the reference quotes from scipy's L-BFGS-B (ftol 1e-12, gtol
    # 1e-10), src/numeric/lbfgsb.rs:1229-1256
    x, f, _ = oracle.lbfgsb(nebula_shaped, [0.0, 0.0, 1.0, 1.0], [-100.0, -100.0, 1e-4, 1e-4],
                            [100.0, 100.0, 10.0, 1000.0], 1e-12, 1e-10, 500)
    np.testing.assert_allclose(x[0], -5.992406750800, rtol=1e-6)
    np.testing.assert_allclose(x[1], 0.074133661253, rtol=1e-5)
    np.testing.assert_allclose(x[2], 1e-4, rtol=1e-12)
    np.testing.assert_allclose(x[3], 1000.0, rtol=1e-12)
    np.testing.assert_allclose(f, 0.516012646087, rtol=1e-8)


def test_search_stops_at_the_boundary_rather_than_past_it():
    """src/numeric/lbfgsb.rs:1267-1296: badly scaled and separable, every
    optimum on the upper bound; with the line search capped at the feasible
    step the generalised Cauchy point reaches the corner (1, 1, 1, 1) to
    1e-12 in well under 100 evaluations."""
    lo, hi = np.zeros(4), np.ones(4)
    calls = [0]

    def counted(x):
        calls[0] += 1
        return decaying(x)
    x, f, ok = oracle.lbfgsb(counted, [0.0] * 4, lo, hi, 0.0, 1e-10, 200)
    assert ok
    np.testing.assert_allclose(x, np.ones(4), rtol=1e-12)
    assert calls[0] < 100


@pytest.mark.parametrize("seed", range(6))
def test_boxed_quadratics_match_scipy(seed):
    """A random SPD quadratic in a box that cuts off its free minimum: the
    constrained optimum is unique, so both methods must land on it."""
    rng = np.random.default_rng(seed)
    n = 5
    a = rng.normal(size=(n, n))
    h = a @ a.T + 0.5 * np.eye(n)
    c = rng.normal(scale=3.0, size=n)

    def fg(x):
        return 0.5 * x @ h @ x - c @ x, h @ x - c

    lo, hi = -np.ones(n), np.ones(n)
    x0 = rng.uniform(-1, 1, size=n)
    x, f, ok = oracle.lbfgsb(fg, x0, lo, hi, 1e-15, 1e-12, 1000)
    ref = minimize(fg, x0, jac=True, method="L-BFGS-B", bounds=list(zip(lo, hi)),
                   options=dict(ftol=1e-15, gtol=1e-12, maxiter=1000))
    assert ok
    np.testing.assert_allclose(x, ref.x, atol=1e-6)
    assert f <= ref.fun + 1e-10 * abs(ref.fun)
