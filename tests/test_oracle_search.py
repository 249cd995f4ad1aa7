"""The oracle's stage-two search state machine on synthetic objectives, as
the reference pins its own (src/sc/nebula.rs:3516-3619): an interior minimum,
the order-one retry after a rejected fit at a raised order, abandoning a
non-finite order-one pass, the roundoff-limited BOBYQA pass reported as
failed, and the one-component search keeping a genuine bound after its
restart.  The one-component simplex is also checked against scipy's
Nelder-Mead started from the same simplex, and the polish against the
closed-form minimum of a quadratic."""
import numpy as np
import pytest
from scipy.optimize import minimize

import oracle

BOX2 = ([1e-4, 1e-4], [10.0, 1000.0])
BOX1 = ([1e-4], [10.0])


def bowl(x, *_):
    return (x[0] - 0.3) ** 2 + (0.01 * (x[1] - 4.0) ** 2 if len(x) > 1 else 0.0)


def test_finds_an_interior_minimum():
    x, failed, order, _ = oracle.stage_two_search(bowl, [0.5, 2.0], *BOX2, ord=1)
    assert x is not None and not failed and order == 1
    np.testing.assert_allclose(x, [0.3, 4.0], rtol=1e-4)


def test_retries_at_order_one_after_a_rejected_fit():
    def f(x, order, calls):
        return np.inf if order > 1 and calls >= 5 else bowl(x)
    x, failed, order, _ = oracle.stage_two_search(f, [0.5, 2.0], *BOX2, ord=3)
    assert order == 1 and x is not None and not failed
    np.testing.assert_allclose(x[0], 0.3, rtol=1e-4)


def test_abandons_a_non_finite_order_one_pass():
    def f(x, order, calls):
        return np.nan if calls == 7 else bowl(x)
    x, _, _, _ = oracle.stage_two_search(f, [0.5, 2.0], *BOX2, ord=1)
    assert x is None


def test_roundoff_limited_search_is_reported_failed():
    x, failed, _, _ = oracle.stage_two_search(lambda x, *_: (x[0] + x[1] - 1.0) ** 2,
                                              [0.0, 0.0], [-2.0, -2.0], [2.0, 2.0], ord=1)
    assert x is not None and failed
    assert abs(x[0] + x[1] - 1.0) < 1e-6


def test_one_component_keeps_a_genuine_bound_after_the_restart():
    x, failed, _, _ = oracle.stage_two_search(lambda x, *_: x[0], [0.5], *BOX1, ord=1)
    assert x is not None and not failed
    assert x[0] == 1e-4


def test_one_component_interior_minimum_and_polish():
    """A quadratic's minimum: the polish stencils fit it exactly, so the
    search lands on 0.3 to rounding, not just to the simplex's 1e-7."""
    x, failed, order, evals = oracle.stage_two_search(lambda x, *_: (x[0] - 0.3) ** 2 + 2.0,
                                                      [0.5], *BOX1, ord=1)
    assert x is not None and not failed and order == 1
    assert abs(x[0] - 0.3) < 1e-12
    # start, simplex, check, three stencils of three
    assert evals > 1 + 2 + 1 + 9


def test_one_component_order_one_retry_and_abandon():
    """A raised order rejected once: the pass reruns at order one from the
    start.  At order one a rejected simplex point only marks the pass (its
    result is kept, nebula.rs:1756-1765), while a rejected start or check
    point abandons the search."""
    seen = []

    def f(x, order, calls):
        seen.append(order)
        return np.inf if order == 3 and calls == 4 else (x[0] - 0.3) ** 2
    x, _, order, _ = oracle.stage_two_search(f, [0.5], *BOX1, ord=3)
    assert order == 1 and x is not None and abs(x[0] - 0.3) < 1e-9
    assert seen[0] == 3 and seen[-1] == 1
    x, _, _, _ = oracle.stage_two_search(lambda x, o, c: np.nan if c == 3 else x[0] ** 2,
                                         [0.5], *BOX1, ord=1)
    assert x is not None and x[0] == 1e-4
    x, _, _, _ = oracle.stage_two_search(lambda x, o, c: np.nan if c == 1 else x[0] ** 2,
                                         [0.5], *BOX1, ord=1)
    assert x is None


def test_one_component_bound_restart_recovers_the_interior_basin():
    """An overshooting first simplex pinned on the lower bound is restarted at
    the geometric mean of the bound and the start and finds the better
    interior minimum (src/sc/nebula.rs:1760-1790)."""
    # a spurious well at the bound, a deeper minimum at 0.05
    def f(x, *_):
        s = x[0]
        return -0.5 * np.exp(-((s - 1e-4) / 2e-3) ** 2) - 1.0 * np.exp(-((s - 0.05) / 0.02) ** 2)
    x, _, _, _ = oracle.stage_two_search(f, [0.5], *BOX1, ord=1)
    assert x is not None
    assert f(x) <= -0.999


@pytest.mark.parametrize("x0", [0.7, 2.5])
def test_simplex_agrees_with_scipy_nelder_mead(x0):
    """Without bounds in play the simplex is scipy's: from the same initial
    simplex (x0, 1.05 x0) at the same tolerances both stop within the
    tolerance of each other on a smooth convex function."""
    f = lambda x: (np.log(x[0]) - np.log(1.3)) ** 2 + 0.1 * x[0]
    r = minimize(f, [x0], method="Nelder-Mead",
                 options=dict(xatol=1e-7, fatol=1e-7, initial_simplex=[[x0], [1.05 * x0]]))
    calls = []

    def g(x, order, c):
        calls.append(x[0])
        return f(x)
    x, _, _, _ = oracle.stage_two_search(g, [x0], [1e-4], [100.0], ord=1)
    # the simplex stage's own minimiser is in the calls before the check
    assert abs(x[0] - r.x[0]) < 1e-5
