"""The library's stage-one quasi-Newton (bound restarts around L-BFGS-B, the
search stage one runs where zero-count tables apply; nh_stage_one_search on
the host, the same code the driver runs on device sums) against the oracle's
independent restatement of L-BFGS-B (oracle/lbfgsb.hpp,
src/numeric/lbfgsb.rs): for the same objective the library's first descent
asks for the oracle's points bit for bit, and both land on the reference's
known answers.  CPU only."""
import numpy as np
import pytest

import nebula_hip as nh
import oracle
from test_oracle_lbfgsb import decaying, nebula_shaped, quadratic, rosenbrock

INF = np.inf


def boxed_quadratic(seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(4, 4))
    h = a @ a.T + 0.5 * np.eye(4)
    c = rng.normal(size=4) * 3.0

    def f(x):
        d = x - c
        return float(0.5 * d @ h @ d), h @ d
    return f


PROBLEMS = {
    "rosenbrock": (rosenbrock, [-1.2, 1.0], [-5.0, -5.0], [5.0, 5.0]),
    "quadratic_active_bound": (quadratic, [5.0, 0.0, 0.0], [2.0, -10.0, -10.0], [10.0] * 3),
    "decaying_corner": (decaying, [0.0] * 4, [0.0] * 4, [1.0] * 4),
    "nebula_shaped": (nebula_shaped, [0.0, 0.0, 1.0, 1.0], [-100.0, -100.0, 1e-4, 1e-4],
                      [100.0, 100.0, 10.0, 1000.0]),
    **{f"boxed_quadratic_{s}": (boxed_quadratic(s), [0.1] * 4, [-1.0] * 4, [1.0] * 4)
       for s in range(4)},
}


@pytest.mark.parametrize("name", list(PROBLEMS))
def test_first_descent_retraces_the_oracle(name):
    fg, x0, lo, hi = PROBLEMS[name]
    lib_trace, orc_trace = [], []

    def lf(x):
        lib_trace.append(np.array(x))
        return fg(x)

    def of(x):
        orc_trace.append(np.array(x))
        return fg(x)
    x, failed, evals = nh.stage_one_search(lf, x0, lo, hi)
    assert not failed and evals == len(lib_trace)
    x0c = np.clip(x0, lo, hi)
    f0, _ = fg(x0c)
    ox, of_, ok = oracle.lbfgsb(of, x0c, lo, hi, 1e-6 / max(abs(f0), 1.0), 1e-7, 300)
    assert ok
    # the library evaluates the start once before its descent starts there
    first = np.array(lib_trace[1:1 + len(orc_trace)])
    assert first.shape == np.array(orc_trace).shape
    assert np.array_equal(first, np.array(orc_trace))
    # a restart only ever keeps a lower objective
    assert fg(x)[0] <= of_ + 1e-12 * max(1.0, abs(of_))


def test_known_answers():
    """src/numeric/lbfgsb.rs's NEBULA-shaped problem through the library's
    stage-one search (scipy's optimum as the reference quotes it, at the
    looser stage-one tolerances) and the badly scaled corner."""
    x, _, _ = nh.stage_one_search(nebula_shaped, [0.0, 0.0, 1.0, 1.0],
                                  [-100.0, -100.0, 1e-4, 1e-4], [100.0, 100.0, 10.0, 1000.0])
    np.testing.assert_allclose(x[0], -5.992406750800, rtol=1e-4)
    assert x[2] == 1e-4 and x[3] == 1000.0
    x, _, _ = nh.stage_one_search(decaying, [0.0] * 4, [0.0] * 4, [1.0] * 4)
    np.testing.assert_allclose(x, np.ones(4), rtol=1e-6)


def test_arguments_are_checked():
    with pytest.raises(nh.InvalidBounds):
        nh.stage_one_search(quadratic, [0.0] * 3, [1.0] * 3, [0.0] * 3)
