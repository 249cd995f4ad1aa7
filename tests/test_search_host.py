"""The library's stage-two search control (SearchControl, the state machine
both engines run; nh_stage_two_search) on the host against the oracle's
independent restatement of StageTwoSearch (oracle/search.hpp and the BOBYQA
pass of oracle/nebula.cpp): for the same objective both ask for the same
points, in the same order, at the same Laplace order, bit for bit; and the
reference's own state-machine cases (src/sc/nebula.rs:3516-3619).  CPU only."""
import math

import numpy as np
import pytest

import nebula_hip as nh
import oracle


def objective(problem):
    """Built-in objectives of the device hook, in Python (no exp in 0-3, so
    host and device arithmetic agree)."""
    def f(s, c):
        if problem == 0:
            return (s - 0.3) ** 2 + 0.01 * (c - 4.0) ** 2
        if problem == 1:
            return (s + c - 1.0) ** 2
        if problem == 2:
            return s
        if problem == 3:
            return (s - 0.3) ** 2 + 2.0
        a, b = (s - 1e-4) / 2e-3, (s - 0.05) / 0.02
        return -0.5 * math.exp(-(a * a)) - 1.0 * math.exp(-(b * b))
    return f


def run_both(f, kind, start, lo, hi, ord):
    """Library trace and oracle trace of the same search: (subject, cell,
    order, value) per evaluation."""
    lib_trace, orc_trace = [], []

    def lf(s, c, o):
        v = f(s, c, o, len(lib_trace) + 1)
        lib_trace.append((s, c, o, v))
        return v
    r = nh.stage_two_search(lf, kind, start, lo, hi, ord)
    held = start[1]

    def of(x, o, call):
        s, c = (x[0], x[1]) if kind == 2 else (x[0], held)
        v = f(s, c, o, call)
        orc_trace.append((s, c, o, v))
        return v
    ox, ofailed, oorder, oevals = oracle.stage_two_search(
        of, start[:kind], lo[:kind], hi[:kind], ord)
    return r, np.array(lib_trace), (ox, ofailed, oorder, oevals), np.array(orc_trace)


BOX2 = ([1e-4, 1e-4], [10.0, 1000.0])
BOX1 = ([1e-4, 0.0], [10.0, 1.0])

CASES = {
    # name: (kind, problem, start, box, ord, injection(call, order) -> value or None)
    "both_interior": (2, 0, [0.5, 2.0], BOX2, 1, None),
    "both_retry_order_one": (2, 0, [0.5, 2.0], BOX2, 3,
                             lambda c, o: math.inf if o > 1 and c >= 5 else None),
    "both_abandon_nan": (2, 0, [0.5, 2.0], BOX2, 1, lambda c, o: math.nan if c == 7 else None),
    "both_roundoff": (2, 1, [0.0, 0.0], ([-2.0, -2.0], [2.0, 2.0]), 1, None),
    "subject_interior": (1, 3, [0.5, 2.0], BOX1, 1, None),
    "subject_bound_after_restart": (1, 2, [0.5, 2.0], BOX1, 1, None),
    "subject_two_wells": (1, 4, [0.5, 2.0], BOX1, 1, None),
    "subject_retry_order_one": (1, 3, [0.5, 2.0], BOX1, 3,
                                lambda c, o: math.inf if o == 3 and c == 4 else None),
    "subject_nan_in_simplex": (1, 3, [0.5, 2.0], BOX1, 1,
                               lambda c, o: math.nan if c == 3 else None),
    "subject_nan_at_start": (1, 3, [0.5, 2.0], BOX1, 1,
                             lambda c, o: math.nan if c == 1 else None),
    "subject_polish_inf": (1, 3, [0.5, 2.0], BOX1, 1,
                           lambda c, o: math.inf if c in (40, 41) else None),
    "subject_start_outside": (1, 3, [25.0, 2.0], BOX1, 1, None),
}


def make_f(problem, inject):
    base = objective(problem)

    def f(s, c, o, call):
        if inject is not None:
            v = inject(call, o)
            if v is not None:
                return v
        return base(s, c)
    return f


@pytest.mark.parametrize("name", list(CASES))
def test_library_search_retraces_the_oracle(name):
    kind, problem, start, box, ord, inject = CASES[name]
    r, lt, (ox, ofailed, oorder, oevals), ot = run_both(make_f(problem, inject), kind, start,
                                                        box[0], box[1], ord)
    assert lt.shape == ot.shape, (lt.shape, ot.shape)
    assert np.array_equal(lt[:, :3], ot[:, :3])
    assert np.array_equal(np.isnan(lt[:, 3]), np.isnan(ot[:, 3]))
    assert r["found"] == (ox is not None)
    if ox is not None:
        assert r["x"][0] == ox[0]
        if kind == 2:
            assert r["x"][1] == ox[1]
    assert r["failed"] == ofailed and r["order"] == oorder and r["evaluations"] == oevals


def test_reference_outcomes():
    """The outcomes src/sc/nebula.rs:3544-3619 asserts, on the library."""
    f = make_f(0, None)
    r = nh.stage_two_search(lambda s, c, o: f(s, c, o, 0), 2, [0.5, 2.0], *BOX2, ord=1)
    assert r["found"] and not r["failed"] and r["order"] == 1
    np.testing.assert_allclose(r["x"], [0.3, 4.0], rtol=1e-4)
    calls = [0]

    def retry(s, c, o):
        calls[0] += 1
        return math.inf if o > 1 and calls[0] >= 5 else f(s, c, o, 0)
    r = nh.stage_two_search(retry, 2, [0.5, 2.0], *BOX2, ord=3)
    assert r["order"] == 1 and r["found"] and not r["failed"]
    np.testing.assert_allclose(r["x"][0], 0.3, rtol=1e-4)
    calls[0] = 0

    def nan7(s, c, o):
        calls[0] += 1
        return math.nan if calls[0] == 7 else f(s, c, o, 0)
    assert not nh.stage_two_search(nan7, 2, [0.5, 2.0], *BOX2, ord=1)["found"]
    r = nh.stage_two_search(lambda s, c, o: (s + c - 1.0) ** 2, 2, [0.0, 0.0], [-2.0, -2.0],
                            [2.0, 2.0], ord=1)
    assert r["found"] and r["failed"] and abs(r["x"].sum() - 1.0) < 1e-6
    r = nh.stage_two_search(lambda s, c, o: s, 1, [0.5, 2.0], *BOX1, ord=1)
    assert r["found"] and not r["failed"] and r["x"][0] == 1e-4 and r["x"][1] == 2.0


def test_search_arguments_are_checked():
    with pytest.raises(nh.InvalidArgument):
        nh.stage_two_search(lambda *a: 0.0, 3, [0.5, 2.0], *BOX2)
    with pytest.raises(nh.InvalidBounds):
        nh.stage_two_search(lambda *a: 0.0, 1, [0.5, 2.0], [1.0, 0.0], [1.0, 1.0])
