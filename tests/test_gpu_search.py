"""The persistent kernel's search control on the device (nh_stage_two_search_device:
the same ctl_begin / ctl_tell / ctl_out the kernel calls, driven by built-in
objectives) against the same control on the host, and the reference's
state-machine outcomes (src/sc/nebula.rs:3516-3619) on the device: the
order-one retry after a rejected fit at a raised order, abandoning a
non-finite order-one pass, the roundoff-limited pass kept but failed, the
one-component search's restart off a bound, its polish, and a fit whose
Laplace correction falls below -1 (+inf by the profile objective)."""
import math

import numpy as np
import pytest

import nebula_hip as nh
from test_search_host import BOX1, BOX2, objective

pytestmark = pytest.mark.gpu

# name: (kind, problem, start, box, ord, inject (from, to, min_order, kind) or None)
CASES = {
    "both_interior": (2, 0, [0.5, 2.0], BOX2, 1, None),
    "both_retry_order_one": (2, 0, [0.5, 2.0], BOX2, 3, (5, 10 ** 6, 2, 0)),
    "both_abandon_nan": (2, 0, [0.5, 2.0], BOX2, 1, (7, 7, 1, 1)),
    "both_second_below_minus_one": (2, 0, [0.5, 2.0], BOX2, 3, (6, 6, 3, 2)),
    "both_roundoff": (2, 1, [0.0, 0.0], ([-2.0, -2.0], [2.0, 2.0]), 1, None),
    "subject_interior": (1, 3, [0.5, 2.0], BOX1, 1, None),
    "subject_bound_after_restart": (1, 2, [0.5, 2.0], BOX1, 1, None),
    "subject_retry_order_one": (1, 3, [0.5, 2.0], BOX1, 3, (4, 4, 3, 0)),
    "subject_nan_in_simplex": (1, 3, [0.5, 2.0], BOX1, 1, (3, 3, 1, 1)),
    "subject_nan_at_start": (1, 3, [0.5, 2.0], BOX1, 1, (1, 1, 1, 1)),
    "subject_second_below_minus_one": (1, 3, [0.5, 2.0], BOX1, 3, (2, 2, 3, 2)),
}


def host_run(kind, problem, start, box, ord, inject):
    base = objective(problem)
    trace = []

    def f(s, c, o):
        call = len(trace) + 1
        v = base(s, c)
        if inject is not None and inject[0] <= call <= inject[1] and o >= inject[2]:
            v = math.nan if inject[3] == 1 else math.inf
        trace.append((s, c, o, v))
        return v
    r = nh.stage_two_search(f, kind, start, box[0], box[1], ord)
    r["trace"] = np.array(trace)
    return r


@pytest.mark.parametrize("name", list(CASES))
def test_device_control_retraces_the_host(name):
    kind, problem, start, box, ord, inject = CASES[name]
    print(f"case {name}", flush=True)
    d = nh.stage_two_search_device(problem, kind, start, box[0], box[1], ord, inject)
    h = host_run(kind, problem, start, box, ord, inject)
    assert d["trace"].shape == h["trace"].shape
    assert np.array_equal(d["trace"][:, :3], h["trace"][:, :3])
    # values: the device evaluates the same expressions; +inf and NaN agree
    dv, hv = d["trace"][:, 3], h["trace"][:, 3]
    same = (dv == hv) | (np.isnan(dv) & np.isnan(hv))
    assert same.all()
    for key in ("found", "failed", "order", "evaluations"):
        assert d[key] == h[key], key
    if d["found"]:
        assert np.array_equal(d["x"], h["x"])


def test_device_outcomes_match_the_reference_cases():
    r = nh.stage_two_search_device(0, 2, [0.5, 2.0], *BOX2, ord=1)
    assert r["found"] and not r["failed"] and r["order"] == 1
    np.testing.assert_allclose(r["x"], [0.3, 4.0], rtol=1e-4)
    r = nh.stage_two_search_device(0, 2, [0.5, 2.0], *BOX2, ord=3, inject=(5, 10 ** 6, 2, 0))
    assert r["order"] == 1 and r["found"] and not r["failed"]
    assert r["trace"][0, 2] == 3 and r["trace"][-1, 2] == 1
    r = nh.stage_two_search_device(0, 2, [0.5, 2.0], *BOX2, ord=1, inject=(7, 7, 1, 1))
    assert not r["found"] and r["evaluations"] == 7
    r = nh.stage_two_search_device(1, 2, [0.0, 0.0], [-2.0, -2.0], [2.0, 2.0], ord=1)
    assert r["found"] and r["failed"] and abs(r["x"].sum() - 1.0) < 1e-6
    r = nh.stage_two_search_device(2, 1, [0.5, 2.0], *BOX1, ord=1)
    assert r["found"] and r["x"][0] == 1e-4 and r["x"][1] == 2.0
    # the Laplace correction below -1 at order 3: +inf, so the pass reruns at order 1
    r = nh.stage_two_search_device(0, 2, [0.5, 2.0], *BOX2, ord=3, inject=(6, 6, 3, 2))
    assert math.isinf(r["trace"][5, 3]) and r["order"] == 1 and r["found"]
    # one component: the deeper interior well is found after the restart
    r = nh.stage_two_search_device(4, 1, [0.5, 2.0], *BOX1, ord=1)
    assert r["found"] and abs(r["x"][0] - 0.05) < 1e-3
