"""The FP64 sweep's table-driven exp and log, its Newton-refined
reciprocal and the device objective's log Gamma, run on the device through
the C ABI, against numpy's libm and scipy: exp
within 2 ulp, reciprocal within 1 ulp, log within 2 ulp of its larger part
(relative accuracy kept within 1/512 of t = 1, where log 1 = 0 exactly),
overflow, underflow, inf and NaN as libm gives them.  Marked gpu."""
import numpy as np
import pytest

import nebula_hip as nh

pytestmark = pytest.mark.gpu


def _ulp_err(got, ref):
    return np.abs(got - ref) / np.spacing(np.abs(ref))


def test_exp_matches_libm():
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.uniform(-708.0, 709.0, 400_000), rng.uniform(-40.0, 20.0, 400_000),
                        np.arange(-256, 257) * np.log(2.0) / 64.0,  # table knots
                        [0.0, -0.0, 1e-300, -1e-300, 1.0, -1.0]])
    got = nh.selftest_f64_math(x, 0)
    ref = np.exp(x)
    assert _ulp_err(got, ref).max() <= 2.0
    # overflow, underflow, deep underflow, NaN
    edge = np.array([709.9, 1e4, 1e10, -746.0, -1e4, -1e10, np.nan])
    e = nh.selftest_f64_math(edge, 0)
    assert np.all(np.isinf(e[:3]) & (e[:3] > 0))
    assert np.all(e[3:6] == 0.0)
    assert np.isnan(e[6])
    # subnormal results to within a few units of the smallest subnormal
    xs = np.linspace(-744.0, -709.0, 1001)
    assert np.max(np.abs(nh.selftest_f64_math(xs, 0) - np.exp(xs)) /
                  np.maximum(np.spacing(np.exp(xs)), 5e-324)) <= 2.0


def test_log_matches_libm():
    rng = np.random.default_rng(12)
    t = np.concatenate([np.exp(rng.uniform(-700.0, 700.0, 400_000)),
                        rng.uniform(1e-4, 50.0, 400_000),
                        1.0 + rng.uniform(-1e-3, 1e-3, 100_000)])
    # mantissas at and beside the interval edges 1 + (j +- 1/2) / 128
    j = np.arange(0, 129)
    edges = 1.0 + (j - 0.5) / 128.0
    t = np.concatenate([t, edges, np.nextafter(edges, 0), np.nextafter(edges, 3),
                        edges * 2.0 ** 40, edges * 2.0 ** -40, [1.0, 2.0, 0.5, 1e-300, 1e300]])
    got = nh.selftest_f64_math(t, 1)
    ref = np.log(t)
    big = np.maximum(np.abs(ref), np.abs(np.round(np.log2(t))) * np.log(2.0))
    err = np.abs(got - ref) / np.spacing(np.maximum(big, 0.5))
    assert err.max() <= 2.0, (err.max(), t[np.argmax(err)])
    # within 1/512 of 1 the table point is c = 1 (log c = 0): relative accuracy
    near = np.abs(t - 1.0) < 1.0 / 512
    assert _ulp_err(got[near & (t != 1.0)], ref[near & (t != 1.0)]).max() <= 4.0
    assert nh.selftest_f64_math(np.array([1.0]), 1)[0] == 0.0
    e = nh.selftest_f64_math(np.array([np.inf, np.nan]), 1)
    assert np.isinf(e[0]) and e[0] > 0 and np.isnan(e[1])


def test_reciprocal_matches_libm():
    rng = np.random.default_rng(13)
    t = np.concatenate([np.exp(rng.uniform(-690.0, 690.0, 400_000)),
                        rng.uniform(1e-4, 1e4, 400_000), [1.0, 3.0, 0.1, 1e-300, 1e300]])
    got = nh.selftest_f64_math(t, 2)
    assert _ulp_err(got, 1.0 / t).max() <= 1.0
    assert np.isnan(nh.selftest_f64_math(np.array([np.nan]), 2)[0])


def test_lgamma_matches_scipy():
    """The device objective's log Gamma (shift to x >= 10, Stirling's series)
    against scipy's gammaln over the arguments the objective uses: alpha and
    the cell parameter over their boxes, count values + cell."""
    from scipy.special import gammaln
    rng = np.random.default_rng(14)
    x = np.concatenate([np.exp(rng.uniform(np.log(1e-4), np.log(1e4), 300_000)),
                        3.0 + rng.integers(0, 500, 200_000) + rng.uniform(1e-4, 1e3, 200_000),
                        [1e-4, 0.5, 1.0, 2.0, 3.0, 9.999999, 10.0, 10.000001, 1e4]])
    got = nh.selftest_f64_math(x, 3)
    ref = gammaln(x)
    err = np.abs(got - ref) / np.maximum(1.0, np.abs(ref))
    assert err.max() < 2e-14, (err.max(), x[np.argmax(err)])
