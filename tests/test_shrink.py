"""Empirical-Bayes shrinkage of the cell-level overdispersions
(nh_shrink_sc_dispersion / nh_squeeze_var, src/sc/shrink.rs:90-226 over
limma's squeezeVar in src/limma/squeeze_var.rs) against R: limma's outputs
that the reference's tests hold (tests/golden/limma_squeeze_var.json, from
src/limma/squeeze_var.rs's test module) and limma's squeezeVar on the R
NEBULA fit of the sc dataset (tests/golden/sc_shrink.csv,
tests/e2e_nebula.rs:875-930); the reference's unit tests of shrink.rs
(:262-357); the distribution functions against 40-digit mpmath values and
scipy; lowess against a matrix least-squares restatement.  CPU only."""
import csv
import json
import os

import mpmath as mp
import numpy as np
import pytest
from scipy import special, stats

import nebula_hip as nh

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
R = json.load(open(os.path.join(GOLDEN, "limma_squeeze_var.json")))


def fixture_a():
    i = np.arange(1, 25)
    return (((i * 37) % 97) + 1) / 64.0, ((i * 53) % 101) / 8.0


def fixture_b():
    i = np.arange(1, 41)
    x = (((i * 37) % 251) + 1) / 64.0
    x[6], x[22], x[30] = 40.0, 60.0, 25.0
    return x, ((i * 29) % 97) / 8.0


def fixture_d():
    j = np.arange(1, 61)
    return 1.0 + (((j * 17) % 7) - 3.0) / 4096.0


def table(name):
    with open(os.path.join(GOLDEN, name)) as fh:
        rows = list(csv.reader(fh))
    return {h: np.array([float(v) for v in col]) for h, col in zip(rows[0], zip(*rows[1:]))}


def close(got, want, rtol):
    got, want = np.asarray(got, float), np.asarray(want, float)
    assert got.shape == want.shape
    inf = np.isinf(want)
    assert np.array_equal(got[inf], want[inf])
    np.testing.assert_allclose(got[~inf], want[~inf], rtol=rtol, atol=0)


# --------------------------------------------------------------- squeezeVar

def test_plain_fit_matches_limma():
    """fitFDist and squeezeVar, no covariate (squeeze_var.rs:2329-2395)."""
    x, _ = fixture_a()
    post, prior, dfp = nh.squeeze_var(x, 4.0)
    assert prior.size == 1 and dfp.size == 1
    close(prior, [R["A_FIT_SCALE"]], 1e-12)
    close(dfp, [R["A_FIT_DF2"]], 1e-12)
    close(post, R["A_VAR_POST"], 1e-12)


def test_trended_fit_matches_limma():
    """fitFDist with a natural-spline trend (squeeze_var.rs:2450-2519)."""
    x, cov = fixture_a()
    post, prior, dfp = nh.squeeze_var(x, 4.0, cov)
    assert prior.size == 24 and dfp.size == 1
    close(dfp, [R["A_TREND_DF_PRIOR"]], 1e-12)
    close(prior, R["A_TREND_VAR_PRIOR"], 1e-12)
    close(post, R["A_TREND_VAR_POST"], 1e-12)


def test_robust_fit_matches_limma():
    """fitFDistRobustly untrended: the root of the Winsorised variance, the
    outliers' discounted prior df (squeeze_var.rs:2555-2660)."""
    x, _ = fixture_b()
    post, prior, dfp = nh.squeeze_var(x, 4.0, robust=True)
    assert prior.size == 40 and dfp.size == 40
    close(prior, np.full(40, R["B_ROBUST_SCALE"]), 1e-12)
    close(dfp[[6, 22, 30]], R["B_ROBUST_SHRUNK_OUTLIERS"], 1e-12)
    close(np.delete(dfp, [6, 22, 30]), np.full(37, R["B_ROBUST_DF2"]), 5e-12)
    close(post, R["B_ROBUST_VAR_POST"], 1e-12)
    # closer to limma's converged root than stock limma's uniroot(tol = 1e-8)
    assert abs(dfp[0] - R["B_ROBUST_DF2"]) < abs(R["B_ROBUST_DF2_STOCK"] - R["B_ROBUST_DF2"])


def test_robust_trended_fit_takes_the_infinite_branch_as_limma():
    """Lowess trend, Winsorised spread below its df2 = Inf value: infinite
    prior df except the three outliers (squeeze_var.rs:2748-2834)."""
    x, cov = fixture_b()
    post, prior, dfp = nh.squeeze_var(x, 4.0, cov, robust=True)
    close(prior, R["B_ROBUST_TREND_VAR_PRIOR"], 1e-12)
    close(dfp[[6, 22, 30]], R["B_ROBUST_TREND_SHRUNK"], 1e-12)
    rest = np.delete(np.arange(40), [6, 22, 30])
    assert np.all(np.isinf(dfp[rest]))
    assert np.array_equal(post[rest], prior[rest])


def test_plain_fit_on_outliers_matches_limma():
    """fitFDist on fixture B: what limma's robust fit returns when nothing is
    Winsorised (squeeze_var.rs:2864-2890)."""
    x, _ = fixture_b()
    _, prior, dfp = nh.squeeze_var(x, 4.0)
    close(prior, [R["B_TINY_WINSOR_SCALE"]], 1e-12)
    close(dfp, [R["B_TINY_WINSOR_DF2"]], 1e-12)


def test_flat_variances_give_an_infinite_robust_prior():
    """Spread below any F's: df2 = Inf for every gene, posterior = prior
    (squeeze_var.rs:3083-3110)."""
    post, prior, dfp = nh.squeeze_var(fixture_d(), 6.0, robust=True)
    close(prior, np.full(60, R["D_ROBUST_SCALE"]), 1e-12)
    assert np.all(np.isinf(dfp))
    assert np.array_equal(post, prior)


def test_constant_variances_give_an_infinite_prior():
    post, prior, dfp = nh.squeeze_var(np.full(20, 0.5), 4.0)
    assert np.isinf(dfp[0]) and prior[0] == 0.5 and np.all(post == 0.5)


def test_fewer_than_three_genes_come_back_unmoderated():
    post, prior, dfp = nh.squeeze_var([0.5, 2.0], 4.0)
    assert list(post) == [0.5, 2.0] and list(prior) == [0.5, 2.0] and list(dfp) == [0.0]


def test_posterior_is_the_precision_weighted_mean():
    """var_post = (df s2 + d0 s0^2) / (df + d0), from the returned prior."""
    x, cov = fixture_a()
    for c in (None, cov):
        post, prior, dfp = nh.squeeze_var(x, 4.0, c)
        np.testing.assert_allclose(post, (4.0 * x + dfp * prior) / (4.0 + dfp), rtol=1e-14)


def test_squeeze_var_rejections():
    with pytest.raises(nh.InvalidArgument):
        nh.squeeze_var([], 4.0)
    with pytest.raises(nh.InvalidArgument):
        nh.squeeze_var([1.0, -1.0, 2.0], 4.0)
    with pytest.raises(nh.InvalidArgument):
        nh.squeeze_var([1.0, np.inf, 2.0], 4.0)
    with pytest.raises(nh.InvalidArgument):
        nh.squeeze_var([1.0, 2.0, 3.0], 0.0)
    with pytest.raises(nh.InvalidArgument):
        nh.squeeze_var([1.0, 2.0, 3.0], 4.0, [0.0, np.nan, 1.0])
    x, cov = fixture_b()
    cov = cov.copy()
    cov[3] = np.inf
    with pytest.raises(nh.InvalidArgument):  # the robust fit needs a finite covariate
        nh.squeeze_var(x, 4.0, cov, robust=True)
    # the spline trend moves it one unit past the finite range instead
    finite = cov.copy()
    finite[3] = np.max(np.delete(cov, 3)) + 1.0
    a = nh.squeeze_var(x, 4.0, cov)
    b = nh.squeeze_var(x, 4.0, finite)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


# ------------------------------------------------------ shrink_sc_dispersion

def test_residual_df():
    """src/sc/shrink.rs:269-273, and the sc dataset's 988 (1,005 cells, 3
    coefficients, 15 subjects; tests/e2e_nebula.rs:884-890)."""
    assert nh.sc_residual_df(100, 2, 6) == 93.0
    assert nh.sc_residual_df(10, 2, 20) == 1.0
    assert nh.sc_residual_df(1005, 3, 15) == 988.0


def test_matches_limma_on_the_r_nebula_fit():
    """limma's squeezeVar (trended on logFC_int, robust) on R's cell
    overdispersions of the sc dataset: phi to 1e-12, df_prior to 1e-11, one
    df_prior per usable gene (tests/e2e_nebula.rs:875-983)."""
    neb = table("sc_nebula.csv")
    want = table("sc_shrink.csv")
    scal = {(r[0], r[1]): float(r[2]) for r in csv.reader(open(os.path.join(GOLDEN,
                                                                            "sc_scalars.csv")))
            if r[0] == "sc_shrink"}
    df = nh.sc_residual_df(1005, 3, 15)
    assert df == scal[("sc_shrink", "df_residual")]
    out = nh.shrink_sc_dispersion(neb["Cell"], neb["convergence"].astype(np.int32),
                                  neb["logFC_int"], df)
    pos = {int(g): i for i, g in enumerate(neb["gene_id"])}
    idx = np.array([pos[int(g)] for g in want["gene_id"]])
    assert idx.size == int(scal[("sc_shrink", "n_usable")]) < neb["Cell"].size
    assert out.df_prior.size == idx.size and out.df_residual == df
    close(np.maximum(out.phi_raw[idx], 1e-8), want["phi_raw"], 1e-4)
    close(out.phi_post[idx], want["var_post"], 1e-12)
    close(out.phi_prior[idx], want["var_prior"], 1e-12)
    close(out.df_prior, want["df_prior"], 1e-11)
    # genes outside the fit take the median prior
    rest = np.setdiff1d(np.arange(neb["Cell"].size), idx)
    med = np.median(out.phi_prior[idx])
    assert np.all(out.phi_post[rest] == med) and np.all(out.phi_prior[rest] == med)


def _ref_fixture():
    return (np.array([0.5, 0.4, 0.6, 0.55, 0.45, 0.52, 0.48, 5.0, 0.51, 0.49]),
            np.ones(10, dtype=np.int32))


def test_shrinkage_pulls_the_outlier_towards_the_others():
    d, c = _ref_fixture()
    out = nh.shrink_sc_dispersion(d, c, None, 93.0)
    assert out.phi_raw[7] < out.phi_post[7] < out.phi_prior[7]


def test_non_converged_genes_get_the_prior():
    d, c = _ref_fixture()
    c[3] = -20
    out = nh.shrink_sc_dispersion(d, c, None, 93.0)
    assert out.phi_post[3] == pytest.approx(out.phi_prior[3], rel=1e-15)


def test_zero_dispersion_gives_infinite_phi_and_is_excluded():
    d, c = _ref_fixture()
    d[0] = 0.0
    out = nh.shrink_sc_dispersion(d, c, None, 93.0)
    assert np.isinf(out.phi_raw[0]) and np.isfinite(out.phi_post[0])
    assert out.phi_post[0] == pytest.approx(out.phi_prior[0], rel=1e-15)


def test_too_few_converged_genes_returns_unshrunk():
    d, c = _ref_fixture()
    c[2:] = -20
    out = nh.shrink_sc_dispersion(d, c, None, 93.0)
    assert np.array_equal(out.dispersion_shrunk, d) and list(out.df_prior) == [0.0]
    assert np.all(out.phi_prior == np.median(1.0 / d))


def test_a_covariate_gives_a_trended_prior():
    d, c = _ref_fixture()
    out = nh.shrink_sc_dispersion(d, c, np.arange(10.0), 93.0)
    assert np.any(np.abs(out.phi_prior - out.phi_prior[0]) > 1e-12)


def test_shrunk_dispersion_is_the_reciprocal_of_the_posterior():
    d, c = _ref_fixture()
    out = nh.shrink_sc_dispersion(d, c, None, 93.0)
    np.testing.assert_allclose(out.dispersion_shrunk, 1.0 / out.phi_post, rtol=1e-14)


def test_each_rung_is_squeeze_var_on_the_usable_phi():
    """The ladder's rungs: robust trended, robust, plain, each on phi = 1 /
    dispersion; a covariate the robust fit refuses (infinite) drops to the
    untrended robust fit, as in shrink.rs:170-196."""
    x, cov = fixture_b()
    conv = np.ones(40, dtype=np.int32)
    for c, robust in ((cov, True), (None, True), (None, False), (cov, False)):
        out = nh.shrink_sc_dispersion(1.0 / x, conv, c, 4.0, nh.ShrinkParams(robust=robust))
        post, prior, dfp = nh.squeeze_var(1.0 / (1.0 / x), 4.0, c, robust=robust)
        np.testing.assert_allclose(out.phi_post, post, rtol=1e-15)
        np.testing.assert_allclose(out.df_prior, dfp, rtol=1e-15)
    bad = cov.copy()
    bad[0] = np.inf
    out = nh.shrink_sc_dispersion(1.0 / x, conv, bad, 4.0)
    post, _, dfp = nh.squeeze_var(1.0 / (1.0 / x), 4.0, None, robust=True)
    np.testing.assert_allclose(out.phi_post, post, rtol=1e-15)
    close(out.phi_prior[[0]], [R["B_ROBUST_SCALE"]], 1e-12)


def test_shrink_rejections():
    d, c = _ref_fixture()
    with pytest.raises(nh.InvalidArgument):
        nh.shrink_sc_dispersion([], [], None, 93.0)
    with pytest.raises(nh.LengthMismatch):
        nh.shrink_sc_dispersion(d, [1, 1], None, 93.0)
    with pytest.raises(nh.LengthMismatch):
        nh.shrink_sc_dispersion(d, c, [1.0], 93.0)
    for bad in (0.0, np.nan, np.inf, -1.0):
        with pytest.raises(nh.InvalidArgument):
            nh.shrink_sc_dispersion(d, c, None, bad)


# ------------------------------------------------------- special functions

def _mp_f(m, n):
    """F(m, n) density, upper tail and quantile in 40-digit arithmetic."""
    mp.mp.dps = 40
    m_, n_ = mp.mpf(m), mp.mpf(n)

    def pdf(x):
        x = mp.mpf(x)
        return mp.exp(m_ / 2 * mp.log(m_ * x) + n_ / 2 * mp.log(n_) -
                      (m_ + n_) / 2 * mp.log(m_ * x + n_) - mp.log(x) - mp.log(mp.beta(m_ / 2,
                                                                                         n_ / 2)))

    def sf(x):
        x = mp.mpf(x)
        return mp.betainc(n_ / 2, m_ / 2, 0, n_ / (m_ * x + n_), regularized=True)

    def ppf(p, guess):
        cdf = lambda x: mp.betainc(m_ / 2, n_ / 2, 0, m_ * x / (m_ * x + n_), regularized=True)
        return mp.findroot(lambda x: cdf(x) - mp.mpf(p), (mp.mpf(guess) * (1 - mp.mpf(1e-6)),
                                                          mp.mpf(guess) * (1 + mp.mpf(1e-6))),
                           solver="anderson")
    return pdf, sf, ppf


@pytest.mark.parametrize("m,n", [(4.0, 8.77), (988.0, 7.5), (988.0, 2500.0), (0.7, 3.0),
                                 (6.0, 40.8), (2.0, 1.0)])
def test_f_distribution_against_high_precision(m, n):
    """Density, upper tail and quantile against 40-digit values (scipy's
    quantile is itself off by 1e-11 at p = 1e-6 for F(4, 8.77))."""
    pdf, sf, ppf = _mp_f(m, n)
    x = np.concatenate([np.geomspace(1e-4, 1e3, 15), [1.0]])
    np.testing.assert_allclose(nh.shrink_selftest(0, x, a=m, b=n),
                               [float(pdf(v)) for v in x], rtol=5e-13)
    np.testing.assert_allclose(nh.shrink_selftest(1, x, a=m, b=n),
                               [float(sf(v)) for v in x], rtol=5e-13)
    p = np.array([0.05, 0.1, 0.5, 0.9, 0.95, 1e-6, 1 - 1e-6])
    q = nh.shrink_selftest(2, p, a=m, b=n)
    want = [float(ppf(pi, stats.f.ppf(pi, m, n))) for pi in p]
    np.testing.assert_allclose(q, want, rtol=1e-13)


@pytest.mark.parametrize("m", [4.0, 6.0, 988.0, 0.5])
def test_infinite_denominator_is_the_chi_square_limit(m):
    x = np.geomspace(1e-3, 10.0, 31)
    np.testing.assert_allclose(nh.shrink_selftest(0, x, a=m, b=np.inf),
                               m * stats.chi2.pdf(m * x, m), rtol=1e-12)
    np.testing.assert_allclose(nh.shrink_selftest(1, x, a=m, b=np.inf),
                               stats.chi2.sf(m * x, m), rtol=1e-11)
    p = np.array([0.05, 0.5, 0.9])
    np.testing.assert_allclose(nh.shrink_selftest(2, p, a=m, b=np.inf),
                               stats.chi2.ppf(p, m) / m, rtol=1e-11)
    np.testing.assert_allclose(nh.shrink_selftest(3, m * x, a=m), stats.chi2.sf(m * x, m),
                               rtol=1e-11)


def test_polygamma_against_high_precision():
    """log x - digamma x (whose difference cancels in double precision for
    large x), trigamma against 40-digit values; inverse trigamma round trip."""
    mp.mp.dps = 40
    x = np.concatenate([np.geomspace(1e-3, 1e6, 40), [0.5, 1.0, 2.0, 4.37, 19.9, 20.0, 494.0]])
    want = [float(mp.log(mp.mpf(v)) - mp.digamma(mp.mpf(v))) for v in x]
    np.testing.assert_allclose(nh.shrink_selftest(5, x), want, rtol=1e-14)
    want = [float(mp.polygamma(1, mp.mpf(v))) for v in x]
    np.testing.assert_allclose(nh.shrink_selftest(6, x), want, rtol=1e-14)
    np.testing.assert_allclose(nh.shrink_selftest(6, x), special.polygamma(1, x), rtol=1e-13)
    inner = x[(x > 1e-2) & (x < 1e5)]
    np.testing.assert_allclose(nh.shrink_selftest(7, special.polygamma(1, inner)), inner,
                               rtol=1e-10)


# ------------------------------------------------------------------ lowess

def _lowess_by_matrix(x, y, f, iters):
    """stats::lowess restated as explicit weighted least squares at every
    point (points spaced wider than 1% of the range, so none is skipped):
    window of the ns = floor(f n) nearest points, tricube weights on the
    distance over the window's radius (1 inside 0.001 radius, 0 beyond 0.999),
    a local line unless the weighted spread is under 0.001 of the range;
    bisquare robustness weights on 6 x the median absolute residual."""
    n = x.size
    ns = max(2, min(n, int(f * n + 1e-7)))
    rw = np.ones(n)
    fit = np.zeros(n)
    for it in range(iters + 1):
        for i in range(n):
            d = np.abs(x - x[i])
            h = np.sort(d)[ns - 1]
            w = np.where(d <= 0.001 * h, 1.0, np.where(d <= 0.999 * h,
                                                       (1 - (d / h) ** 3) ** 3, 0.0))
            w = w * (rw if it > 0 else 1.0)
            w = w / w.sum()
            xm = np.sum(w * x)
            c = np.sum(w * (x - xm) ** 2)
            if np.sqrt(c) > 0.001 * (x.max() - x.min()):
                beta = np.sum(w * (x - xm) * y) / c
                fit[i] = np.sum(w * y) + beta * (x[i] - xm)
            else:
                fit[i] = np.sum(w * y)
        if it == iters:
            break
        r = np.abs(y - fit)
        cmad = 6.0 * np.median(r)
        u = r / cmad
        rw = np.where(r <= 0.001 * cmad, 1.0, np.where(r <= 0.999 * cmad, (1 - u ** 2) ** 2, 0.0))
    return fit


def test_lowess_against_a_matrix_restatement():
    i = np.arange(1, 51)
    x = i / 64.0
    y = 3 * x - 2 * x * x + ((i * i * 37) % 101) / 500.0 - 0.1
    y[[7, 30]] += [1.5, -2.0]
    perm = np.random.default_rng(3).permutation(50)
    for f, iters in ((0.4, 3), (0.3, 0), (0.8, 2)):
        got = nh.shrink_selftest(4, x[perm], y[perm], a=f, b=iters)
        want = _lowess_by_matrix(x, y, f, iters)[perm]
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)


def test_lowess_reproduces_a_line():
    x = np.sort(np.random.default_rng(5).uniform(0, 10, 300))
    y = 2.5 - 0.75 * x
    np.testing.assert_allclose(nh.shrink_selftest(4, x, y, a=0.4, b=3), y, rtol=1e-12,
                               atol=1e-12)
