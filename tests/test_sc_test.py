"""The Wald test on a fit (nh_glm_sc_test, glm_sc_test in src/sc/test.rs:98)
against nebula's numbers quoted by the reference's tests, the oracle's numpy
restatement, scipy and the definition of Benjamini-Hochberg.  CPU only."""
import numpy as np
import pytest

import nebula_hip as nh
import sc_test as oracle_test

# src/sc/test.rs:214-229: two genes, two coefficients, packed V11, V12, V22
COEF = np.array([[-3.625170366404492, 0.751566378617826],
                 [-5.04618066368606, 1.161381538814725]])
COV = np.array([[0.014446194037806379, -0.001444153266646014, 0.057784776151225514],
                [0.02024037603384674, -0.00954150218919818, 0.08096150413538698]])


@pytest.mark.parametrize("impl", ["library", "oracle"])
def test_single_coefficient_matches_nebula(impl):
    """se and p of the group coefficient as nebula reports them
    (src/sc/test.rs:237-244, 276-290)."""
    f = nh.glm_sc_test if impl == "library" else oracle_test.glm_sc_test
    out = f(COEF, COV, coef=1)
    np.testing.assert_allclose(out["se"], [0.24038464208685528, 0.28453735103741123], rtol=1e-12)
    np.testing.assert_allclose(out["log_fc"][0], 0.751566378617826, rtol=1e-12)
    np.testing.assert_allclose(out["p_value"], [0.0017689100639910616, 0.0000447173328216282],
                               rtol=1e-9)


def test_contrast_uses_the_off_diagonal_term():
    out = nh.glm_sc_test(COEF, COV, contrast=[1.0, 1.0])
    want = np.sqrt(0.014446194037806379 + 2 * -0.001444153266646014 + 0.057784776151225514)
    np.testing.assert_allclose(out["se"][0], want, rtol=1e-12)
    assert np.sqrt(0.014446194037806379 + 0.057784776151225514) > out["se"][0] * (1 + 1e-3)


def test_unit_contrast_agrees_with_the_coefficient():
    a = nh.glm_sc_test(COEF, COV, coef=1)
    b = nh.glm_sc_test(COEF, COV, contrast=[0.0, 1.0])
    for k in a:
        np.testing.assert_allclose(a[k], b[k], rtol=1e-15)


def test_degenerate_covariance_gives_a_p_value_of_one():
    out = nh.glm_sc_test(np.array([[1.0, 2.0]]), np.zeros((1, 3)), coef=1)
    assert out["se"][0] == 0.0 and out["p_value"][0] == 1.0 and np.isfinite(out["z"][0])


def test_rejections():
    with pytest.raises(nh.InvalidArgument):
        nh.glm_sc_test(COEF, COV, contrast=[0.0, 0.0])
    with pytest.raises(nh.CoefOutOfRange):
        nh.glm_sc_test(COEF, COV, coef=5)
    with pytest.raises(nh.LengthMismatch):
        nh.glm_sc_test(COEF, COV[:, :2], coef=0)
    with pytest.raises(nh.MustBePositive):
        nh.glm_sc_test(np.zeros((0, 2)), np.zeros((0, 3)), coef=0)


def _bh_definition(p):
    """fdr_(i) = min over j >= i of min(1, n p_(j) / j), ranks by ascending p."""
    p = np.asarray(p)
    n = len(p)
    order = np.argsort(p, kind="stable")
    s = p[order]
    out = np.empty(n)
    for i in range(n):
        out[order[i]] = min(1.0, min(n * s[j] / (j + 1) for j in range(i, n)))
    return out


def test_bh_matches_its_definition_and_scipy_chi2():
    rng = np.random.default_rng(5)
    p = np.concatenate([rng.uniform(size=60), [0.01, 0.01, 0.5, 0.5, 1.0, 0.0]])
    np.testing.assert_allclose(oracle_test.p_adjust_bh(p), _bh_definition(p), rtol=1e-15)


@pytest.mark.parametrize("n_coef", [1, 2, 4, 8])
def test_random_fits_match_the_oracle(n_coef):
    """Random fits (SPD covariances, a NaN gene, tied p-values): library
    against the oracle to 1e-13, for a coefficient and a contrast."""
    rng = np.random.default_rng(100 + n_coef)
    g = 300
    coef = rng.normal(scale=0.5, size=(g, n_coef))
    packed = n_coef * (n_coef + 1) // 2
    cov = np.zeros((g, packed))
    for i in range(g):
        a = rng.normal(size=(n_coef, n_coef)) * 0.1
        v = a @ a.T + 1e-3 * np.eye(n_coef)
        for jj in range(n_coef):
            for ii in range(jj + 1):
                cov[i, jj * (jj + 1) // 2 + ii] = v[ii, jj]
    coef[7] = np.nan                  # a failed gene
    coef[11], cov[11] = coef[10], cov[10]  # tied p-values
    for kw in ({"coef": n_coef - 1}, {"contrast": rng.normal(size=n_coef)}):
        a = nh.glm_sc_test(coef, cov, **kw)
        b = oracle_test.glm_sc_test(coef, cov, **kw)
        for k in ("log_fc", "se", "z", "p_value", "fdr"):
            np.testing.assert_allclose(a[k], b[k], rtol=1e-13, atol=1e-300, equal_nan=True)
