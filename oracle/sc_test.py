"""numpy f64 restatement of edge-rs's Wald test on a NEBULA fit, for checking
the library's nh_glm_sc_test.

Test infrastructure: only tests/ import this module, as the checker.

glm_sc_test (src/sc/test.rs:98-201): a coefficient index becomes a unit
weight vector; effect = c'beta; variance = c'Vc over the non-zero weights with
the off-diagonal terms (packed upper triangle, (i, j) at j(j+1)/2 + i,
src/sc/test.rs:78-81); se = sqrt(variance) when positive, else 0; z =
effect / se when se > 0, else 0; p = chisq_sf(z^2, 1) (src/numeric/dist.rs:
737-746) when se > 0, else 1; fdr = p_adjust_bh (src/numeric/stats.rs:40-61).

Pinned by nebula's own numbers the reference's tests quote
(src/sc/test.rs:214-305: se and p of the group coefficient on a two-gene
fit), by scipy's chi2.sf, and by BH's definition fdr_(i) = min_{j >= i}
min(1, n p_(j) / j) evaluated by brute force (tests/test_sc_test.py).
"""
import numpy as np
from scipy.stats import chi2


def packed_get(v, i, j):
    lo, hi = min(i, j), max(i, j)
    return v[hi * (hi + 1) // 2 + lo]


def p_adjust_bh(p):
    """src/numeric/stats.rs:40-61: stable order by descending p (NaN first,
    as f64::total_cmp puts positive NaN above +inf), running minimum of
    n / rank * p capped at 1; a NaN candidate never lowers the minimum."""
    p = np.asarray(p, dtype=np.float64)
    n = len(p)
    key = np.where(np.isnan(p), np.inf, p)
    nan_first = np.isnan(p)
    # descending, NaN (above +inf) first, ties in input order
    order = sorted(range(n), key=lambda i: (0 if nan_first[i] else 1, -key[i] if not nan_first[i] else 0, i))
    out = np.zeros(n)
    running = np.inf
    for j, idx in enumerate(order):
        cand = n / (n - j) * p[idx]
        if not np.isnan(cand):
            running = min(running, cand)
        out[idx] = min(running, 1.0)
    return out


def glm_sc_test(coefficients, covariance, coef=None, contrast=None):
    b = np.asarray(coefficients, dtype=np.float64)
    v = np.asarray(covariance, dtype=np.float64)
    n_genes, n_coef = b.shape
    if contrast is None:
        w = np.zeros(n_coef)
        w[coef] = 1.0
    else:
        w = np.asarray(contrast, dtype=np.float64)
    out = {k: np.zeros(n_genes) for k in ("log_fc", "se", "z", "p_value")}
    for g in range(n_genes):
        effect = float(np.sum(b[g] * w))
        var = 0.0
        for i in range(n_coef):
            if w[i] == 0.0:
                continue
            for j in range(n_coef):
                if w[j] != 0.0:
                    var += w[i] * w[j] * packed_get(v[g], i, j)
        se = np.sqrt(var) if var > 0.0 else 0.0
        z = effect / se if se > 0.0 else 0.0
        p = float(chi2.sf(z * z, 1)) if se > 0.0 else 1.0
        out["log_fc"][g], out["se"][g], out["z"][g], out["p_value"][g] = effect, se, z, p
    out["fdr"] = p_adjust_bh(out["p_value"])
    return out
