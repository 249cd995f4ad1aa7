"""Pins the oracle's zero-count tables (src/sc/zeros.rs restated in
oracle/zeros.cpp): the table against its own compensated direct sums, the
direct sums against an independent numpy form and against finite differences
(each of the six sums is the s-derivative of the one before), the grouping
rules of ZeroCells::build, and the tabled penalised fit and marginal
likelihood against the dense sweep they replace (src/sc/pml.rs:2884-2960)."""
import numpy as np
import pytest

import oracle


def _offsets(n, seed):
    # log offsets spread like library sizes over two orders of magnitude
    return np.random.default_rng(seed).uniform(6.0, 10.6, n)


def _numpy_sums(log_o, s):
    """The six sums by their defining formulas, in plain numpy (float64)."""
    v = np.exp(s + log_o)
    w = 1.0 / (1.0 + v)
    u = v * w
    return np.array([
        np.sum(np.log1p(v)),
        np.sum(u),
        np.sum(u * w),
        np.sum(u * w * (w - u)),
        np.sum(u * w * (w * w - 4 * u * w + u * u)),
        np.sum(u * w * (w ** 3 - 11 * w * w * u + 11 * w * u * u - u ** 3)),
    ])


def _scaled_error(got, want):
    # C, D and E are judged against B, the scale they are stored at
    scale = np.abs(want).copy()
    scale[..., 3:] = np.abs(want[..., 2:3])
    return np.abs(got - want) / scale


def test_table_matches_direct_sums_over_the_whole_range():
    """zeros.rs:512-535: every sum within 1e-13 of the compensated direct sum
    from 5 below the table to 5 above it."""
    log_o = _offsets(500, 7)
    _, _, info = oracle.zero_group(log_o, [0.0], direct=False)
    s = np.arange(info[0] - 5.0, info[1] + 5.0, 0.0137)
    got, want, _ = oracle.zero_group(log_o, s)
    worst = _scaled_error(got, want).max(axis=0)
    assert np.all(worst < 1e-13), worst


def test_direct_sums_match_numpy_and_are_successive_derivatives():
    """The direct sums against their formulas summed by numpy, and each sum
    against a central difference of the one before it in s: a wrong sign or
    coefficient in C, D or E (or a dropped term) fails one of these."""
    log_o = _offsets(64, 3)
    s = np.array([-12.0, -9.3, -8.0, -7.1, -5.0, -2.0])
    _, direct, _ = oracle.zero_group(log_o, s)
    for i, si in enumerate(s):
        np.testing.assert_allclose(direct[i], _numpy_sums(log_o, si), rtol=1e-12)
    h = 1e-4
    _, up, _ = oracle.zero_group(log_o, s + h)
    _, dn, _ = oracle.zero_group(log_o, s - h)
    fd = (up - dn) / (2 * h)
    for f in range(5):
        np.testing.assert_allclose(fd[:, f], direct[:, f + 1], rtol=1e-6,
                                   atol=1e-7 * np.abs(direct[:, 2]).max())


def test_tails_follow_the_closed_forms():
    """Past the table every v is tiny (all sums are sum v) or huge (L grows as
    n s + sum ln O, A tends to n); infinities and NaN as zeros.rs:575-596."""
    log_o = _offsets(100, 13)
    got, _, info = oracle.zero_group(log_o, [info_s for info_s in (-80.0, 80.0)], direct=False)
    small = np.exp(-80.0) * np.sum(np.exp(log_o))
    np.testing.assert_allclose(got[0], small, rtol=1e-13)
    np.testing.assert_allclose(got[1, 0], 100 * 80.0 + np.sum(log_o), rtol=1e-13)
    np.testing.assert_allclose(got[1, 1], 100.0, rtol=1e-13)
    lo, _, _ = oracle.zero_group(log_o, [-np.inf], direct=False)
    assert np.all(lo == 0.0)
    hi, _, _ = oracle.zero_group(log_o, [np.inf], m=3, direct=False)
    assert hi[0, 0] == np.inf and hi[0, 1] == 100.0
    nan, _, _ = oracle.zero_group(log_o, [np.nan], direct=False)
    assert np.all(np.isnan(nan))


def test_table_is_continuous_across_its_ends():
    """zeros.rs:560-580: within 1e-12 of the direct sums on either side of
    s_lo and s_hi."""
    log_o = _offsets(300, 11)
    _, _, info = oracle.zero_group(log_o, [0.0], direct=False)
    pts = [e + d for e in info[:2] for d in (-1e-9, 0.0, 1e-9)]
    got, want, _ = oracle.zero_group(log_o, pts)
    assert _scaled_error(got, want).max() <= 1e-12


def test_groups_of_a_categorical_design():
    """zeros.rs:543-557: three design rows per subject, the rows with the odd
    third column too few to table."""
    n = 400
    log_o = _offsets(n, 3)
    design = np.array([[1.0, r % 3, 0.5 if r % 50 == 0 else 0.0] for r in range(n)])
    z = oracle.zero_cells(design, log_o, [0, 200, 400])
    assert z is not None
    assert list(z["subject_groups"]) == [0, 3, 6]
    assert list(z["loose"]) == [0, 50, 100, 150, 200, 250, 300, 350]
    assert int(z["group_n"].sum()) + len(z["loose"]) == n
    # groups ordered by first cell, each keyed by its first member's row
    assert list(z["group_row"]) == [1, 2, 3, 201, 202, 203]


def test_a_continuous_design_is_not_tabled():
    n = 400
    log_o = _offsets(n, 3)
    x = np.random.default_rng(9).uniform(-1, 1, n)
    design = np.column_stack([np.ones(n), x])
    assert oracle.zero_cells(design, log_o, [0, 200, 400]) is None


def _categorical_gene(seed=11, n_per=400, k=3):
    """Three subjects of 400 cells: intercept, a subject-level indicator and a
    cell-level indicator, offsets uniform on (-2.3, 2.3), about 12 % of the
    cells positive (the shape of pml.rs:2813-2844)."""
    rng = np.random.default_rng(seed)
    n = n_per * k
    design, offset, counts, cells = [], [], [], []
    totals = np.zeros(k)
    for c in range(n):
        s = c // n_per
        kind = float(rng.random() < 0.4)
        design.append([1.0, float(s % 2), kind])
        offset.append(rng.uniform(-2.3, 2.3))
        if rng.random() < 0.1 + 0.05 * kind:
            y = float(rng.integers(1, 6))
            counts.append(y)
            cells.append(c)
            totals[s] += y
    starts = [n_per * s for s in range(k + 1)]
    return (np.array(design), np.array(offset), np.array(counts), np.array(cells), totals,
            starts)


@pytest.mark.parametrize("subject,cell", [(0.3, 0.05), (0.8, 2.0), (0.05, 40.0)])
def test_tabled_fit_matches_the_dense_fit(subject, cell):
    """pml.rs:2926-2960: same iterations, beta and information to 1e-11,
    log-likelihood to 1e-13, log-determinant to 1e-12 (REML, Laplace order 3,
    which reads the tables' B..E)."""
    design, offset, counts, cells, totals, starts = _categorical_gene()
    start = [-1.5, 0.2, 0.1]
    kw = dict(reml=True, ord=3)
    dense = oracle.opt_pml(design, offset, counts, cells, starts, totals, start, subject, cell,
                           **kw)
    tabled = oracle.opt_pml(design, offset, counts, cells, starts, totals, start, subject, cell,
                            tabled=True, **kw)
    assert dense["iterations"] == tabled["iterations"]
    np.testing.assert_allclose(tabled["beta"], dense["beta"], rtol=1e-11)
    np.testing.assert_allclose(tabled["information"], dense["information"], rtol=1e-11)
    np.testing.assert_allclose(tabled["log_likelihood"], dense["log_likelihood"], rtol=1e-13)
    np.testing.assert_allclose(tabled["log_det"], dense["log_det"], rtol=1e-12)
    np.testing.assert_allclose(tabled["second_order"], dense["second_order"], rtol=1e-11)


def test_tabled_marginal_likelihood_matches_the_dense_one():
    """Stage one's value and gradient from the tables (ptmg.rs:1462) against
    the dense pass at several points."""
    design, offset, counts, cells, totals, starts = _categorical_gene(seed=5)
    for params in ([-1.0, 0.3, -0.2, 0.5, 2.0], [0.2, -0.1, 0.4, 1.5, 0.3],
                   [-3.0, 0.0, 0.0, 0.05, 40.0]):
        v0, g0, _ = oracle.ptmg(design, offset, counts, cells, totals, starts, params,
                                hessian=False)
        v1, g1, _ = oracle.ptmg(design, offset, counts, cells, totals, starts, params,
                                hessian=False, tabled=True)
        np.testing.assert_allclose(v1, v0, rtol=1e-12)
        np.testing.assert_allclose(g1, g0, rtol=1e-10, atol=1e-10 * np.abs(g0).max())


def test_tabled_and_dense_runs_agree_end_to_end():
    """nebula_sparse on a subject-level-only design (the shape of the
    200-gene x 5k-cell workload, scaled down): the tabled run and the run that
    sweeps every cell reach the same fits."""
    from nebula_hip import synth
    w = synth.config(1, n_genes=12, n_cells=2000, seed=4)
    prm = oracle.params(method="HL")
    args = (w.indptr, w.indices, w.data, w.n_cells, w.subject_id, w.design, w.n_coef, w.offset)
    fid = np.r_[0, np.cumsum(np.bincount(w.subject_id.astype(np.int64)))]
    z = oracle.zero_cells(w.design, np.log(w.offset), fid)
    assert z is not None and len(z["loose"]) == 0
    a = oracle.nebula_sparse(*args, prm=prm, tables=True, with_stage_two=True)
    b = oracle.nebula_sparse(*args, prm=prm, tables=False, with_stage_two=True)
    np.testing.assert_array_equal(a["gene_index"], b["gene_index"])
    np.testing.assert_array_equal(a["convergence"], b["convergence"])
    # the profile objective at each minimiser agrees to rounding; the
    # minimisers themselves only to the search's stopping resolution
    np.testing.assert_allclose(a["stage_two"][:, 3], b["stage_two"][:, 3], rtol=1e-10)
    rel = lambda k: np.max(np.abs(a[k] - b[k]) / np.maximum(np.abs(b[k]), 1e-3))
    assert rel("coefficients") < 1e-5
    assert rel("se") < 5e-5
    assert rel("cell_overdispersion") < 5e-5
    assert rel("subject_overdispersion") < 2e-4


def test_library_tables_match_the_oracle():
    """The library's host-built tables (the ones the device sweeps read): the
    same groups as the oracle's ZeroCells::build, and every sum of every group
    within 1e-13 of the oracle's compensated direct sums over the group's
    offsets, across the whole table and both tails."""
    import nebula_hip as nh
    n = 600
    rng = np.random.default_rng(21)
    log_o = rng.uniform(-1.5, 2.5, n)
    design = np.array([[1.0, r % 3, 1.0 if r % 97 == 0 else 0.0, (r // 300) * 0.5]
                       for r in range(n)])
    fid = [0, 300, 600]
    want = oracle.zero_cells(design, log_o, fid)
    got_grouping = nh.zero_tables_probe(design, log_o, fid, [0.0])
    np.testing.assert_array_equal(got_grouping["group_row"], want["group_row"])
    np.testing.assert_array_equal(got_grouping["subject_groups"], want["subject_groups"])
    members = {}
    for r in range(n):
        key = tuple(design[r])
        s = 0 if r < 300 else 1
        members.setdefault((s, key), []).append(r)
    s_pts = np.arange(-50.0, 50.0, 0.173)
    got = nh.zero_tables_probe(design, log_o, fid, s_pts, frm=0, m=6)["values"]
    for g, row in enumerate(want["group_row"]):
        s = 0 if row < 300 else 1
        cells = members[(s, tuple(design[row]))]
        _, direct, _ = oracle.zero_group(log_o[cells], s_pts)
        err = _scaled_error(got[g], direct)
        assert err.max() < 1e-13, (g, err.max(axis=0))
    # (2, 4): B..E, as the Laplace terms read them
    b_e = nh.zero_tables_probe(design, log_o, fid, s_pts, frm=2, m=4)["values"]
    np.testing.assert_allclose(b_e, got[:, :, 2:6], rtol=1e-15, atol=0)
    x = np.random.default_rng(9).uniform(-1, 1, n)
    assert nh.zero_tables_probe(np.column_stack([np.ones(n), x]), log_o, fid) is None
