"""Pins the oracle's opt_pml against nebula 1.5.8 (the values edge-rs quotes in
src/sc/pml.rs:2151-2700) and against closed forms / brute force."""
import numpy as np
import pytest

import oracle

# Fixture A (pml.rs:2151-2180): 12 cells, 3 subjects of 4, 5, 3 cells.
DESIGN = np.array([1, 0, 1, 1, 1, 0, 1, 1, 1, 1, 1, 0, 1, 1, 1, 0, 1, 1, 1, 0, 1, 1, 1, 1],
                  dtype=float).reshape(12, 2)
OFFSET = np.array([0.25, -0.5, 0.75, 0.0, 0.5, -0.25, 0.125, 0.375, -0.125, 0.25, 0.0, -0.375])
START = np.array([0, 4, 9, 12], dtype=np.uint64)
COUNTS = np.array([3.0, 5.0, 2.0, 4.0, 1.0, 7.0, 2.0, 6.0])
CELLS = np.array([0, 2, 3, 5, 6, 7, 9, 10], dtype=np.uint64)
TOTAL = np.array([10.0, 12.0, 8.0])
COUNTS_B = np.array([3.0, 5.0, 2.0, 2.0, 6.0])
CELLS_B = np.array([0, 2, 3, 9, 10], dtype=np.uint64)
TOTAL_B = np.array([10.0, 0.0, 8.0])
TOL = 1e-11


def close(got, want, rel=TOL, abs_=1e-12):
    np.testing.assert_allclose(np.asarray(got), np.asarray(want), rtol=rel, atol=abs_)


def fit_a(beta, subject=0.25, cell=2.0, **kw):
    return oracle.opt_pml(DESIGN, OFFSET, COUNTS, CELLS, START, TOTAL, beta, subject, cell, **kw)


def test_well_behaved_matches_nebula():
    f = fit_a([0.5, -0.25])
    close(f["beta"], [1.1055943907637618, -0.98348970166026206])
    close(f["log_w"], [-0.025429869309964156, 0.022419629678247836, 0.33781545877640862])
    close(f["information"], [5.4708370512237092, 2.6193993778249318, 2.6193993778249318,
                             4.0223986651682804])
    close(f["log_likelihood"], -60.378642466146786)
    close(f["log_det"], 6.0421923783970968)
    assert f["second_order"] == 0.0
    assert (f["iterations"], f["backtracks"], f["convergence"]) == (4, 0, 1)


def test_third_order_laplace():
    f = fit_a([0.5, -0.25], ord=3)
    close(f["beta"], [1.1055943907637618, -0.98348970166026206])
    close(f["second_order"], 0.0028301625845031939)


def test_reml_log_det():
    f = fit_a([0.5, -0.25], reml=True)
    close(f["log_det"], 8.759838692617901)


def test_all_zero_subject():
    f = oracle.opt_pml(DESIGN, OFFSET, COUNTS_B, CELLS_B, START, TOTAL_B, [0.5, -0.25], 0.25, 2.0)
    close(f["beta"], [0.15781792039265752, -0.098178037869394941])
    close(f["log_w"], [0.31781101972092152, -0.83477809302857631, 0.46458776408134972])
    close(f["information"], [4.9151627959404793, 2.59325292186585, 2.59325292186585,
                             3.5781777249175555])
    close(f["log_likelihood"], -50.769290856943599)
    close(f["log_det"], 5.3802869340559738)
    f3 = oracle.opt_pml(DESIGN, OFFSET, COUNTS_B, CELLS_B, START, TOTAL_B, [0.5, -0.25], 0.25,
                        2.0, ord=3)
    close(f3["second_order"], 0.029026901671004347)


def test_backtracking_near_bound():
    f = fit_a([90.0, 0.0])
    close(f["beta"], [1.1055943759446161, -0.98348968952057669])
    close(f["log_w"], [-0.025429869876909816, 0.022419639320108459, 0.3378154594265026])
    close(f["information"], [5.4711020927179739, 2.6193969725187061, 2.6193969725187061,
                             4.0224446400576319])
    assert f["iterations"] == 9 and f["convergence"] == 1


def test_fixed_displacement():
    f = fit_a([-90.0, 60.0])
    close(f["beta"], [1.1055943725487707, -0.98348968650982715])
    close(f["log_det"], 6.0423215051590002)
    assert f["iterations"] == 12


def test_critical_point_minus_ten():
    f = fit_a([15.0, 5.0], subject=8.0, cell=0.001)
    close(f["beta"], [4.9609375, -4.9609375])
    close(f["log_w"], [0.0390625, 0.0390625, 4.5234136840078936])
    close(f["information"], [7.0733722800021132e-05, 7.0105348391562988e-05,
                             7.0105348391562988e-05, 0.00016310178559274507])
    close(f["log_det"], -20.515549220945573)
    assert (f["iterations"], f["backtracks"], f["convergence"]) == (3, 11, -10)


def test_stalled_minus_forty():
    f = fit_a([20.0, 0.0], subject=8.0, cell=0.001)
    close(f["beta"], [14.9609375, -4.9609375])
    close(f["log_w"], [-4.9609375, -4.9609375, 4.9962571608576489])
    assert (f["iterations"], f["backtracks"], f["convergence"]) == (2, 12, -40)


def test_iteration_budget_minus_twenty():
    f = fit_a([0.5, -0.25], eps=-1e-300)
    assert f["iterations"] == 50 and f["convergence"] == -20
    close(f["beta"], [1.1055943907637629, -0.98348970166026151])
    close(f["information"], [5.4708368635191373, 2.619399280826618, 2.619399280826618,
                             4.0223986123865263])


def test_collinear_minus_twentyfive():
    design = np.column_stack([DESIGN, DESIGN[:, 1]])
    f = oracle.opt_pml(design, OFFSET, COUNTS, CELLS, START, TOTAL, [0.5, -0.25, 0.1], 0.25, 2.0)
    assert f["convergence"] == -25
    close(f["beta"], [1.1055943330380458, -1.0834897078722583, 0.10000000000000001])
    close(f["log_det"], 6.042689573294485)


def test_single_coefficient():
    f = oracle.opt_pml(np.ones((12, 1)), OFFSET, COUNTS, CELLS, START, TOTAL, [0.0], 0.25, 2.0,
                       reml=True, ord=3)
    close(f["beta"], [0.64526354365336647])
    close(f["log_w"], [0.040002950937536336, 0.10805651074009694, 0.21875064209166895])
    close(f["information"], [5.7533150000267472])
    close(f["log_det"], 7.9235375511695514)
    close(f["second_order"], 0.0044134970888637676)


def test_nbm_matches_nebula():
    f = fit_a([0.5, -0.25], subject=0.3, reml=True, nbm=True)
    close(f["beta"], [1.2213403122060453, -0.98985988766009869])
    close(f["log_w"], [-0.14301073703498665, -0.095894342874540239, 0.23890507990952689])
    close(f["information"], [5.4025646031315233, 2.5076353884976541, 2.5076353884976541,
                             3.9141513741254204])
    close(f["log_likelihood"], -49.318101708775217)
    close(f["log_det"], 8.6579114426780528)
    assert f["iterations"] == 3


def test_pml_ll_grad_matches_nebula():
    obj, grad = oracle.pml_ll_grad(DESIGN, OFFSET, COUNTS, CELLS, START, TOTAL, [0.5, -0.25],
                                   [0.1, -0.2, 0.05], 0.25, 2.0)
    close(obj, 53.926203132942852)
    close(grad, [-5.2597994101747254, -0.19846989257393055, -0.40624529144311583,
                 -2.8367693486273806, -2.0229915782218155])


def fixture_large():
    starts = [0, 12, 32, 48, 64]
    design, offset, counts, cells = [], [], [], []
    total = np.zeros(4)
    for i in range(64):
        design.append([1.0, 1.0 if i % 3 == 0 else 0.0, (i % 7) / 8.0])
        offset.append(((i % 9) - 4.0) / 8.0)
        y = float((i * 7 + 3) % 11)
        if y > 0:
            counts.append(y)
            cells.append(i)
        s = max(j for j in range(4) if starts[j] <= i)
        total[s] += y
    return (np.array(design), np.array(offset), np.array(counts),
            np.array(cells, dtype=np.uint64), np.array(starts, dtype=np.uint64), total)


def test_wider_design_more_cells():
    d, o, y, c, s, t = fixture_large()
    assert list(t) == [58.0, 100.0, 83.0, 82.0]
    f = oracle.opt_pml(d, o, y, c, s, t, [0.5, -0.25, 0.75], 0.35, 1.5, reml=True, ord=3)
    close(f["beta"], [1.4272053191863423, 0.20743588590645803, -0.017982095723719935])
    close(f["log_w"], [0.21944093016930771, 0.12415265795703513, 0.20913336580790309,
                       0.14393280971073011])
    close(f["information"], [8.398261243665857, 2.7062153540842324, 3.0637064599168475,
                             2.7062153540842324, 16.226178650135836, 1.2821611476475319,
                             3.0637064599168475, 1.2821611476475319, 5.5864643706193053])
    close(f["log_likelihood"], -276.5244116459354)
    close(f["log_det"], 18.407453390744379)
    close(f["second_order"], 0.0029993806316514779)
    assert f["iterations"] == 4 and f["convergence"] == 1
    plain = oracle.opt_pml(d, o, y, c, s, t, [0.5, -0.25, 0.75], 0.35, 1.5)
    close(plain["log_det"], 12.052188343414912)


def test_rejects_malformed_input():
    with pytest.raises(oracle.OracleError):
        fit_a([0.5, -0.25], cell=0.0)
    with pytest.raises(oracle.OracleError):
        fit_a([0.5, -0.25], subject=0.0)


def test_linear_algebra_closed_forms():
    lib = oracle.lib()
    a = np.array([2.0, 1.0, 1.0, 2.0])
    assert abs(lib.oracle_min_eigenvalue(a.ctypes.data_as(oracle._dp), 2) - 1.0) < 1e-13
    a = np.array([1.0, 2.0, 2.0, 1.0])  # indefinite: eigenvalues 3, -1
    b = np.array([3.0, 3.0])
    lib.oracle_ldlt_solve(a.ctypes.data_as(oracle._dp), b.ctypes.data_as(oracle._dp), 2)
    close(b, [1.0, 1.0], rel=1e-13)
    m = np.array([4.0, 1.0, 2.0, 1.0, 5.0, 3.0, 2.0, 3.0, 6.0])
    assert abs(lib.oracle_determinant(m.ctypes.data_as(oracle._dp), 3) - 70.0) < 1e-12
    rng = np.random.default_rng(3)
    for n in range(1, 9):
        q = rng.normal(size=(n, n))
        spd = q @ q.T + n * np.eye(n)
        rhs = rng.normal(size=n)
        x = rhs.copy()
        lib.oracle_ldlt_solve(spd.copy().ravel().ctypes.data_as(oracle._dp),
                              x.ctypes.data_as(oracle._dp), n)
        close(x, np.linalg.solve(spd, rhs), rel=1e-10)
        assert abs(lib.oracle_min_eigenvalue(np.ascontiguousarray(spd.ravel()).ctypes.data_as(
            oracle._dp), n) - np.linalg.eigvalsh(spd)[0]) < 1e-9 * np.abs(spd).max()


def test_optimum_is_a_stationary_point_of_the_penalised_likelihood():
    """Brute-force pin independent of the quoted values: at opt_pml's optimum
    the gradient of the penalised log-likelihood vanishes (gamma prior
    gradient alpha - lambda*w added to the data part)."""
    d, o, y, c, s, t = fixture_large()
    f = oracle.opt_pml(d, o, y, c, s, t, [0.5, -0.25, 0.75], 0.35, 1.5, eps=1e-12)
    beta, lw = f["beta"], f["log_w"]
    # finite differences of the penalised log-likelihood via pml_ll_grad's
    # data part with prior (alpha, lambda) from gamma_prior.
    exps = np.exp(0.35)
    alpha, lam = 1 / (exps - 1), 1 / (np.sqrt(exps) * (exps - 1))

    def pll(b, w):
        eta = o + d @ b
        sub = np.repeat(np.arange(4), np.diff(s.astype(int)))
        e = np.exp(eta + w[sub])
        yy = np.zeros(64)
        yy[c.astype(int)] = y
        ll = np.sum(yy * eta) + np.sum(w * t) - np.sum((1.5 + yy) * np.log(e + 1.5))
        return ll + alpha * w.sum() - lam * np.exp(w).sum()

    h = 1e-6
    for j in range(3):
        e = np.zeros(3)
        e[j] = h
        g = (pll(beta + e, lw) - pll(beta - e, lw)) / (2 * h)
        assert abs(g) < 1e-5
    for j in range(4):
        e = np.zeros(4)
        e[j] = h
        g = (pll(beta, lw + e) - pll(beta, lw - e)) / (2 * h)
        assert abs(g) < 1e-5
    assert abs(pll(beta, lw) - f["log_likelihood"]) < 1e-9 * abs(f["log_likelihood"])
