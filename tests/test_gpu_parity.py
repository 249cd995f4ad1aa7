"""HIP path against the oracle and the reference's golden data, through the
C ABI (libnebula_hip.so).  Marked gpu: these run on an MI355X."""
import numpy as np
import pytest

import oracle
from fixtures import SC_DATASETS, assert_close, dense_to_csr, golden8, load_sc, read_matrix
import nebula_hip as nh
from nebula_hip import synth

pytestmark = pytest.mark.gpu

PROBLEMS = {"rosen": (0, [-1.2, 1.0], [-2.0, -1.0], [2.0, 3.0], 4),
            "rosen_edge": (0, [0.4001, 0.2], [0.4, 0.2], [2.0, 2.0], 4),
            "corner": (1, [1.0, 1.0], [0.0, 0.0], [2.0, 2.0], 4),
            "rescue": (2, [-1.0, 0.25], [-2.0, -2.0], [2.0, 2.0], 4),
            "roundoff": (3, [0.0, 0.0], [-2.0, -2.0], [2.0, 2.0], -4),
            "chain3": (4, [0.5, 2.0, -1.0], [-3.0] * 3, [3.0] * 3, 4)}


@pytest.mark.parametrize("tag", list(PROBLEMS))
def test_device_bobyqa_retraces_nloptr(tag):
    """The stepper running on the device asks for nloptr's points bit for bit
    (traces from src/numeric/bobyqa.rs:2947-2970's fixtures)."""
    prob, x0, lo, hi, status = PROBLEMS[tag]
    want = read_matrix(f"bobyqa_{tag}.csv")
    got, st = nh.bobyqa_device_trace(prob, x0, lo, hi, max_trace=len(want) + 10)
    assert st == status
    assert got.shape == want.shape
    assert np.array_equal(got, want)


# pml fixture A (src/sc/pml.rs:2151-2180)
DESIGN = np.array([1, 0, 1, 1, 1, 0, 1, 1, 1, 1, 1, 0, 1, 1, 1, 0, 1, 1, 1, 0, 1, 1, 1, 1],
                  dtype=float).reshape(12, 2)
OFFSET = np.array([0.25, -0.5, 0.75, 0.0, 0.5, -0.25, 0.125, 0.375, -0.125, 0.25, 0.0, -0.375])
START = np.array([0, 4, 9, 12], dtype=np.uint64)
COUNTS = np.array([3.0, 5.0, 2.0, 4.0, 1.0, 7.0, 2.0, 6.0])
CELLS = np.array([0, 2, 3, 5, 6, 7, 9, 10], dtype=np.uint32)


def test_device_f64_fit_reproduces_nebula_values():
    """The device opt_pml in f64 (stage three's mode) follows opt_pml's control
    flow exactly: nebula 1.5.8's optimum at 1e-11 and the same iteration count."""
    rb = nh.ResidentBatch(DESIGN, OFFSET, START, [0, len(COUNTS)], CELLS, COUNTS)
    r = rb.solve([0, 0, 0], [0.25, 0.25, 8.0], [2.0, 2.0, 0.001],
                 [[0.5, -0.25], [90.0, 0.0], [15.0, 5.0]], mode=nh.ResidentBatch.SOLVE_F64)
    np.testing.assert_allclose(r["beta"][0], [1.1055943907637618, -0.98348970166026206], rtol=1e-11)
    np.testing.assert_allclose(r["log_w"][0], [-0.025429869309964156, 0.022419629678247836,
                                               0.33781545877640862], rtol=1e-11)
    np.testing.assert_allclose(r["information"][0], [5.4708370512237092, 2.6193993778249318,
                                                     2.6193993778249318, 4.0223986651682804],
                               rtol=1e-11)
    np.testing.assert_allclose(r["log_likelihood"][0], -60.378642466146786, rtol=1e-11)
    np.testing.assert_allclose(r["log_det"][0], 6.0421923783970968, rtol=1e-11)
    assert r["iterations"][0] == 4
    np.testing.assert_allclose(r["beta"][1], [1.1055943759446161, -0.98348968952057669], rtol=1e-11)
    assert r["iterations"][1] == 9
    np.testing.assert_allclose(r["beta"][2], [4.9609375, -4.9609375], rtol=1e-11)
    assert (r["iterations"][2], r["backtracks"][2]) == (3, 11)
    # check_convergence codes of the same fits (src/sc/pml.rs:1111-1139)
    assert r["convergence"].tolist() == [1, 1, -10]


def test_device_f64_fit_convergence_codes():
    """The reply's convergence code for the oracle's -40 (stalled), -20
    (iteration budget, eps < 0) and -25 (collinear design) fixtures
    (src/sc/pml.rs:2151-2700), with the oracle's iterations and backtracks."""
    rb = nh.ResidentBatch(DESIGN, OFFSET, START, [0, len(COUNTS)], CELLS, COUNTS)
    r = rb.solve([0], [8.0], [0.001], [[20.0, 0.0]], mode=nh.ResidentBatch.SOLVE_F64)
    o = oracle.opt_pml(DESIGN, OFFSET, COUNTS, CELLS.astype(np.uint64), START,
                       np.array([10.0, 12.0, 8.0]), [20.0, 0.0], 8.0, 0.001)
    assert (r["iterations"][0], r["backtracks"][0], r["convergence"][0]) == (2, 12, -40)
    assert (o["iterations"], o["backtracks"], o["convergence"]) == (2, 12, -40)
    r = rb.solve([0], [0.25], [2.0], [[0.5, -0.25]], eps=-1e-300, mode=nh.ResidentBatch.SOLVE_F64)
    assert (r["iterations"][0], r["convergence"][0]) == (50, -20)
    rb.close()
    design = np.column_stack([DESIGN, DESIGN[:, 1]])
    rb = nh.ResidentBatch(design, OFFSET, START, [0, len(COUNTS)], CELLS, COUNTS)
    r = rb.solve([0], [0.25], [2.0], [[0.5, -0.25, 0.1]], mode=nh.ResidentBatch.SOLVE_F64)
    assert r["convergence"][0] == -25
    rb.close()


def test_resident_submit_collect_matches_solve():
    """submit returns with the launch in flight; collect gives the same
    replies as solve; a second submit before collect and a collect with the
    wrong count are rejected, and the handle stays usable."""
    w, d, lo, fid = synthetic_gene_batch()
    rb = nh.ResidentBatch(d, lo, fid, w.indptr.astype(np.uint64), w.indices, w.data)
    genes = np.arange(w.n_genes)
    subj = np.linspace(0.05, 0.6, w.n_genes)
    cell = np.linspace(0.3, 4.0, w.n_genes)
    beta0 = np.zeros((w.n_genes, w.n_coef))
    beta0[:, 0] = -1.0
    a = rb.solve(genes, subj, cell, beta0)
    rb.submit(genes, subj, cell, beta0)
    with pytest.raises(nh.InvalidArgument):
        rb.submit(genes, subj, cell, beta0)
    with pytest.raises(nh.InvalidArgument):
        rb.collect(n=len(genes) - 1)
    b = rb.collect()
    for key in a:
        np.testing.assert_array_equal(a[key], b[key])
    with pytest.raises(nh.InvalidArgument):
        rb.collect()
    c = rb.solve(genes, subj, cell, beta0)
    np.testing.assert_array_equal(a["log_likelihood"], c["log_likelihood"])
    rb.close()


def synthetic_gene_batch(n_cells=6000, k=12, p=4, n_genes=6, seed=7):
    w = synth.generate("t", n_genes, n_cells, k, p, "dirichlet2", ("normal", -1.0, 0.5), seed)
    fid = np.concatenate([[0], np.cumsum(w.truth["sizes"])]).astype(np.uint64)
    # centre the design as the driver does
    d = w.design.copy()
    for j in range(1, p):
        d[:, j] = (d[:, j] - d[:, j].mean()) / d[:, j].std()
    return w, d, np.log(w.offset), fid


def test_fp32_fit_with_fp64_finish_matches_oracle():
    """Stage two's device fit: FP32 Newton then the FP64 finish, against the
    oracle's f64 opt_pml run to a tight tolerance (so its penultimate-iterate
    log-determinant is the optimum's)."""
    w, d, lo, fid = synthetic_gene_batch()
    rb = nh.ResidentBatch(d, lo, fid, w.indptr.astype(np.uint64), w.indices, w.data)
    genes = np.arange(w.n_genes)
    subj = np.linspace(0.05, 0.6, w.n_genes)
    cell = np.linspace(0.3, 4.0, w.n_genes)
    beta0 = np.zeros((w.n_genes, w.n_coef))
    beta0[:, 0] = -1.0
    r = rb.solve(genes, subj, cell, beta0, mode=nh.ResidentBatch.SOLVE_F32_FINISH)
    for g in genes:
        a, b = int(w.indptr[g]), int(w.indptr[g + 1])
        cells = w.indices[a:b].astype(np.uint64)
        tot = np.zeros(len(fid) - 1)
        np.add.at(tot, np.searchsorted(fid, cells, side="right") - 1, w.data[a:b])
        o = oracle.opt_pml(d, lo, w.data[a:b], cells, fid, tot, beta0[g], subj[g], cell[g],
                           eps=1e-13)
        assert abs(r["log_likelihood"][g] - o["log_likelihood"]) <= 1e-10 * abs(o["log_likelihood"])
        np.testing.assert_allclose(r["beta"][g], o["beta"], rtol=1e-6, atol=1e-8)
        np.testing.assert_allclose(r["log_det"][g], o["log_det"], rtol=1e-8)
        assert r["finish_steps"][g] >= 1


@pytest.mark.parametrize("cell", [1e-4, 0.5, 2.0, 1000.0])
def test_fp32_fit_holds_across_the_variance_box(cell):
    """The stage-two fit (FP32 Newton + FP64 finish) over the subject variances
    the search visits, sigma^2 from 1e-4 to 1 (tests/e2e_nebula_gpu.rs:600-691:
    subject = ln(1 + sigma^2)), at cell sizes from the box's lower corner 1e-4
    to its upper 1000: the log-likelihood within 1e-9 of the oracle's f64
    optimum, the log-determinant within 1e-6."""
    w, d, lo, fid = synthetic_gene_batch(seed=17)
    rb = nh.ResidentBatch(d, lo, fid, w.indptr.astype(np.uint64), w.indices, w.data)
    worst_ll = worst_det = 0.0
    for sigma2 in (1e-4, 1e-3, 1e-2, 1e-1, 1.0):
        subject = np.log1p(sigma2)
        genes = np.arange(w.n_genes)
        beta0 = np.zeros((w.n_genes, w.n_coef))
        r = rb.solve(genes, subject, cell, beta0, mode=nh.ResidentBatch.SOLVE_F32_FINISH)
        for g in genes:
            a, b = int(w.indptr[g]), int(w.indptr[g + 1])
            cells = w.indices[a:b].astype(np.uint64)
            tot = np.zeros(len(fid) - 1)
            np.add.at(tot, np.searchsorted(fid, cells, side="right") - 1, w.data[a:b])
            o = oracle.opt_pml(d, lo, w.data[a:b], cells, fid, tot, beta0[g], subject, cell,
                               eps=1e-13)
            dll = abs(r["log_likelihood"][g] - o["log_likelihood"]) / abs(o["log_likelihood"])
            ddet = abs(r["log_det"][g] - o["log_det"]) / max(abs(o["log_det"]), 1.0)
            worst_ll, worst_det = max(worst_ll, dll), max(worst_det, ddet)
            assert dll <= 1e-9, (sigma2, cell, g, dll)
            assert ddet <= 1e-6, (sigma2, cell, g, ddet)
    print(f"cell {cell}: worst |dll| {worst_ll:.2e}, worst |dlogdet| {worst_det:.2e}")
    rb.close()


def test_f64_mode_matches_oracle_iterations():
    """Stage three's mode (f64 opt_pml, ord 3) against the oracle on 6000-cell
    genes: same iteration count, log-likelihood to 1e-12, and the quantities
    taken at the penultimate iterate (log-det, information) to 1e-9.

    The last step's accept test is `likdif < 0`; when the last likdif is at
    rounding level (1.8e-12 on gene 4 here) the two summation orders can
    take different numbers of halvings, moving beta by a fraction of that
    last step.  Where the backtrack counts differ, beta and the Laplace term
    at the final point are held to 1e-6 instead of 1e-10."""
    w, d, lo, fid = synthetic_gene_batch(seed=11)
    rb = nh.ResidentBatch(d, lo, fid, w.indptr.astype(np.uint64), w.indices, w.data)
    genes = np.arange(w.n_genes)
    beta0 = np.zeros((w.n_genes, w.n_coef))
    r = rb.solve(genes, 0.3, 1.5, beta0, mode=nh.ResidentBatch.SOLVE_F64, ord=3)
    same = 0
    for g in genes:
        a, b = int(w.indptr[g]), int(w.indptr[g + 1])
        cells = w.indices[a:b].astype(np.uint64)
        tot = np.zeros(len(fid) - 1)
        np.add.at(tot, np.searchsorted(fid, cells, side="right") - 1, w.data[a:b])
        o = oracle.opt_pml(d, lo, w.data[a:b], cells, fid, tot, beta0[g], 0.3, 1.5, ord=3)
        assert r["iterations"][g] == o["iterations"]
        tie = r["backtracks"][g] != o["backtracks"]
        if tie:
            assert abs(o["log_likelihood"] - o["log_likelihood_prev"]) < 1e-9
        else:
            same += 1
        rt = 1e-6 if tie else 1e-10
        np.testing.assert_allclose(r["beta"][g], o["beta"], rtol=rt, atol=1e-12)
        np.testing.assert_allclose(r["log_likelihood"][g], o["log_likelihood"], rtol=1e-12)
        np.testing.assert_allclose(r["log_det"][g], o["log_det"], rtol=1e-10)
        np.testing.assert_allclose(r["second_order"][g], o["second_order"],
                                   rtol=1e-6 if tie else 1e-9)
        np.testing.assert_allclose(r["information"][g], o["information"], rtol=1e-9, atol=1e-10)
    assert same >= w.n_genes - 1


def test_counts_without_an_f32_value_reach_the_device_in_f64():
    """Whole counts cross to the device as f32 and are widened there; counts
    that f32 cannot hold (thirds here) cross in f64.  Either way the f64 fit
    matches the oracle's log-likelihood to 1e-12, which a count rounded to f32
    (relative error 3e-8) would miss."""
    w, d, lo, fid = synthetic_gene_batch(seed=13)
    for scale in (1.0, 1.0 / 3.0):
        data = w.data * scale
        if scale != 1.0:
            assert np.any(data.astype(np.float32).astype(np.float64) != data)
        rb = nh.ResidentBatch(d, lo, fid, w.indptr.astype(np.uint64), w.indices, data)
        genes = np.arange(w.n_genes)
        beta0 = np.zeros((w.n_genes, w.n_coef))
        r = rb.solve(genes, 0.3, 1.5, beta0, mode=nh.ResidentBatch.SOLVE_F64)
        rb.close()
        for g in genes:
            a, b = int(w.indptr[g]), int(w.indptr[g + 1])
            cells = w.indices[a:b].astype(np.uint64)
            tot = np.zeros(len(fid) - 1)
            np.add.at(tot, np.searchsorted(fid, cells, side="right") - 1, data[a:b])
            o = oracle.opt_pml(d, lo, data[a:b], cells, fid, tot, beta0[g], 0.3, 1.5)
            np.testing.assert_allclose(r["log_likelihood"][g], o["log_likelihood"], rtol=1e-12,
                                       err_msg=f"scale {scale} gene {g}")


ENGINES = {"plain": nh.ENGINE_PLAIN, "batched": nh.ENGINE_BATCHED}


def run_hip(ip, ix, dv, n_cells, subject, design, n_coef, offset, params, engine):
    counts = nh.CompressedSparse(dv, ix, ip, (len(ip) - 1, n_cells))
    return nh.nebula_sparse_hip(counts, subject, design, n_coef, offset, params, engine=engine)


def assert_objective_parity(ip, ix, dv, n_cells, subject, design, n_coef, offset, prm, fit,
                            bound=1e-6):
    """The stage-two criterion: the oracle's f64 profile objective at the HIP
    minimiser and at the oracle's own minimiser agree to `bound` relative,
    gene by gene, over the genes the oracle searched."""
    ref = oracle.nebula_sparse(ip, ix, dv, n_cells, subject, design, n_coef, offset, prm,
                               with_stage_two=True)
    assert np.array_equal(fit.gene_index, ref["gene_index"])
    s2 = ref["stage_two"]
    g = np.nonzero(s2[:, 0] > 0)[0]
    assert len(g) > 0
    pts_hip = np.c_[fit.subject_overdispersion, 1.0 / fit.cell_overdispersion][g]
    v_hip = oracle.profile_objective(ip, ix, dv, n_cells, subject, design, n_coef, offset, prm,
                                     g, pts_hip)
    v_cpu = oracle.profile_objective(ip, ix, dv, n_cells, subject, design, n_coef, offset, prm,
                                     g, s2[g, 1:3])
    assert np.all(np.isfinite(v_cpu))
    rel = np.abs(v_hip - v_cpu) / np.abs(v_cpu)
    assert np.all(np.isfinite(rel)) and rel.max() < bound, (rel.max(), int(g[np.argmax(rel)]))
    return rel.max()


@pytest.mark.parametrize("engine", list(ENGINES))
def test_golden8_matches_r(engine):
    """The 8-gene golden (src/sc/nebula.rs:2457-2560) under the GPU path's
    gates (tests/e2e_nebula.rs:1010-1020), and the stage-two objective at the
    oracle's precision."""
    from test_oracle_nebula import GPU_TOLS
    c, subject, d, offset = golden8()
    ip, ix, dv = dense_to_csr(c)
    fit = run_hip(ip, ix, dv, 150, subject, d, 3, offset,
                  nh.NebulaParams(cutoff_cell=0.0), ENGINES[engine])
    from fixtures import read_table
    _, exp = read_table("nebula_expected.csv")
    assert list(fit.gene_index) == list(range(8))
    assert_close(fit.coefficients, exp[:, 0:3], *GPU_TOLS["coef"], label="coef")
    assert_close(fit.se, exp[:, 3:6], *GPU_TOLS["se"], label="se")
    assert_close(fit.subject_overdispersion, exp[:, 10], *GPU_TOLS["subject"], label="subject")
    assert_close(fit.cell_overdispersion, exp[:, 11], *GPU_TOLS["cell"], label="cell")
    assert list(np.nonzero(fit.sigma_at_bound)[0]) == [1, 3, 6]
    assert_objective_parity(ip, ix, dv, 150, subject, d, 3, offset,
                            oracle.params(cutoff_cell=0.0), fit)


@pytest.mark.parametrize("engine", list(ENGINES))
@pytest.mark.parametrize("tag", [t for t in SC_DATASETS if SC_DATASETS[t][2] <= 8])
def test_sc_dataset_matches_r(tag, engine):
    """R nebula's values on the e2e datasets under the GPU path's gates
    (tests/e2e_nebula.rs:1010-1074, per-dataset loosening :232-254): NEBULA-HL
    and NEBULA-LN (its LN+HL genes refit the subject variance alone), the
    count of genes pinned at the variance floor, the Wald test of the HIP fit
    against R's p and se columns; and the stage-two objective parity against
    the oracle's own searches."""
    from test_oracle_nebula import GPU_TOLS, check_sc
    d = load_sc(tag)
    ip, ix, dv = dense_to_csr(d["counts"])
    n_cells = d["counts"].shape[1]
    fit = run_hip(ip, ix, dv, n_cells, d["subject"], d["design"], d["n_coef"],
                  d["offset"], nh.NebulaParams(method=d["method"]), ENGINES[engine])
    print(f"{tag} {engine}: {len(fit.gene_index)} genes", flush=True)
    check_sc(tag, dict(coefficients=fit.coefficients, se=fit.se, covariance=fit.covariance,
                       subject_overdispersion=fit.subject_overdispersion,
                       cell_overdispersion=fit.cell_overdispersion, convergence=fit.convergence,
                       gene_index=fit.gene_index, sigma_at_bound=fit.sigma_at_bound),
             base=GPU_TOLS, wald=nh.glm_sc_test)
    # stage two (both components, or NEBULA-LN's subject variance alone)
    assert_objective_parity(ip, ix, dv, n_cells, d["subject"], d["design"], d["n_coef"],
                            d["offset"], oracle.params(method=d["method"]), fit)


@pytest.mark.parametrize("engine", list(ENGINES))
def test_synthetic_config2_shape_matches_oracle(engine):
    """A scaled config-2 workload (Dirichlet subject sizes, p = 4): stage-two
    minimisers re-evaluated by the oracle's f64 profile objective agree to
    |dloglik| < 1e-6 relative; coefficients and variances to search accuracy."""
    w = synth.config(2, seed=21, n_genes=24, n_cells=20000)
    prm = oracle.params(method="HL")
    ref = oracle.nebula_sparse(w.indptr, w.indices, w.data, w.n_cells, w.subject_id, w.design,
                               w.n_coef, w.offset, prm, with_stage_two=True)
    fit = run_hip(w.indptr, w.indices, w.data, w.n_cells, w.subject_id, w.design, w.n_coef,
                  w.offset, nh.NebulaParams(method="HL"), ENGINES[engine])
    assert np.array_equal(fit.gene_index, ref["gene_index"])
    g = np.arange(len(fit.gene_index))
    pts_hip = np.c_[fit.subject_overdispersion, 1.0 / fit.cell_overdispersion]
    pts_cpu = ref["stage_two"][:, 1:3]
    v_hip = oracle.profile_objective(w.indptr, w.indices, w.data, w.n_cells, w.subject_id,
                                     w.design, w.n_coef, w.offset, prm, g, pts_hip)
    v_cpu = oracle.profile_objective(w.indptr, w.indices, w.data, w.n_cells, w.subject_id,
                                     w.design, w.n_coef, w.offset, prm, g, pts_cpu)
    rel = np.abs(v_hip - v_cpu) / np.abs(v_cpu)
    assert rel.max() < 1e-6, rel.max()
    assert_close(fit.coefficients, ref["coefficients"], 1e-3, 1e-6, label="coef")
    assert_close(fit.subject_overdispersion, ref["subject_overdispersion"], 2e-2, 1e-4,
                 label="subject")
    assert_close(fit.cell_overdispersion, ref["cell_overdispersion"], 1e-2, label="cell")


@pytest.mark.parametrize("p", [2, 4, 6, 8])
@pytest.mark.parametrize("wg", [256, 512, 1024])
@pytest.mark.parametrize("engine", list(ENGINES))
def test_kernel_objective_matches_oracle_at_every_shape(engine, wg, p):
    """For each design width and workgroup size the engines launch, the
    objective a search reports at its minimiser matches the oracle's f64
    profile objective at that point (the kernel's FP64 finish, not just the
    search, is checked), and the minimisers agree with the oracle's search to
    |dloglik| < 1e-6."""
    w = synth.generate("shape", 40, 16_000, 12, p, "dirichlet2", ("uniform", -4.0, 0.5),
                       seed=300 + p, design_kind="config3" if p == 8 else "generic")
    counts = nh.CompressedSparse(w.data, w.indices, w.indptr, (w.n_genes, w.n_cells))
    s = nh.Session(counts, w.subject_id, w.design, w.n_coef, w.offset,
                   nh.NebulaParams(method="HL"), wg_size=wg)
    try:
        r = s.stage_two(ENGINES[engine])
    finally:
        s.close()
    prm = oracle.params(method="HL")
    g = np.arange(len(r["minimisers"]))
    v = oracle.profile_objective(w.indptr, w.indices, w.data, w.n_cells, w.subject_id, w.design,
                                 w.n_coef, w.offset, prm, g, r["minimisers"])
    # evaluation error of the device objective at the same point: typically
    # 1e-12..1e-8; a single evaluation can move by the stopping rule's jitter
    # (log-determinant at the penultimate Newton iterate, ~1e-6 relative)
    rel = np.abs(r["objectives"] - v) / np.abs(v)
    assert np.all(np.isfinite(rel))
    assert np.median(rel) < 1e-8 and rel.max() < 1e-5, (np.median(rel), rel.max())
    ref = oracle.nebula_sparse(w.indptr, w.indices, w.data, w.n_cells, w.subject_id, w.design,
                               w.n_coef, w.offset, prm, with_stage_two=True)
    vc = oracle.profile_objective(w.indptr, w.indices, w.data, w.n_cells, w.subject_id,
                                  w.design, w.n_coef, w.offset, prm, g, ref["stage_two"][:, 1:3])
    assert np.max(np.abs(v - vc) / np.abs(vc)) < 1e-6


@pytest.mark.parametrize("engine", ["plain", "batched", "learned"])
def test_many_subjects_work_arrays_in_global_memory(engine):
    """180 subjects at p = 8: a fit's work arrays (63 k + 53 W + 36 doubles)
    exceed the 48 KB LDS budget, so every engine runs them in global memory;
    objectives at the minimisers match the oracle's f64 profile objective and
    the minimisers the oracle's own search, as at the other shapes."""
    w = synth.generate("manyk", 8, 14_400, 180, 8, "equal", ("uniform", -2.5, 0.5), seed=77,
                       design_kind="config3")
    assert 63 * 180 + 53 * 4 + 36 > 48 * 1024 // 8
    counts = nh.CompressedSparse(w.data, w.indices, w.indptr, (w.n_genes, w.n_cells))
    weights = nh.load_weights() if engine == "learned" else None
    s = nh.Session(counts, w.subject_id, w.design, w.n_coef, w.offset,
                   nh.NebulaParams(method="HL"), wg_size=512, mlp_weights=weights)
    try:
        eng = {"plain": nh.ENGINE_PLAIN, "batched": nh.ENGINE_BATCHED,
               "learned": nh.ENGINE_LEARNED}[engine]
        r = s.stage_two(eng)
    finally:
        s.close()
    prm = oracle.params(method="HL")
    g = np.arange(len(r["minimisers"]))
    v = oracle.profile_objective(w.indptr, w.indices, w.data, w.n_cells, w.subject_id, w.design,
                                 w.n_coef, w.offset, prm, g, r["minimisers"])
    # the log-determinant sums 180 subjects, and the oracle's is taken at its
    # penultimate Newton iterate: evaluations differ by ~1e-8 here (1e-10 at 12)
    rel = np.abs(r["objectives"] - v) / np.abs(v)
    assert np.all(np.isfinite(rel))
    assert np.median(rel) < 1e-7 and rel.max() < 1e-5, (np.median(rel), rel.max())
    ref = oracle.nebula_sparse(w.indptr, w.indices, w.data, w.n_cells, w.subject_id, w.design,
                               w.n_coef, w.offset, prm, with_stage_two=True)
    vc = oracle.profile_objective(w.indptr, w.indices, w.data, w.n_cells, w.subject_id,
                                  w.design, w.n_coef, w.offset, prm, g, ref["stage_two"][:, 1:3])
    assert np.max(np.abs(v - vc) / np.abs(vc)) < 1e-6


@pytest.mark.parametrize("engine", list(ENGINES))
def test_at_bound_code_matches_oracle(engine):
    """check_convergence's -60 (a variance component on its upper subject or
    lower cell bound after stage two, src/sc/pml.rs:1111-1139 called from
    finish_gene, src/sc/nebula.rs:966-968): with the subject variance capped
    at 0.02 on sc_small, the genes the oracle reports at -60 are the HIP
    path's, gene for gene."""
    d = load_sc("sc_small")
    ip, ix, dv = dense_to_csr(d["counts"])
    n_cells = d["counts"].shape[1]
    fit = run_hip(ip, ix, dv, n_cells, d["subject"], d["design"], d["n_coef"], d["offset"],
                  nh.NebulaParams(method="LN", max=(0.02, 1000.0)), ENGINES[engine])
    ref = oracle.nebula_sparse(ip, ix, dv, n_cells, d["subject"], d["design"], d["n_coef"],
                               d["offset"], oracle.params(method="LN", max=(0.02, 1000.0)))
    assert np.array_equal(fit.gene_index, ref["gene_index"])
    want = ref["convergence"] == -60
    assert want.sum() > 20
    np.testing.assert_array_equal(fit.convergence == -60, want)
    assert np.all(fit.subject_overdispersion[want] == 0.02)
