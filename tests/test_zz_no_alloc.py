"""nh_resident_solve_noalloc against nh_resident_solve through the C ABI: the
same fits, bit for bit, on two calls in a row (the second smaller), with the
fit's work arrays in LDS (20 subjects) and in the scratch array's global
work area (180 subjects, 8 coefficients); a batch beyond the scratch array's
size is refused before any device work."""
import numpy as np
import pytest

import nebula_hip as nh
from nebula_hip import synth

pytestmark = pytest.mark.gpu

KEYS = ("beta", "log_w", "information", "log_likelihood", "log_likelihood_prev", "log_det",
        "second_order", "iterations", "backtracks", "finish_steps", "convergence")


def resident(n_genes, n_cells, k, p, sizes, design_kind, seed):
    w = synth.generate("na", n_genes, n_cells, k, p, sizes, ("normal", -1.0, 0.5), seed,
                       design_kind=design_kind)
    fid = np.concatenate([[0], np.cumsum(w.truth["sizes"])]).astype(np.uint64)
    d = w.design.copy()
    for j in range(1, p):
        d[:, j] = (d[:, j] - d[:, j].mean()) / d[:, j].std()
    return w, nh.ResidentBatch(d, np.log(w.offset), fid, w.indptr.astype(np.uint64),
                               w.indices, w.data)


def assert_same(a, b, label):
    for key in KEYS:
        assert np.array_equal(a[key], b[key]), f"{label}: {key}"


@pytest.mark.parametrize("shape", ["lds_work", "global_work"])
def test_noalloc_matches_solve_on_two_calls(shape):
    if shape == "lds_work":
        w, rb = resident(8, 6000, 20, 4, "dirichlet2", "generic", 31)
    else:
        w, rb = resident(6, 14_400, 180, 8, "equal", "config3", 32)
    print(f"no-alloc {shape}: {w.n_genes} genes, k={len(w.truth['sizes'])}, p={w.n_coef}",
          flush=True)
    try:
        for n in (w.n_genes, w.n_genes // 2):
            genes = np.arange(n)
            subj = np.linspace(0.05, 0.6, n)
            cell = np.linspace(0.3, 4.0, n)
            beta0 = np.zeros((n, w.n_coef))
            beta0[:, 0] = -1.0
            lw0 = np.full((n, len(w.truth["sizes"])), 0.01)
            for mode in (nh.ResidentBatch.SOLVE_F32_FINISH, nh.ResidentBatch.SOLVE_F64):
                a = rb.solve(genes, subj, cell, beta0, mode=mode)
                b = rb.solve_noalloc(genes, subj, cell, beta0, mode=mode)
                assert_same(a, b, f"{shape} n={n} mode={mode}")
            a = rb.solve(genes, subj, cell, beta0, log_w_init=lw0)
            b = rb.solve_noalloc(genes, subj, cell, beta0, log_w_init=lw0)
            assert_same(a, b, f"{shape} n={n} log_w_init")
    finally:
        rb.close()


def test_noalloc_refuses_a_batch_beyond_its_scratch():
    w, rb = resident(2, 14_400, 180, 8, "equal", "config3", 33)
    try:
        # at 180 subjects and p = 8 a fit's work area (~72 KB) lives in global
        # memory: 1,000 of them need about twice the 32 MiB array
        n = 1000
        with pytest.raises(nh.InvalidArgument, match="scratch"):
            rb.solve_noalloc(np.zeros(n, dtype=np.int64), 0.3, 1.5, np.zeros(w.n_coef))
        # the handle and the entry point still work afterwards
        a = rb.solve([0, 1], 0.3, 1.5, np.zeros(w.n_coef))
        b = rb.solve_noalloc([0, 1], 0.3, 1.5, np.zeros(w.n_coef))
        assert_same(a, b, "after refusal")
    finally:
        rb.close()
