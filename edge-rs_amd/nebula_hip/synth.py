"""Seeded synthetic single-cell data for NEBULA workloads.

Counts ~ NB(mean = offset * exp(x'beta + b_s), size = 1/phi), subject effects
b_s ~ N(0, sigma^2), cells grouped contiguously by subject.  The offset is
library size over its mean (log-offset = log(lib / mean lib)), the convention
under which the stated non-zero fractions hold.  Genes come out as a
gene-major CSR (u32 indptr/indices, f64 values), the input of
nebula_sparse_hip.
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class Workload:
    name: str
    indptr: np.ndarray
    indices: np.ndarray
    data: np.ndarray
    n_genes: int
    n_cells: int
    subject_id: np.ndarray
    design: np.ndarray
    n_coef: int
    offset: np.ndarray
    truth: dict

    @property
    def nnz(self):
        return int(self.indptr[-1])


def _subject_sizes(rng, n_cells, k, kind, min_size=2):
    if kind == "equal":
        w = np.full(k, 1.0 / k)
    elif kind == "dirichlet2":
        w = rng.dirichlet(np.full(k, 2.0))
    elif kind.startswith("lognormal"):
        s = float(kind[len("lognormal"):] or 1.0)
        w = rng.lognormal(0.0, s, size=k)
        w /= w.sum()
    else:
        raise ValueError(kind)
    sizes = np.maximum(np.floor(w * n_cells).astype(np.int64), min_size)
    # fix the total, trimming the largest subjects
    diff = int(sizes.sum() - n_cells)
    order = np.argsort(-sizes)
    i = 0
    while diff != 0:
        j = order[i % k]
        if diff > 0 and sizes[j] > min_size:
            sizes[j] -= 1
            diff -= 1
        elif diff < 0:
            sizes[j] += 1
            diff += 1
        i += 1
    return sizes


def generate(name, n_genes, n_cells, k, p, sizes="equal", beta0=("normal", -3.0, 1.5),
             seed=0, design_kind="config2", lib_mu=8.0, lib_sd=0.3):
    rng = np.random.default_rng(seed)
    sz = _subject_sizes(rng, n_cells, k, sizes)
    subject = np.repeat(np.arange(k, dtype=np.uint64), sz)
    # design
    treat_subject = np.zeros(k)
    treat_subject[rng.permutation(k)[: k // 2]] = 1.0
    cols = [np.ones(n_cells), np.repeat(treat_subject, sz)]
    if design_kind == "config1":
        pass
    elif design_kind == "config2":
        cols.append(rng.normal(size=n_cells))
        cols.append((rng.random(n_cells) < 0.3).astype(float))
    elif design_kind == "config3":
        for _ in range(3):
            cols.append(rng.normal(size=n_cells))
        batch = rng.integers(0, 4, size=n_cells)
        for level in (1, 2, 3):
            cols.append((batch == level).astype(float))
    elif design_kind == "categorical":
        # a three-level cell type (two indicator columns) and a rare flag
        # (~2 % of cells): every column but the flag is shared by large groups
        # of a subject's cells, so the zero-count tables apply and the flagged
        # cells stay loose
        ctype = rng.integers(0, 3, size=n_cells)
        cols.append((ctype == 1).astype(float))
        cols.append((ctype == 2).astype(float))
        cols.append((rng.random(n_cells) < 0.02).astype(float))
    elif design_kind == "generic":
        while len(cols) < p:
            cols.append(rng.normal(size=n_cells))
    else:
        raise ValueError(design_kind)
    design = np.column_stack(cols[:p])
    assert design.shape[1] == p
    lib = rng.lognormal(lib_mu, lib_sd, size=n_cells)
    offset = lib / lib.mean()
    log_off = np.log(offset)
    # per-gene truth
    kind, a, b = beta0
    b0 = rng.normal(a, b, size=n_genes) if kind == "normal" else rng.uniform(a, b, size=n_genes)
    betas = np.column_stack([b0] + [rng.normal(0.0, 0.5, size=n_genes) for _ in range(p - 1)])
    sigma2 = rng.uniform(0.05, 0.5, size=n_genes)
    phi = rng.uniform(0.2, 5.0, size=n_genes)
    indptr = np.zeros(n_genes + 1, dtype=np.uint64)
    idx_parts, val_parts = [], []
    chunk = max(1, int(4_000_000 // n_cells))
    for g0 in range(0, n_genes, chunk):
        g1 = min(n_genes, g0 + chunk)
        eta = log_off[None, :] + betas[g0:g1] @ design.T
        bs = rng.normal(size=(g1 - g0, k)) * np.sqrt(sigma2[g0:g1, None])
        eta += np.repeat(bs, sz, axis=1)
        mu = np.exp(np.minimum(eta, 30.0))
        size = 1.0 / phi[g0:g1, None]
        lam = rng.gamma(size, mu / size)
        y = rng.poisson(lam)
        for r in range(g1 - g0):
            nz = np.nonzero(y[r])[0]
            idx_parts.append(nz.astype(np.uint32))
            val_parts.append(y[r, nz].astype(np.float64))
            indptr[g0 + r + 1] = indptr[g0 + r] + len(nz)
    indices = np.concatenate(idx_parts) if idx_parts else np.zeros(0, np.uint32)
    data = np.concatenate(val_parts) if val_parts else np.zeros(0)
    if indptr[-1] >= 2 ** 32:
        raise ValueError("u32 indptr overflow")
    return Workload(name, indptr.astype(np.uint32), indices, data, n_genes, n_cells, subject,
                    np.ascontiguousarray(design), p, offset,
                    dict(beta=betas, sigma2=sigma2, phi=phi, sizes=sz))


def config(i, seed=None, n_genes=None, n_cells=None):
    """The five workload shapes; n_genes / n_cells override for scaled-down
    parity cases."""
    if i == 1:
        return generate("config1", n_genes or 200, n_cells or 5000, 10, 2, "equal",
                        ("normal", -2.0, 1.0), 1 if seed is None else seed, "config1")
    if i == 2:
        return generate("config2", n_genes or 500, n_cells or 100_000, 20, 4, "dirichlet2",
                        ("normal", -3.0, 1.5), 2 if seed is None else seed, "config2")
    if i == 3:
        return generate("config3", n_genes or 2000, n_cells or 200_000, 40, 8, "lognormal1",
                        ("uniform", -6.0, 1.0), 3 if seed is None else seed, "config3")
    if i == 4:
        return generate("config4", n_genes or 5000, n_cells or 100_000, 30, 6, "lognormal1",
                        ("uniform", -6.0, 1.0), 4 if seed is None else seed, "generic")
    if i == 5:
        return generate("config5", n_genes or 20000, n_cells or 500_000, 100, 6, "lognormal1.2",
                        ("uniform", -7.0, 1.0), 5 if seed is None else seed, "generic")
    raise ValueError(i)


def reference_bench(n_genes=200, n_cells=20_000, k=20, p=3, seed=0x5EED1234):
    """The shape of the reference's own NEBULA bench (benches/nebula_bench.rs:
    62-78, make_problem :136-220): balanced subjects in contiguous blocks,
    offset ~ LogNormal(0, 0.4), design [1, subject parity, U(-1, 1)...],
    beta0 ~ U(-2, 1.5), other coefficients ~ U(-0.5, 0.5), per-subject
    frailty ~ Gamma(shape 4, scale 0.25), per-cell noise ~ Gamma(shape 2,
    scale 0.5), counts ~ Poisson(offset * exp(x'beta) * frailty * noise).
    The same distributions drawn with numpy's generator, not the reference's
    random stream."""
    rng = np.random.default_rng(seed)
    sz = np.full(k, n_cells // k, dtype=np.int64)
    sz[-1] += n_cells - sz.sum()
    subject = np.repeat(np.arange(k, dtype=np.uint64), sz)
    offset = rng.lognormal(0.0, 0.4, size=n_cells)
    cols = [np.ones(n_cells), (subject % 2).astype(float)]
    while len(cols) < p:
        cols.append(rng.uniform(-1.0, 1.0, size=n_cells))
    design = np.column_stack(cols)
    indptr = np.zeros(n_genes + 1, dtype=np.uint64)
    idx_parts, val_parts = [], []
    for g in range(n_genes):
        beta = np.concatenate([[rng.uniform(-2.0, 1.5)], rng.uniform(-0.5, 0.5, size=p - 1)])
        frailty = rng.gamma(4.0, 0.25, size=k)
        noise = rng.gamma(2.0, 0.5, size=n_cells)
        mu = offset * np.exp(design @ beta) * np.repeat(frailty, sz) * noise
        y = rng.poisson(np.maximum(mu, 1e-12))
        nz = np.nonzero(y)[0]
        idx_parts.append(nz.astype(np.uint32))
        val_parts.append(y[nz].astype(np.float64))
        indptr[g + 1] = indptr[g] + len(nz)
    return Workload("reference_bench", indptr.astype(np.uint32), np.concatenate(idx_parts),
                    np.concatenate(val_parts), n_genes, n_cells, subject,
                    np.ascontiguousarray(design), p, offset, dict(sizes=sz))
