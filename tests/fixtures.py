"""Loaders for the committed golden data in tests/golden (copied from the
reference's tests/data by tests/golden/fetch_fixtures.py)."""
import csv
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def read_table(name):
    with open(os.path.join(GOLDEN, name)) as fh:
        rows = list(csv.reader(fh))
    header = [h.strip().strip('"') for h in rows[0]]
    data = np.array([[float(v.strip().strip('"')) for v in r] for r in rows[1:] if r],
                    dtype=np.float64)
    return header, data


def read_matrix(name):
    with open(os.path.join(GOLDEN, name)) as fh:
        rows = [r for r in csv.reader(fh) if r]
    return np.array([[float(v) for v in r] for r in rows], dtype=np.float64)


def dense_to_csr(counts):
    """Gene-major dense counts -> (indptr u32, indices u32, data f64), zeros
    dropped, as edge-rs's build_csr (src/sc/nebula.rs:2170-2205)."""
    counts = np.asarray(counts, dtype=np.float64)
    indptr = [0]
    indices = []
    data = []
    for row in counts:
        nz = np.nonzero(row > 0)[0]
        indices.extend(nz.tolist())
        data.extend(row[nz].tolist())
        indptr.append(len(data))
    return (np.array(indptr, dtype=np.uint32), np.array(indices, dtype=np.uint32),
            np.array(data, dtype=np.float64))


def golden8():
    """The 8 x 150 fixture (src/sc/nebula.rs:2378-2396): counts gene-major,
    subject per cell, design [1, grp, cov2], offsets."""
    counts = read_matrix("nebula_counts.csv")
    header, d = read_table("nebula_design.csv")
    subject = d[:, 0].astype(np.uint64)
    design = np.column_stack([np.ones(len(d)), d[:, 1], d[:, 2]])
    offset = d[:, 3].copy()
    return counts, subject, design, offset


SC_DATASETS = {
    # tag: (method, has_offset, n_coef, design_file) -- tests/e2e_nebula.rs:312-401
    # without sc_xwide (13 columns, beyond the device fit's 8)
    "sc": ("LN", True, 3, False),
    "sc_small": ("LN", True, 3, False),
    "sc_blocks": ("LN", True, 3, True),
    "sc_cat_ln": ("LN", True, 4, True),
    "sc_k2": ("HL", True, 3, True),
    "sc_intercept": ("HL", False, 1, True),
    "sc_wide": ("HL", True, 8, True),
    "sc_high": ("HL", True, 3, True),
    "sc_cat_hl": ("HL", True, 4, True),
    "sc_cat_sparse": ("HL", True, 3, True),
}


def load_sc(tag):
    """One R-parity dataset (tests/e2e_nebula.rs:400-450 loading rules)."""
    method, has_offset, n_coef, design_file = SC_DATASETS[tag]
    header, counts = read_table(f"{tag}_counts.csv")
    mh, meta = read_table(f"{tag}_meta.csv")
    col = {h: i for i, h in enumerate(mh)}
    subject = (meta[:, col["subject"]] - 1).astype(np.uint64)
    if design_file:
        _, design = read_table(f"{tag}_design.csv")
    else:
        design = np.column_stack([np.ones(len(meta)), meta[:, col["grp"]], meta[:, col["cov2"]]])
    offset = meta[:, col["offset"]].copy() if has_offset else None
    nh, neb = read_table(f"{tag}_nebula.csv")
    ch, cov = read_table(f"{tag}_covariance.csv")
    return dict(counts=counts, subject=subject, design=design, offset=offset, method=method,
                n_coef=n_coef, expected=(nh, neb), covariance=cov)


def read_scalars():
    """tests/golden/sc_scalars.csv as {(scenario, name): value}."""
    out = {}
    with open(os.path.join(GOLDEN, "sc_scalars.csv")) as fh:
        for r in list(csv.reader(fh))[1:]:
            if r:
                out[(r[0], r[1])] = float(r[2])
    return out


def n_pinned(tag):
    """Genes whose subject variance finishes on its lower bound, per dataset
    (tests/e2e_nebula.rs:786-795: recorded 18 for sc and 40 for sc_small, R's
    n_pinned scalar for the others)."""
    if tag == "sc":
        return 18
    if tag == "sc_small":
        return 40
    return int(read_scalars()[(tag, "n_pinned")])


def repack_r_cov(row, n):
    """R's lower.tri column-major -> edge-rs packed upper (tests/e2e_nebula.rs:515-526)."""
    out = np.zeros(n * (n + 1) // 2)
    kk = 0
    for j in range(n):
        for i in range(j, n):
            out[i * (i + 1) // 2 + j] = row[kk]
            kk += 1
    return out


def assert_close(got, want, rel, eps=0.0, label=""):
    """edge-rs's assert_close (tests/common/mod.rs:152-240): a value passes when
    its absolute difference is within eps or its relative difference,
    normalised by max(|got|, |want|), is within rel; NaN matches NaN."""
    got = np.asarray(got, dtype=np.float64).ravel()
    want = np.asarray(want, dtype=np.float64).ravel()
    assert got.shape == want.shape and got.size > 0, f"{label}: shape {got.shape} vs {want.shape}"
    both_nan = np.isnan(got) & np.isnan(want)
    finite = np.isfinite(got) & np.isfinite(want)
    absd = np.where(finite, np.abs(got - want), np.inf)
    scale = np.maximum(np.abs(got), np.abs(want))
    with np.errstate(invalid="ignore", divide="ignore"):
        reld = np.where(got == want, 0.0, np.where(finite, absd / np.where(scale == 0, 1, scale),
                                                  np.inf))
    ok = both_nan | (absd <= eps) | (reld <= rel)
    if not ok.all():
        i = int(np.argmax(np.where(ok, -1, reld)))
        raise AssertionError(f"{label}: {int((~ok).sum())} of {got.size} outside tolerance "
                             f"(rel {rel:.1e}, eps {eps:.1e}); worst idx {i} got {got[i]!r} "
                             f"want {want[i]!r} rel {reld[i]:.3e}")
    return float(np.max(np.where(both_nan, 0.0, reld)))
