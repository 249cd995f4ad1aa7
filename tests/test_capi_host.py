"""The C ABI without a GPU: libnebula_hip.so loads and exports every function
include/nebula_hip.h declares; the host BOBYQA stepper retraces nloptr point
for point (src/numeric/bobyqa.rs:2947-2970 fixtures); input validation maps
to the EdgeErrors variants before any device work (src/sc/nebula.rs:3303-3514,
tests/e2e_nebula_gpu.rs:854-903)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import nebula_hip as nh
from fixtures import dense_to_csr, golden8, read_matrix

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "nebula_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set()
    for m in re.finditer(r"^\s*(?:int|void|size_t|double)\s*\**\s*(\w+)\s*\(", text, re.M):
        names.add(m.group(1))
    return names


def test_library_exports_every_declared_function():
    names = declared_functions()
    assert "nebula_sparse_hip" in names and "nh_resident_solve" in names
    lib = C.CDLL(nh.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


PROBLEMS = {
    "rosen": (lambda x: 100.0 * (x[1] - x[0] * x[0]) ** 2 + (1.0 - x[0]) ** 2,
              [-1.2, 1.0], [-2.0, -1.0], [2.0, 3.0], 4),
    "rosen_edge": (lambda x: 100.0 * (x[1] - x[0] * x[0]) ** 2 + (1.0 - x[0]) ** 2,
                   [0.4001, 0.2], [0.4, 0.2], [2.0, 2.0], 4),
    "corner": (lambda x: (x[0] - 3.0) ** 2 + 10.0 * (x[1] + 1.0) ** 2 + x[0] * x[1],
               [1.0, 1.0], [0.0, 0.0], [2.0, 2.0], 4),
    "roundoff": (lambda x: (x[0] + x[1] - 1.0) ** 2, [0.0, 0.0], [-2.0, -2.0], [2.0, 2.0], -4),
    "chain3": (lambda x: (x[0] - 1.0) ** 2 + 100.0 * (x[1] - x[0] * x[0]) ** 2 + (x[2] - x[1]) ** 2
               + 0.5 * (x[2] * x[2]) ** 2, [0.5, 2.0, -1.0], [-3.0] * 3, [3.0] * 3, 4),
}


def ripple(x):
    u = 10000.0 * x[0] + 30000.0 * x[1]
    t = abs(u - 2.0 * np.floor((u + 1.0) / 2.0))
    return (x[0] - 0.25) ** 2 + 2.0 * (x[1] + 0.125) ** 2 + 0.5 * x[0] * x[1] + 0.5 * (t * t)


PROBLEMS["rescue"] = (ripple, [-1.0, 0.25], [-2.0, -2.0], [2.0, 2.0], 4)


@pytest.mark.parametrize("tag", list(PROBLEMS))
def test_host_bobyqa_retraces_nloptr(tag):
    """Every point the host stepper asks for, and its status, equal nloptr's."""
    f, x0, lo, hi, status = PROBLEMS[tag]
    want = read_matrix(f"bobyqa_{tag}.csv")
    trace = []

    def obj(x):
        v = float(f(x))
        trace.append(list(x) + [v])
        return v

    x, fx, evals, st = nh.bobyqa(obj, x0, lo, hi)
    assert st == status
    got = np.array(trace)
    assert got.shape == want.shape
    assert np.array_equal(got, want)
    assert evals == len(want)


def _golden_inputs():
    counts, subject, design, offset = golden8()
    ip, ix, dv = dense_to_csr(counts)
    return nh.CompressedSparse(dv, ix, ip, counts.shape), subject, design, offset


def test_reml_is_rejected():
    cs, subject, design, offset = _golden_inputs()
    with pytest.raises(nh.InvalidArgument):
        nh.nebula_sparse_hip(cs, subject, design, 3, offset, nh.NebulaParams(reml=True))


def test_more_than_eight_columns_rejected():
    cs, subject, design, offset = _golden_inputs()
    rng = np.random.default_rng(0)
    wide = np.column_stack([design] + [rng.normal(size=design.shape[0]) for _ in range(6)])
    with pytest.raises(nh.InvalidArgument):
        nh.nebula_sparse_hip(cs, subject, wide, 9, offset)


def test_split_subject_rejected():
    cs, subject, design, offset = _golden_inputs()
    s = subject.copy()
    s[0], s[-1] = s[-1], s[0]
    with pytest.raises(nh.SubjectsNotContiguous):
        nh.nebula_sparse_hip(cs, s, design, 3, offset)


def test_missing_intercept_rejected():
    cs, subject, design, offset = _golden_inputs()
    d = design.copy()
    d[:, 0] = np.arange(d.shape[0], dtype=float)
    with pytest.raises(nh.MissingIntercept):
        nh.nebula_sparse_hip(cs, subject, d, 3, offset)


def test_zero_column_rejected():
    cs, subject, design, offset = _golden_inputs()
    d = design.copy()
    d[:, 2] = 0.0
    with pytest.raises(nh.InvalidArgument):
        nh.nebula_sparse_hip(cs, subject, d, 3, offset)


def test_single_subject_rejected():
    cs, subject, design, offset = _golden_inputs()
    with pytest.raises(nh.TooFewSubjects):
        nh.nebula_sparse_hip(cs, np.zeros_like(subject), design, 3, offset)


def test_negative_count_rejected():
    cs, subject, design, offset = _golden_inputs()
    bad = nh.CompressedSparse(cs.data.copy(), cs.indices, cs.indptr, cs.shape)
    bad.data[3] = -1.0
    with pytest.raises(nh.InvalidArgument):
        nh.nebula_sparse_hip(bad, subject, design, 3, offset)


def test_malformed_csr_rejected():
    cs, subject, design, offset = _golden_inputs()
    ix = cs.indices.copy()
    ix[0] = 10_000
    with pytest.raises(nh.MalformedSparse):
        nh.nebula_sparse_hip(nh.CompressedSparse(cs.data, ix, cs.indptr, cs.shape), subject,
                             design, 3, offset)
    with pytest.raises(nh.MalformedSparse):
        nh.nebula_sparse_hip(nh.CompressedSparse(cs.data, cs.indices, cs.indptr, cs.shape,
                                                 format="csc"), subject, design, 3, offset)


def test_first_offending_gene_decides_and_structure_precedes_values():
    """Genes are checked in parallel; the report is the one a gene-by-gene
    scan of the structure followed by a pass over the values gives: a bad
    count in the first gene loses to a bad index in a later gene, and of two
    structure errors the earlier gene's is named."""
    cs, subject, design, offset = _golden_inputs()
    ip = cs.indptr
    data = cs.data.copy()
    data[ip[0]] = -1.0  # gene 0: a negative count
    ix = cs.indices.copy()
    ix[ip[3]] = 10_000  # gene 3: a cell out of range
    j = ip[5] + 1
    ix[j] = ix[j - 1]   # gene 5: indices not increasing
    with pytest.raises(nh.MalformedSparse, match="out of range"):
        nh.nebula_sparse_hip(nh.CompressedSparse(data, ix, ip, cs.shape), subject, design, 3,
                             offset)
    ix2 = cs.indices.copy()
    ix2[j] = ix2[j - 1]
    with pytest.raises(nh.MalformedSparse, match="increase"):
        nh.nebula_sparse_hip(nh.CompressedSparse(data, ix2, ip, cs.shape), subject, design, 3,
                             offset)
    with pytest.raises(nh.InvalidArgument, match="non-negative"):
        nh.nebula_sparse_hip(nh.CompressedSparse(data, cs.indices, ip, cs.shape), subject,
                             design, 3, offset)


def test_invalid_bounds_and_tolerances():
    cs, subject, design, offset = _golden_inputs()
    with pytest.raises(nh.InvalidBounds):
        nh.nebula_sparse_hip(cs, subject, design, 3, offset, nh.NebulaParams(max=(1e-5, 1000.0)))
    with pytest.raises(nh.MustBePositive):
        nh.nebula_sparse_hip(cs, subject, design, 3, offset, nh.NebulaParams(min=(0.0, 1e-4)))
    with pytest.raises(nh.MustBePositive):
        nh.nebula_sparse_hip(cs, subject, design, 3, offset, nh.NebulaParams(eps=0.0))
    with pytest.raises(nh.MustBePositive):
        nh.nebula_sparse_hip(cs, subject, design, 3, -np.asarray(offset))


def test_shape_and_length_mismatch():
    cs, subject, design, offset = _golden_inputs()
    with pytest.raises(nh.ShapeMismatch):
        nh.nebula_sparse_hip(cs, subject, design[:-1], 3, offset)
    with pytest.raises(nh.LengthMismatch):
        nh.nebula_sparse_hip(cs, subject[:-1], design, 3, offset)


def test_everything_filtered():
    cs, subject, design, offset = _golden_inputs()
    with pytest.raises(nh.NoGenesAfterFiltering):
        nh.nebula_sparse_hip(cs, subject, design, 3, offset, nh.NebulaParams(cpc=1e9))


def test_f64_math_selftest_argument_checks():
    """nh_selftest_f64_math: an empty input is a no-op and an unknown function
    is rejected, both before any device call."""
    assert nh.selftest_f64_math(np.zeros(0), 0).size == 0
    with pytest.raises(nh.InvalidArgument):
        nh.selftest_f64_math(np.ones(3), 7)


def test_resident_submit_and_collect_reject_a_null_handle():
    """nh_resident_submit / nh_resident_collect check the handle before any
    device call."""
    L = nh.lib()
    err = C.create_string_buffer(256)
    assert L.nh_resident_submit(None, None, 0, 1e-6, 0, err, 256) == nh.InvalidArgument.code
    assert b"null handle" in err.value
    assert L.nh_resident_collect(None, None, 0, err, 256) == nh.InvalidArgument.code


def _upload(gene_ptr, cells, counts, n_cells=12):
    """nh_resident_upload through ctypes with raw arrays (None = NULL)."""
    L = nh.lib()
    design = np.ones(n_cells, dtype=np.float64)
    lo = np.zeros(n_cells, dtype=np.float64)
    ss = np.array([0, 6, n_cells], dtype=np.uint64)
    gp = np.ascontiguousarray(gene_ptr, dtype=np.uint64)
    ce = None if cells is None else np.ascontiguousarray(cells, dtype=np.uint32)
    y = None if counts is None else np.ascontiguousarray(counts, dtype=np.float64)
    h = C.c_void_p()
    err = C.create_string_buffer(256)
    rc = L.nh_resident_upload(n_cells, 1, 2, design.ctypes.data_as(C.POINTER(C.c_double)),
                              lo.ctypes.data_as(C.POINTER(C.c_double)),
                              ss.ctypes.data_as(C.POINTER(C.c_uint64)), len(gp) - 1,
                              gp.ctypes.data_as(C.POINTER(C.c_uint64)),
                              None if ce is None else ce.ctypes.data_as(C.POINTER(C.c_uint32)),
                              None if y is None else y.ctypes.data_as(C.POINTER(C.c_double)),
                              0, C.byref(h), err, 256)
    assert not h.value, "no handle may be made for malformed input"
    return rc, err.value.decode()


def test_resident_upload_rejects_malformed_gene_ptr_before_any_device_call():
    """gene_ptr must start at 0 and never decrease, and cells / counts may be
    NULL only when no gene stores an entry: each is an error code, not a
    device read out of bounds."""
    cells = np.array([0, 3, 5, 7], dtype=np.uint32)
    y = np.ones(4)
    rc, msg = _upload([1, 4], cells, y)
    assert rc == nh.MalformedSparse.code and "start at zero" in msg
    rc, msg = _upload([0, 3, 1, 4], cells, y)
    assert rc == nh.MalformedSparse.code and "non-decreasing" in msg
    rc, msg = _upload([0, 4], None, y)
    assert rc == nh.InvalidArgument.code and "null" in msg
    rc, msg = _upload([0, 4], cells, None)
    assert rc == nh.InvalidArgument.code and "null" in msg


def test_null_indices_or_data_rejected():
    """nebula_sparse_hip with stored entries but NULL indices or data."""
    L = nh.lib()
    cs, subject, design, offset = _golden_inputs()
    for which in ("indices", "data"):
        fit = nh._Fit()
        err = C.create_string_buffer(256)
        ip = np.ascontiguousarray(cs.indptr, dtype=np.uint32)
        ix = np.ascontiguousarray(cs.indices, dtype=np.uint32)
        dv = np.ascontiguousarray(cs.data, dtype=np.float64)
        sid = np.ascontiguousarray(subject, dtype=np.uint64)
        d = np.ascontiguousarray(design, dtype=np.float64)
        rc = L.nebula_sparse_hip(cs.shape[0], cs.shape[1], ip.ctypes.data_as(C.POINTER(C.c_uint32)),
                                 None if which == "indices" else ix.ctypes.data_as(C.POINTER(C.c_uint32)),
                                 None if which == "data" else dv.ctypes.data_as(C.POINTER(C.c_double)),
                                 sid.ctypes.data_as(C.POINTER(C.c_uint64)),
                                 d.ctypes.data_as(C.POINTER(C.c_double)), 3, None, None,
                                 C.byref(fit), err, 256)
        assert rc == nh.InvalidArgument.code, (which, rc, err.value)
        assert b"null" in err.value
