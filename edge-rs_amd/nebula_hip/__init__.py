"""Python mirror of edge-rs's NEBULA stage-two interface on the HIP library.

    from nebula_hip import nebula_sparse_hip, NebulaParams, CompressedSparse

Same names, argument meaning and errors as edge-rs's ``nebula_sparse`` /
``nebula_sparse_gpu`` (src/sc/nebula.rs:434-445, src/gpu/stage_two.rs:233-265)
and ``ResidentBatch`` (src/gpu/nebula_gpu.rs:195-648).  Every call goes through
the C ABI of libnebula_hip.so (include/nebula_hip.h); there is no host
fallback: if the library is missing, loading raises.
"""
import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# NH_LIBRARY selects another build of the library (the sanitizer build of
# edge-rs_amd/build.py --sanitize); the in-tree build otherwise
LIB_PATH = os.environ.get("NH_LIBRARY") or os.path.join(HERE, "libnebula_hip.so")

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_u64p = C.POINTER(C.c_uint64)
_u32p = C.POINTER(C.c_uint32)

ENGINE_PLAIN = 0
ENGINE_BATCHED = 1
ENGINE_LEARNED = 2
WEIGHTS_PATH = os.path.join(HERE, "warmstart_weights.npy")


# ---------------------------------------------------------------------------
# Errors: one class per EdgeErrors variant on this path (src/errors.rs:13-223).
class EdgeError(RuntimeError):
    code = -1

    def __init__(self, message=""):
        super().__init__(message)


def _err(name, code):
    return type(name, (EdgeError,), {"code": code})


EmptyCounts = _err("EmptyCounts", 1)
InvalidArgument = _err("InvalidArgument", 2)
MustBePositive = _err("MustBePositive", 3)
LengthMismatch = _err("LengthMismatch", 4)
ShapeMismatch = _err("ShapeMismatch", 5)
MalformedSparse = _err("MalformedSparse", 6)
SubjectsNotContiguous = _err("SubjectsNotContiguous", 7)
TooFewSubjects = _err("TooFewSubjects", 8)
MissingIntercept = _err("MissingIntercept", 9)
NoGenesAfterFiltering = _err("NoGenesAfterFiltering", 10)
InvalidBounds = _err("InvalidBounds", 11)
InvalidDispersion = _err("InvalidDispersion", 12)
Gpu = _err("Gpu", 13)
CoefOutOfRange = _err("CoefOutOfRange", 14)
_BY_CODE = {c.code: c for c in (EmptyCounts, InvalidArgument, MustBePositive, LengthMismatch,
                                ShapeMismatch, MalformedSparse, SubjectsNotContiguous,
                                TooFewSubjects, MissingIntercept, NoGenesAfterFiltering,
                                InvalidBounds, InvalidDispersion, Gpu, CoefOutOfRange)}


def _raise(code, err):
    msg = err.value.decode(errors="replace") if hasattr(err, "value") else str(err)
    raise _BY_CODE.get(code, EdgeError)(msg)


# ---------------------------------------------------------------------------
class _Params(C.Structure):
    _fields_ = [("min_subject", C.c_double), ("min_cell", C.c_double),
                ("max_subject", C.c_double), ("max_cell", C.c_double),
                ("method", C.c_int32), ("reml", C.c_int32),
                ("cutoff_cell", C.c_double), ("kappa", C.c_double),
                ("cpc", C.c_double), ("mincp", C.c_uint64),
                ("min_subjects", C.c_uint64), ("eps", C.c_double)]


class _Fit(C.Structure):
    _fields_ = [("n_genes", C.c_size_t), ("n_coef", C.c_size_t),
                ("coefficients", _dp), ("covariance", _dp), ("se", _dp),
                ("subject_overdispersion", _dp), ("cell_overdispersion", _dp),
                ("convergence", C.POINTER(C.c_int32)),
                ("sigma_at_bound", C.POINTER(C.c_uint8)),
                ("gene_index", _u64p)]


class _Options(C.Structure):
    _fields_ = [("engine", C.c_int32), ("learned", C.c_int32),
                ("mlp_weights", _fp), ("wg_size", C.c_int32),
                ("device", C.c_int32), ("host_threads", C.c_int32),
                ("learned_step", C.c_double), ("learned_xtol", C.c_double),
                ("zero_tables", C.c_int32)]


class RunStats(C.Structure):
    """nebula_run_stats_t: per-phase wall clock and device counters."""
    _fields_ = [("t_setup", C.c_double), ("t_stage_one", C.c_double),
                ("t_stage_two", C.c_double), ("t_stage_three", C.c_double),
                ("t_total", C.c_double), ("t_upload", C.c_double),
                ("genes_searched", C.c_uint64), ("evaluations", C.c_uint64),
                ("rounds", C.c_uint64), ("sweeps_f32", C.c_uint64),
                ("sweeps_f64", C.c_uint64), ("cell_sweeps", C.c_uint64),
                ("nnz_sweeps", C.c_uint64), ("warm_fallbacks", C.c_uint64),
                ("fit_kernel_ms", C.c_double), ("fit_kernel_launches", C.c_uint64),
                ("cell_sweeps_f64", C.c_uint64), ("nnz_sweeps_f64", C.c_uint64),
                ("t_plan", C.c_double), ("zero_groups", C.c_uint64),
                ("t_tables", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class _Request(C.Structure):
    _fields_ = [("gene", C.c_uint32), ("ord", C.c_uint32), ("subject", C.c_double),
                ("cell", C.c_double), ("beta_init", _dp), ("log_w_init", _dp)]


class _Reply(C.Structure):
    _fields_ = [("log_likelihood", C.c_double), ("log_likelihood_prev", C.c_double),
                ("log_det", C.c_double), ("second_order", C.c_double),
                ("convergence", C.c_int32), ("iterations", C.c_int32),
                ("backtracks", C.c_int32), ("finish_steps", C.c_int32),
                ("beta", _dp), ("log_w", _dp), ("information", _dp)]


OBJECTIVE_FN = C.CFUNCTYPE(C.c_double, _dp, C.c_void_p)

_lib = None


def lib():
    """Loads libnebula_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libnebula_hip.so not built at {LIB_PATH}; "
                               "run python edge-rs_amd/build.py")
        L = C.CDLL(LIB_PATH)
        L.nebula_params_default.argtypes = [C.POINTER(_Params)]
        L.nebula_hip_options_default.argtypes = [C.POINTER(_Options)]
        L.nebula_sparse_hip_ex.restype = C.c_int
        L.nebula_sparse_hip_ex.argtypes = [C.c_size_t, C.c_size_t, _u32p, _u32p, _dp, _u64p, _dp,
                                           C.c_size_t, _dp, C.POINTER(_Params),
                                           C.POINTER(_Options), C.POINTER(_Fit),
                                           C.POINTER(RunStats), C.c_char_p, C.c_size_t]
        L.nebula_sparse_hip.restype = C.c_int
        L.nebula_sparse_hip.argtypes = [C.c_size_t, C.c_size_t, _u32p, _u32p, _dp, _u64p, _dp,
                                        C.c_size_t, _dp, C.POINTER(_Params), C.POINTER(_Fit),
                                        C.c_char_p, C.c_size_t]
        L.nebula_fit_free.argtypes = [C.POINTER(_Fit)]
        L.nh_resident_upload.restype = C.c_int
        L.nh_resident_upload.argtypes = [C.c_size_t, C.c_size_t, C.c_size_t, _dp, _dp, _u64p,
                                         C.c_size_t, _u64p, _u32p, _dp, C.c_int32,
                                         C.POINTER(C.c_void_p), C.c_char_p, C.c_size_t]
        L.nh_resident_solve.restype = C.c_int
        L.nh_resident_solve.argtypes = [C.c_void_p, C.POINTER(_Request), C.c_size_t, C.c_double,
                                        C.c_int32, C.POINTER(_Reply), C.c_char_p, C.c_size_t]
        L.nh_resident_submit.restype = C.c_int
        L.nh_resident_submit.argtypes = [C.c_void_p, C.POINTER(_Request), C.c_size_t,
                                         C.c_double, C.c_int32, C.c_char_p, C.c_size_t]
        L.nh_resident_collect.restype = C.c_int
        L.nh_resident_collect.argtypes = [C.c_void_p, C.POINTER(_Reply), C.c_size_t, C.c_char_p,
                                          C.c_size_t]
        L.nh_resident_free.argtypes = [C.c_void_p]
        L.nh_bobyqa.restype = C.c_int
        L.nh_bobyqa.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_int, OBJECTIVE_FN,
                                C.c_void_p, _dp, _dp, C.POINTER(C.c_int)]
        L.nh_mlp_param_count.restype = C.c_size_t
        L.nh_mlp_param_count.argtypes = []
        L.nh_mlp_train.restype = C.c_int
        L.nh_mlp_train.argtypes = [_fp, _fp, C.c_size_t, C.c_int32, C.c_int32, C.c_float,
                                   C.c_uint64, C.c_int32, _fp, C.POINTER(C.c_float),
                                   C.c_char_p, C.c_size_t]
        L.nh_mlp_loss_grad.restype = C.c_int
        L.nh_mlp_loss_grad.argtypes = [_fp, _fp, _fp, C.c_size_t, C.c_int32,
                                       C.POINTER(C.c_float), _fp, C.c_char_p, C.c_size_t]
        L.nh_mlp_predict.restype = C.c_int
        L.nh_mlp_predict.argtypes = [_fp, _fp, C.c_size_t, C.c_int32, _fp, C.c_char_p,
                                     C.c_size_t]
        L.nh_bobyqa_device_trace.restype = C.c_int
        L.nh_bobyqa_device_trace.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_int,
                                             C.POINTER(C.c_int), C.c_char_p, C.c_size_t]
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


# ---------------------------------------------------------------------------
@dataclass
class NebulaParams:
    """NebulaParams (src/sc/nebula.rs:207-258) with the R package defaults."""
    min: tuple = (1e-4, 1e-4)
    max: tuple = (10.0, 1000.0)
    method: str = "LN"
    cutoff_cell: float = 20.0
    kappa: float = 800.0
    cpc: float = 0.005
    mincp: int = 5
    min_subjects: int = 0
    reml: bool = False
    eps: float = 1e-6

    def _c(self):
        return _Params(self.min[0], self.min[1], self.max[0], self.max[1],
                       1 if self.method == "HL" else 0, int(self.reml), self.cutoff_cell,
                       self.kappa, self.cpc, self.mincp, self.min_subjects, self.eps)


@dataclass
class CompressedSparse:
    """Gene-major CSR counts (CompressedSparse<f64>, src/utils/sparse.rs:58-69)."""
    data: np.ndarray
    indices: np.ndarray
    indptr: np.ndarray
    shape: tuple
    format: str = "csr"

    @classmethod
    def from_dense(cls, counts):
        counts = np.asarray(counts, dtype=np.float64)
        indptr = [0]
        indices, data = [], []
        for row in counts:
            nz = np.nonzero(row != 0)[0]
            indices.extend(nz.tolist())
            data.extend(row[nz].tolist())
            indptr.append(len(data))
        return cls(np.array(data, dtype=np.float64), np.array(indices, dtype=np.uint32),
                   np.array(indptr, dtype=np.uint32), counts.shape)


@dataclass
class NebulaFit:
    """NebulaFit (src/sc/nebula.rs:301-344)."""
    coefficients: np.ndarray
    covariance: np.ndarray
    se: np.ndarray
    subject_overdispersion: np.ndarray
    cell_overdispersion: np.ndarray
    convergence: np.ndarray
    sigma_at_bound: np.ndarray
    gene_index: np.ndarray
    n_coef: int
    stats: dict = field(default_factory=dict)


def nebula_sparse_hip(counts: CompressedSparse, subject_id, design, n_coef: int,
                      offset=None, params: Optional[NebulaParams] = None,
                      engine: int = ENGINE_BATCHED, learned: bool = False,
                      mlp_weights=None, wg_size: int = 0, device: int = 0,
                      host_threads: int = 0, learned_step: float = 0.0,
                      learned_xtol: float = 0.0, zero_tables: int = 1) -> NebulaFit:
    """NEBULA with stages one to three on the GPU (edge-rs nebula_sparse_gpu)."""
    if counts.format != "csr":
        raise MalformedSparse("nebula_sparse needs gene-major counts in CSR form.")
    n_genes, n_cells = counts.shape
    design = np.ascontiguousarray(design, dtype=np.float64)
    if design.size != n_cells * n_coef:
        raise ShapeMismatch(f"design has {design.size} values, expected {n_cells}x{n_coef}")
    subject_id = np.ascontiguousarray(subject_id, dtype=np.uint64)
    if subject_id.size != n_cells:
        raise LengthMismatch(f"subject_id has {subject_id.size} entries, expected {n_cells}")
    if offset is not None and np.asarray(offset).size != n_cells:
        raise LengthMismatch(f"offset has {np.asarray(offset).size} entries, expected {n_cells}")
    indptr = np.ascontiguousarray(counts.indptr, dtype=np.uint32)
    indices = np.ascontiguousarray(counts.indices, dtype=np.uint32)
    data = np.ascontiguousarray(counts.data, dtype=np.float64)
    off = None
    offp = None
    if offset is not None:
        off, offp = _d(offset)
    p = (params or NebulaParams())._c()
    o = _Options()
    lib().nebula_hip_options_default(C.byref(o))
    o.engine = engine
    o.learned = int(learned)
    w = None
    if mlp_weights is not None:
        w = np.ascontiguousarray(mlp_weights, dtype=np.float32)
        o.mlp_weights = w.ctypes.data_as(_fp)
    o.wg_size = wg_size
    o.device = device
    o.host_threads = host_threads
    o.learned_step = learned_step
    o.learned_xtol = learned_xtol
    o.zero_tables = int(zero_tables)
    fit = _Fit()
    st = RunStats()
    err = C.create_string_buffer(1024)
    rc = lib().nebula_sparse_hip_ex(n_genes, n_cells, indptr.ctypes.data_as(_u32p),
                                    indices.ctypes.data_as(_u32p), data.ctypes.data_as(_dp),
                                    subject_id.ctypes.data_as(_u64p), design.ctypes.data_as(_dp),
                                    n_coef, offp, C.byref(p), C.byref(o), C.byref(fit),
                                    C.byref(st), err, 1024)
    if rc != 0:
        _raise(rc, err)
    g, nb = fit.n_genes, fit.n_coef
    pk = nb * (nb + 1) // 2
    arr = np.ctypeslib.as_array
    out = NebulaFit(
        coefficients=arr(fit.coefficients, (g * nb,)).reshape(g, nb).copy(),
        covariance=arr(fit.covariance, (g * pk,)).reshape(g, pk).copy(),
        se=arr(fit.se, (g * nb,)).reshape(g, nb).copy(),
        subject_overdispersion=arr(fit.subject_overdispersion, (g,)).copy(),
        cell_overdispersion=arr(fit.cell_overdispersion, (g,)).copy(),
        convergence=arr(fit.convergence, (g,)).copy(),
        sigma_at_bound=arr(fit.sigma_at_bound, (g,)).astype(bool),
        gene_index=arr(fit.gene_index, (g,)).astype(np.int64),
        n_coef=nb, stats=st.as_dict())
    lib().nebula_fit_free(C.byref(fit))
    return out


# ---------------------------------------------------------------------------
class ResidentBatch:
    """Device-resident design, offsets, subjects and per-gene counts; batches
    of penalised fits on them (edge-rs ResidentBatch upload / solve)."""

    SOLVE_F32_FINISH = 0
    SOLVE_F64 = 1

    def __init__(self, design, log_offset, subject_start, gene_ptr, cells, counts, device=0):
        self.design, dp = _d(design)
        self.n_cells = len(log_offset)
        self.n_coef = self.design.size // self.n_cells
        self.k = len(subject_start) - 1
        lo, lop = _d(log_offset)
        ss = np.ascontiguousarray(subject_start, dtype=np.uint64)
        gp = np.ascontiguousarray(gene_ptr, dtype=np.uint64)
        ce = np.ascontiguousarray(cells, dtype=np.uint32)
        y, yp = _d(counts)
        self.n_genes = len(gp) - 1
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        rc = lib().nh_resident_upload(self.n_cells, self.n_coef, self.k, dp, lop,
                                      ss.ctypes.data_as(_u64p), self.n_genes,
                                      gp.ctypes.data_as(_u64p), ce.ctypes.data_as(_u32p), yp,
                                      device, C.byref(h), err, 512)
        if rc != 0:
            _raise(rc, err)
        self._h = h

    def solve(self, genes, subject, cell, beta_init, log_w_init=None, ord=1, eps=1e-6,
              mode=SOLVE_F32_FINISH):
        """One penalised fit per request; returns a dict of arrays
        (nh_resident_solve: submit, then collect)."""
        self.submit(genes, subject, cell, beta_init, log_w_init, ord, eps, mode)
        return self.collect()

    def _requests(self, genes, subject, cell, beta_init, log_w_init, ord):
        """nh_request_t array and the numpy arrays its pointers view."""
        n = len(genes)
        nb, k = self.n_coef, self.k
        bi = np.ascontiguousarray(np.broadcast_to(np.asarray(beta_init, dtype=np.float64),
                                                  (n, nb)))
        lw = None
        if log_w_init is not None:
            lw = np.ascontiguousarray(np.broadcast_to(np.asarray(log_w_init, dtype=np.float64),
                                                      (n, k)))
        ords = np.broadcast_to(np.asarray(ord), (n,))
        subject = np.broadcast_to(np.asarray(subject, dtype=np.float64), (n,))
        cell = np.broadcast_to(np.asarray(cell, dtype=np.float64), (n,))
        reqs = (_Request * n)()
        for q in range(n):
            reqs[q].gene = int(genes[q])
            reqs[q].ord = int(ords[q])
            reqs[q].subject = float(subject[q])
            reqs[q].cell = float(cell[q])
            reqs[q].beta_init = bi[q].ctypes.data_as(_dp)
            reqs[q].log_w_init = lw[q].ctypes.data_as(_dp) if lw is not None else None
        return reqs, (bi, lw)

    def _replies(self, n):
        nb, k = self.n_coef, self.k
        arrays = dict(beta=np.zeros((n, nb)), log_w=np.zeros((n, k)),
                      information=np.zeros((n, nb * nb)))
        reps = (_Reply * n)()
        for q in range(n):
            reps[q].beta = arrays["beta"][q].ctypes.data_as(_dp)
            reps[q].log_w = arrays["log_w"][q].ctypes.data_as(_dp)
            reps[q].information = arrays["information"][q].ctypes.data_as(_dp)
        return reps, arrays

    @staticmethod
    def _reply_dict(reps, arrays):
        return dict(
            **arrays,
            log_likelihood=np.array([r.log_likelihood for r in reps]),
            log_likelihood_prev=np.array([r.log_likelihood_prev for r in reps]),
            log_det=np.array([r.log_det for r in reps]),
            second_order=np.array([r.second_order for r in reps]),
            iterations=np.array([r.iterations for r in reps]),
            backtracks=np.array([r.backtracks for r in reps]),
            finish_steps=np.array([r.finish_steps for r in reps]),
            convergence=np.array([r.convergence for r in reps]))

    def submit(self, genes, subject, cell, beta_init, log_w_init=None, ord=1, eps=1e-6,
               mode=SOLVE_F32_FINISH):
        """Launches the fits and returns at once (nh_resident_submit); one
        launch may be in flight, and collect() returns its replies."""
        reqs, _keep = self._requests(genes, subject, cell, beta_init, log_w_init, ord)
        err = C.create_string_buffer(512)
        rc = lib().nh_resident_submit(self._h, reqs, len(genes), eps, mode, err, 512)
        if rc != 0:
            _raise(rc, err)
        self._pending = len(genes)

    def collect(self, n=None):
        """Waits for the launch in flight and returns its replies as a dict of
        arrays (nh_resident_collect); n defaults to the submitted count."""
        n = getattr(self, "_pending", 0) if n is None else n
        reps, arrays = self._replies(n)
        err = C.create_string_buffer(512)
        rc = lib().nh_resident_collect(self._h, reps, n, err, 512)
        if rc != 0:
            _raise(rc, err)
        self._pending = 0
        return self._reply_dict(reps, arrays)

    def solve_noalloc(self, genes, subject, cell, beta_init, log_w_init=None, ord=1, eps=1e-6,
                      mode=SOLVE_F32_FINISH):
        """solve() through nh_resident_solve_noalloc: the same fits with every
        device buffer in the library's fixed per-device scratch array."""
        reqs, _keep = self._requests(genes, subject, cell, beta_init, log_w_init, ord)
        reps, arrays = self._replies(len(genes))
        L = lib()
        L.nh_resident_solve_noalloc.restype = C.c_int
        L.nh_resident_solve_noalloc.argtypes = [C.c_void_p, C.POINTER(_Request), C.c_size_t,
                                                C.c_double, C.c_int32, C.POINTER(_Reply),
                                                C.c_char_p, C.c_size_t]
        err = C.create_string_buffer(512)
        rc = L.nh_resident_solve_noalloc(self._h, reqs, len(genes), eps, mode, reps, err, 512)
        if rc != 0:
            _raise(rc, err)
        return self._reply_dict(reps, arrays)

    def close(self):
        if getattr(self, "_h", None):
            lib().nh_resident_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bobyqa(f, x0, lower, upper, xtol_rel=1e-6, maxeval=1000):
    """The library's host BOBYQA (NLopt 2.7.1 semantics) on a Python objective."""
    n = len(x0)
    x0, x0p = _d(x0)
    lo, lop = _d(lower)
    hi, hip = _d(upper)

    @OBJECTIVE_FN
    def cb(xp, _ctx):
        return float(f(np.ctypeslib.as_array(xp, (n,)).copy()))

    xo = np.zeros(n)
    fo = C.c_double()
    ev = C.c_int()
    status = lib().nh_bobyqa(n, x0p, lop, hip, xtol_rel, maxeval, cb, None,
                             xo.ctypes.data_as(_dp), C.byref(fo), C.byref(ev))
    return xo, fo.value, ev.value, status


def bobyqa_device_trace(problem, x0, lower, upper, max_trace=2000):
    """Runs the device BOBYQA on built-in problem (0 rosen, 1 corner,
    2 ripple, 3 valley, 4 chain3); returns (trace rows, status)."""
    n = len(x0)
    x0, x0p = _d(x0)
    lo, lop = _d(lower)
    hi, hip = _d(upper)
    tr = np.zeros(max_trace * (n + 1))
    ev = C.c_int()
    err = C.create_string_buffer(512)
    status = lib().nh_bobyqa_device_trace(problem, n, x0p, lop, hip, tr.ctypes.data_as(_dp),
                                          max_trace, C.byref(ev), err, 512)
    if status <= -100:
        _raise(-status - 100, err)
    rows = min(ev.value, max_trace)
    return tr[: rows * (n + 1)].reshape(rows, n + 1), status


def selftest_copy(mib=256, reps=1, device=0):
    """A device-to-device copy of mib MiB, reps times (nh_selftest_copy): the
    yardstick for the HBM byte counters."""
    L = lib()
    L.nh_selftest_copy.restype = C.c_int
    L.nh_selftest_copy.argtypes = [C.c_size_t, C.c_int32, C.c_int32, C.c_char_p, C.c_size_t]
    err = C.create_string_buffer(512)
    rc = L.nh_selftest_copy(mib, reps, device, err, 512)
    if rc != 0:
        _raise(rc, err)


def zero_tables_probe(design, log_offset, subject_start, s=(), frm=0, m=3):
    """The library's zero-count tables built on the host (nh_zero_tables_probe):
    None when they do not apply, else a dict of group_row, subject_groups and
    `values`, the sums from `frm` (m of them) at each point of s, shaped
    (n_groups, len(s), m)."""
    design = np.ascontiguousarray(design, dtype=np.float64)
    lo = np.ascontiguousarray(log_offset, dtype=np.float64)
    n = len(lo)
    nb = design.size // n
    ss = np.ascontiguousarray(subject_start, dtype=np.uint64)
    k = len(ss) - 1
    pts = np.ascontiguousarray(np.atleast_1d(np.asarray(s, dtype=np.float64)))
    cap = n * max(1, len(pts)) * m
    out = np.zeros(cap)
    applies = C.c_int32(0)
    n_groups = C.c_uint64(0)
    rows = np.zeros(n, dtype=np.uint64)
    sg = np.zeros(k + 1, dtype=np.uint64)
    L = lib()
    L.nh_zero_tables_probe.restype = C.c_int
    L.nh_zero_tables_probe.argtypes = [_dp, _dp, C.c_size_t, C.c_size_t, _u64p, C.c_size_t, _dp,
                                       C.c_size_t, C.c_int32, C.c_int32, _dp, C.c_size_t,
                                       C.POINTER(C.c_int32), C.POINTER(C.c_uint64), _u64p, _u64p,
                                       C.c_char_p, C.c_size_t]
    err = C.create_string_buffer(512)
    rc = L.nh_zero_tables_probe(design.ctypes.data_as(_dp), lo.ctypes.data_as(_dp), n, nb,
                                ss.ctypes.data_as(_u64p), k, pts.ctypes.data_as(_dp), len(pts),
                                frm, m, out.ctypes.data_as(_dp), cap, C.byref(applies),
                                C.byref(n_groups), rows.ctypes.data_as(_u64p),
                                sg.ctypes.data_as(_u64p), err, 512)
    if rc != 0:
        _raise(rc, err)
    if not applies.value:
        return None
    g = n_groups.value
    return dict(group_row=rows[:g].astype(np.int64), subject_groups=sg.astype(np.int64),
                values=out[: g * len(pts) * m].reshape(g, len(pts), m))


VALUE_GRAD_FN = C.CFUNCTYPE(C.c_double, _dp, _dp, C.c_void_p)


def stage_one_search(fg, start, lower, upper):
    """Stage one's quasi-Newton search (bound restarts around L-BFGS-B) on the
    host against fg(x) -> (f, g) (nh_stage_one_search).  Returns (x, failed,
    evaluations)."""
    n = len(start)

    @VALUE_GRAD_FN
    def cb(xp, gp, _ctx):
        f, g = fg(np.ctypeslib.as_array(xp, (n,)).copy())
        for i in range(n):
            gp[i] = float(g[i])
        return float(f)

    st, stp = _d(start)
    lo, lop = _d(lower)
    hi, hip = _d(upper)
    x = np.zeros(n)
    failed, ev = C.c_int(), C.c_int()
    L = lib()
    L.nh_stage_one_search.restype = C.c_int
    L.nh_stage_one_search.argtypes = [C.c_int, _dp, _dp, _dp, VALUE_GRAD_FN, C.c_void_p, _dp,
                                      C.POINTER(C.c_int), C.POINTER(C.c_int)]
    rc = L.nh_stage_one_search(n, stp, lop, hip, cb, None, x.ctypes.data_as(_dp),
                               C.byref(failed), C.byref(ev))
    if rc != 0:
        _raise(rc, C.create_string_buffer(b"nh_stage_one_search rejected its arguments"))
    return x, bool(failed.value), ev.value


SEARCH_FN = C.CFUNCTYPE(C.c_double, C.c_double, C.c_double, C.c_int, C.c_void_p)
SEARCH_BOTH = 2
SEARCH_SUBJECT = 1


def stage_two_search(f, kind, start, lower, upper, ord=1):
    """The stage-two search control both engines run (nh_stage_two_search),
    on the host against f(subject, cell, order).  kind 2: BOBYQA over both
    components; kind 1: the subject variance alone, cell held at start[1].
    Returns dict(x, found, failed, order, evaluations)."""
    st = np.zeros(2)
    st[: len(start)] = start
    lo = np.zeros(2)
    lo[: len(lower)] = lower
    hi = np.ones(2)
    hi[: len(upper)] = upper

    @SEARCH_FN
    def cb(sub, cell, order, _ctx):
        return float(f(sub, cell, order))

    x = np.zeros(2)
    found, failed, order, ev = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    L = lib()
    L.nh_stage_two_search.restype = C.c_int
    L.nh_stage_two_search.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, SEARCH_FN, C.c_void_p,
                                      _dp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int), C.POINTER(C.c_int)]
    rc = L.nh_stage_two_search(kind, st.ctypes.data_as(_dp), lo.ctypes.data_as(_dp),
                               hi.ctypes.data_as(_dp), ord, cb, None, x.ctypes.data_as(_dp),
                               C.byref(found), C.byref(failed), C.byref(order), C.byref(ev))
    if rc != 0:
        _raise(rc, C.create_string_buffer(b"nh_stage_two_search rejected its arguments"))
    return dict(x=x, found=bool(found.value), failed=bool(failed.value), order=order.value,
                evaluations=ev.value)


def stage_two_search_device(problem, kind, start, lower, upper, ord=1, inject=None,
                            max_trace=4000, device=0):
    """The same control on the device through the persistent kernel's entry
    points, against built-in objective `problem` (nh_stage_two_search_device).
    inject = (from_call, to_call, min_order, kind) with kind 0 +inf, 1 NaN,
    2 a Laplace correction below -1.  Returns the dict of stage_two_search
    plus trace rows (subject, cell, order, value)."""
    st = np.zeros(2)
    st[: len(start)] = start
    lo = np.zeros(2)
    lo[: len(lower)] = lower
    hi = np.ones(2)
    hi[: len(upper)] = upper
    inj = inject or (0, -1, 99, 0)
    tr = np.zeros(max_trace * 4)
    x = np.zeros(2)
    found, failed, order, ev = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    err = C.create_string_buffer(512)
    L = lib()
    L.nh_stage_two_search_device.restype = C.c_int
    L.nh_stage_two_search_device.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, C.c_int,
                                             C.c_int, C.c_int, C.c_int, C.c_int32, _dp, C.c_int,
                                             _dp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                             C.POINTER(C.c_int), C.POINTER(C.c_int),
                                             C.c_char_p, C.c_size_t]
    rc = L.nh_stage_two_search_device(problem, kind, ord, st.ctypes.data_as(_dp),
                                      lo.ctypes.data_as(_dp), hi.ctypes.data_as(_dp), *inj,
                                      device, tr.ctypes.data_as(_dp), max_trace,
                                      x.ctypes.data_as(_dp), C.byref(found), C.byref(failed),
                                      C.byref(order), C.byref(ev), err, 512)
    if rc != 0:
        _raise(rc, err)
    rows = min(ev.value, max_trace)
    return dict(x=x, found=bool(found.value), failed=bool(failed.value), order=order.value,
                evaluations=ev.value, trace=tr[: rows * 4].reshape(rows, 4))


def glm_sc_test(coefficients, covariance, coef=None, contrast=None):
    """Wald test of one coefficient (coef) or a contrast (n_coef weights) per
    gene on a fit's coefficients (genes x n_coef) and packed covariance
    (genes x n_coef (n_coef + 1) / 2), as NebulaFit holds them
    (nh_glm_sc_test, glm_sc_test in src/sc/test.rs:98).  Returns a dict of
    log_fc, se, z, p_value and fdr arrays."""
    b = np.ascontiguousarray(coefficients, dtype=np.float64)
    v = np.ascontiguousarray(covariance, dtype=np.float64)
    n_genes = b.shape[0] if b.ndim == 2 else (1 if b.size else 0)
    n_coef = b.shape[1] if b.ndim == 2 else b.size
    c = None
    cp = None
    if contrast is not None:
        c, cp = _d(contrast)
        if c.size != n_coef:
            raise LengthMismatch(f"contrast: expected {n_coef}, got {c.size}")
    elif coef is None:
        raise InvalidArgument("give coef or contrast")
    out = {k: np.zeros(n_genes) for k in ("log_fc", "se", "z", "p_value", "fdr")}
    err = C.create_string_buffer(512)
    L = lib()
    L.nh_glm_sc_test.restype = C.c_int
    L.nh_glm_sc_test.argtypes = [_dp, C.c_size_t, _dp, C.c_size_t, C.c_size_t, C.c_size_t,
                                 C.c_size_t, _dp, _dp, _dp, _dp, _dp, _dp, C.c_char_p,
                                 C.c_size_t]
    rc = L.nh_glm_sc_test(b.ctypes.data_as(_dp), b.size, v.ctypes.data_as(_dp), v.size,
                          n_genes, n_coef, int(coef or 0), cp,
                          *[out[k].ctypes.data_as(_dp) for k in ("log_fc", "se", "z",
                                                                 "p_value", "fdr")],
                          err, 512)
    if rc != 0:
        _raise(rc, err)
    return out


def sc_residual_df(n_cells, n_coef, n_subjects):
    """n_cells - n_coef - (n_subjects - 1), at least 1 (nh_sc_residual_df,
    sc_residual_df in src/sc/shrink.rs:90)."""
    L = lib()
    L.nh_sc_residual_df.restype = C.c_double
    L.nh_sc_residual_df.argtypes = [C.c_size_t, C.c_size_t, C.c_size_t]
    return float(L.nh_sc_residual_df(int(n_cells), int(n_coef), int(n_subjects)))


@dataclass
class ShrinkParams:
    """ShrinkParams (src/sc/shrink.rs:38-49): Winsorised (robust) prior fit."""
    robust: bool = True


@dataclass
class ShrinkResult:
    """ShrinkResult (src/sc/shrink.rs:52-71): one value per gene, df_prior one
    per usable gene (robust fit) or one."""
    phi_raw: np.ndarray
    phi_post: np.ndarray
    phi_prior: np.ndarray
    df_residual: float
    df_prior: np.ndarray
    dispersion_shrunk: np.ndarray


def shrink_sc_dispersion(dispersion, convergence, covariate=None, df_residual=None,
                         params: Optional[ShrinkParams] = None) -> ShrinkResult:
    """Empirical-Bayes shrinkage of the cell-level overdispersions
    (nh_shrink_sc_dispersion, shrink_sc_dispersion in src/sc/shrink.rs:118):
    phi = 1 / dispersion through limma's squeezeVar on df_residual degrees of
    freedom, trended on `covariate` when given."""
    params = params or ShrinkParams()
    d = np.ascontiguousarray(dispersion, dtype=np.float64).ravel()
    n = d.size
    conv = np.ascontiguousarray(convergence, dtype=np.int32).ravel()
    if conv.size != n:
        raise LengthMismatch(f"convergence: expected {n}, got {conv.size}")
    cov = None
    if covariate is not None:
        cov = np.ascontiguousarray(covariate, dtype=np.float64).ravel()
        if cov.size != n:
            raise LengthMismatch(f"covariate: expected {n}, got {cov.size}")
    if df_residual is None:
        raise InvalidArgument("shrink_sc_dispersion needs df_residual")
    outs = [np.zeros(n) for _ in range(4)]
    df_prior = np.zeros(max(n, 1))
    n_df = C.c_size_t(0)
    err = C.create_string_buffer(512)
    L = lib()
    L.nh_shrink_sc_dispersion.restype = C.c_int
    L.nh_shrink_sc_dispersion.argtypes = [_dp, C.POINTER(C.c_int32), _dp, C.c_size_t,
                                          C.c_double, C.c_int32, _dp, _dp, _dp, _dp, _dp,
                                          C.POINTER(C.c_size_t), C.c_char_p, C.c_size_t]
    rc = L.nh_shrink_sc_dispersion(
        d.ctypes.data_as(_dp), conv.ctypes.data_as(C.POINTER(C.c_int32)),
        cov.ctypes.data_as(_dp) if cov is not None else None, n, float(df_residual),
        int(bool(params.robust)), *[o.ctypes.data_as(_dp) for o in outs],
        df_prior.ctypes.data_as(_dp), C.byref(n_df), err, 512)
    if rc != 0:
        _raise(rc, err)
    return ShrinkResult(phi_raw=outs[0], phi_post=outs[1], phi_prior=outs[2],
                        df_residual=float(df_residual), df_prior=df_prior[:n_df.value].copy(),
                        dispersion_shrunk=outs[3])


def squeeze_var(var, df, covariate=None, robust=False):
    """limma's squeezeVar for one shared df (nh_squeeze_var, squeeze_var in
    src/limma/squeeze_var.rs:286, the legacy family).  Returns (var_post,
    var_prior, df_prior), the priors of length 1 or n."""
    v = np.ascontiguousarray(var, dtype=np.float64).ravel()
    n = v.size
    cov = None
    if covariate is not None:
        cov = np.ascontiguousarray(covariate, dtype=np.float64).ravel()
        if cov.size != n:
            raise LengthMismatch(f"covariate: expected {n}, got {cov.size}")
    post = np.zeros(n)
    prior = np.zeros(max(n, 1))
    dfp = np.zeros(max(n, 1))
    n_prior = C.c_size_t(0)
    n_dfp = C.c_size_t(0)
    err = C.create_string_buffer(512)
    L = lib()
    L.nh_squeeze_var.restype = C.c_int
    L.nh_squeeze_var.argtypes = [_dp, C.c_size_t, C.c_double, _dp, C.c_int32, _dp, _dp,
                                 C.POINTER(C.c_size_t), _dp, C.POINTER(C.c_size_t), C.c_char_p,
                                 C.c_size_t]
    rc = L.nh_squeeze_var(v.ctypes.data_as(_dp), n, float(df),
                          cov.ctypes.data_as(_dp) if cov is not None else None,
                          int(bool(robust)), post.ctypes.data_as(_dp), prior.ctypes.data_as(_dp),
                          C.byref(n_prior), dfp.ctypes.data_as(_dp), C.byref(n_dfp), err, 512)
    if rc != 0:
        _raise(rc, err)
    return post, prior[:n_prior.value].copy(), dfp[:n_dfp.value].copy()


def shrink_selftest(which, x, y=None, a=0.0, b=0.0):
    """The shrinkage's distribution functions on the host (nh_shrink_selftest):
    which 0 F density, 1 F upper tail, 2 F quantile, 3 chi-square upper tail,
    4 lowess(x, y, span a, b passes), 5 log x - digamma x, 6 trigamma,
    7 inverse trigamma."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    yy = np.ascontiguousarray(y if y is not None else np.zeros_like(x), dtype=np.float64).ravel()
    out = np.zeros_like(x)
    L = lib()
    L.nh_shrink_selftest.restype = C.c_int
    L.nh_shrink_selftest.argtypes = [C.c_int32, _dp, _dp, C.c_size_t, C.c_double, C.c_double,
                                     _dp]
    rc = L.nh_shrink_selftest(which, x.ctypes.data_as(_dp), yy.ctypes.data_as(_dp), x.size,
                              float(a), float(b), out.ctypes.data_as(_dp))
    if rc != 0:
        _raise(rc, C.create_string_buffer(b"shrink selftest failed", 64))
    return out


def selftest_f64_math(x, which, device=0):
    """The FP64 sweep's exp (which 0), log (1) or reciprocal (2) of x,
    evaluated on the device."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    out = np.zeros_like(x)
    err = C.create_string_buffer(512)
    L = lib()
    L.nh_selftest_f64_math.restype = C.c_int
    L.nh_selftest_f64_math.argtypes = [_dp, C.c_size_t, C.c_int32, C.c_int32, _dp, C.c_char_p,
                                       C.c_size_t]
    rc = L.nh_selftest_f64_math(x.ctypes.data_as(_dp), len(x), which, device,
                                out.ctypes.data_as(_dp), err, 512)
    if rc != 0:
        _raise(rc, err)
    return out


# ---------------------------------------------------------------------------
class Session:
    """nebula_sparse_hip split at the device seam: set-up, upload and stage one
    on construction; stage_two() runs the search with inputs resident on the
    device (repeatable, timeable); finish() runs stage three."""

    def __init__(self, counts: CompressedSparse, subject_id, design, n_coef, offset=None,
                 params: Optional[NebulaParams] = None, wg_size=0, device=0, host_threads=0,
                 mlp_weights=None, learned_step=0.0, learned_xtol=0.0, zero_tables=1):
        L = lib()
        if not hasattr(L, "_session_bound"):
            L.nh_session_create.restype = C.c_int
            L.nh_session_create.argtypes = [C.c_size_t, C.c_size_t, _u32p, _u32p, _dp, _u64p, _dp,
                                            C.c_size_t, _dp, C.POINTER(_Params),
                                            C.POINTER(_Options), C.POINTER(C.c_void_p),
                                            C.c_char_p, C.c_size_t]
            L.nh_session_genes.restype = C.c_size_t
            L.nh_session_genes.argtypes = [C.c_void_p]
            L.nh_session_stage_one.restype = C.c_int
            L.nh_session_stage_one.argtypes = [C.c_void_p, _dp, C.POINTER(C.c_int32)]
            L.nh_session_stage_two.restype = C.c_int
            L.nh_session_stage_two.argtypes = [C.c_void_p, C.c_int32, _dp, _dp,
                                               C.POINTER(C.c_int32), C.POINTER(RunStats),
                                               C.c_char_p, C.c_size_t]
            L.nh_session_finish.restype = C.c_int
            L.nh_session_finish.argtypes = [C.c_void_p, C.POINTER(_Fit), _dp, C.c_char_p,
                                            C.c_size_t]
            L.nh_session_free.argtypes = [C.c_void_p]
            L.nh_session_release_plans.restype = C.c_int
            L.nh_session_release_plans.argtypes = [C.c_void_p]
            L.nh_session_features.restype = C.c_int
            L.nh_session_features.argtypes = [C.c_void_p, _fp]
            L.nh_session_set_mlp.restype = C.c_int
            L.nh_session_set_mlp.argtypes = [C.c_void_p, _fp, C.c_char_p, C.c_size_t]
            L.nh_session_warm_rejected.restype = C.c_int
            L.nh_session_warm_rejected.argtypes = [C.c_void_p, C.POINTER(C.c_uint8)]
            L.nh_session_autotune.restype = C.c_int
            L.nh_session_autotune.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int32), _dp,
                                              C.c_char_p, C.c_size_t]
            L._session_bound = True
        n_genes, n_cells = counts.shape
        self._keep = [np.ascontiguousarray(counts.indptr, dtype=np.uint32),
                      np.ascontiguousarray(counts.indices, dtype=np.uint32),
                      np.ascontiguousarray(counts.data, dtype=np.float64),
                      np.ascontiguousarray(subject_id, dtype=np.uint64),
                      np.ascontiguousarray(design, dtype=np.float64)]
        ip, ix, dv, sid, de = self._keep
        off = None
        offp = None
        if offset is not None:
            off, offp = _d(offset)
        p = (params or NebulaParams())._c()
        o = _Options()
        lib().nebula_hip_options_default(C.byref(o))
        o.wg_size = wg_size
        o.device = device
        o.host_threads = host_threads
        o.learned_step = learned_step
        o.learned_xtol = learned_xtol
        o.zero_tables = int(zero_tables)
        w = None
        if mlp_weights is not None:
            w = np.ascontiguousarray(mlp_weights, dtype=np.float32)
            o.mlp_weights = w.ctypes.data_as(_fp)
        h = C.c_void_p()
        err = C.create_string_buffer(1024)
        rc = L.nh_session_create(n_genes, n_cells, ip.ctypes.data_as(_u32p),
                                 ix.ctypes.data_as(_u32p), dv.ctypes.data_as(_dp),
                                 sid.ctypes.data_as(_u64p), de.ctypes.data_as(_dp), n_coef, offp,
                                 C.byref(p), C.byref(o), C.byref(h), err, 1024)
        if rc != 0:
            _raise(rc, err)
        self._h = h
        self.n_kept = L.nh_session_genes(h)

    def stage_one(self):
        sg = np.zeros((self.n_kept, 2))
        ords = np.zeros(self.n_kept, dtype=np.int32)
        lib().nh_session_stage_one(self._h, sg.ctypes.data_as(_dp),
                                   ords.ctypes.data_as(C.POINTER(C.c_int32)))
        return sg, ords

    def features(self):
        """The 12 warm-start features of every kept gene (kept x 12)."""
        f = np.zeros((self.n_kept, 12), dtype=np.float32)
        lib().nh_session_features(self._h, f.ctypes.data_as(_fp))
        return f

    def set_mlp(self, params):
        w = np.ascontiguousarray(params, dtype=np.float32)
        err = C.create_string_buffer(1024)
        rc = lib().nh_session_set_mlp(self._h, w.ctypes.data_as(_fp), err, 1024)
        if rc != 0:
            _raise(rc, err)

    def autotune(self, engine=ENGINE_BATCHED):
        """Times stage two at 4, 8 and 16 waves per gene (the faster of two runs
        each) and keeps the fastest;
        returns (threads per gene, {threads: ms})."""
        ms = np.zeros(3)
        pick = C.c_int32()
        err = C.create_string_buffer(1024)
        rc = lib().nh_session_autotune(self._h, engine, C.byref(pick), ms.ctypes.data_as(_dp),
                                       err, 1024)
        if rc != 0:
            _raise(rc, err)
        return pick.value, {256: ms[0], 512: ms[1], 1024: ms[2]}

    def stage_two(self, engine=ENGINE_BATCHED):
        mins = np.zeros((self.n_kept, 2))
        obj = np.zeros(self.n_kept)
        ev = np.zeros(self.n_kept, dtype=np.int32)
        st = RunStats()
        err = C.create_string_buffer(1024)
        rc = lib().nh_session_stage_two(self._h, engine, mins.ctypes.data_as(_dp),
                                        obj.ctypes.data_as(_dp),
                                        ev.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(st),
                                        err, 1024)
        if rc != 0:
            _raise(rc, err)
        rej = np.zeros(self.n_kept, dtype=np.uint8)
        lib().nh_session_warm_rejected(self._h, rej.ctypes.data_as(C.POINTER(C.c_uint8)))
        return dict(minimisers=mins, objectives=obj, evaluations=ev, stats=st.as_dict(),
                    warm_rejected=rej.astype(bool))

    def release_plans(self):
        """Drops the kept stage-two plans; the next run of each launch shape
        stages its searches and captures its launch again."""
        rc = lib().nh_session_release_plans(self._h)
        if rc != 0:
            _raise(rc, C.create_string_buffer(b"null session"))

    def finish(self):
        fit = _Fit()
        t3 = C.c_double()
        err = C.create_string_buffer(1024)
        rc = lib().nh_session_finish(self._h, C.byref(fit), C.byref(t3), err, 1024)
        if rc != 0:
            _raise(rc, err)
        g, nb = fit.n_genes, fit.n_coef
        pk = nb * (nb + 1) // 2
        arr = np.ctypeslib.as_array
        out = NebulaFit(
            coefficients=arr(fit.coefficients, (g * nb,)).reshape(g, nb).copy(),
            covariance=arr(fit.covariance, (g * pk,)).reshape(g, pk).copy(),
            se=arr(fit.se, (g * nb,)).reshape(g, nb).copy(),
            subject_overdispersion=arr(fit.subject_overdispersion, (g,)).copy(),
            cell_overdispersion=arr(fit.cell_overdispersion, (g,)).copy(),
            convergence=arr(fit.convergence, (g,)).copy(),
            sigma_at_bound=arr(fit.sigma_at_bound, (g,)).astype(bool),
            gene_index=arr(fit.gene_index, (g,)).astype(np.int64),
            n_coef=nb, stats={"t_stage_three": t3.value})
        lib().nebula_fit_free(C.byref(fit))
        return out

    def close(self):
        if getattr(self, "_h", None):
            lib().nh_session_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Warm-start MLP (device kernels behind nh_mlp_*).
def mlp_param_count():
    return int(lib().nh_mlp_param_count())


def _f(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_fp)


def mlp_train(features, targets, epochs=60, batch=256, lr=2e-3, seed=0, device=0):
    """Trains the MLP on the device; returns (parameter blob, final loss)."""
    f, fp = _f(features)
    t, tp = _f(targets)
    assert f.shape[1] == 12 and t.shape[1] == 2 and f.shape[0] == t.shape[0]
    out = np.zeros(mlp_param_count(), dtype=np.float32)
    loss = C.c_float()
    err = C.create_string_buffer(1024)
    rc = lib().nh_mlp_train(fp, tp, f.shape[0], epochs, batch, lr, seed, device,
                            out.ctypes.data_as(_fp), C.byref(loss), err, 1024)
    if rc != 0:
        _raise(rc, err)
    return out, loss.value


def mlp_loss_grad(params, features, targets, device=0):
    """Loss and gradient (network part of the blob) from the device kernels."""
    w, wp = _f(params)
    f, fp = _f(features)
    t, tp = _f(targets)
    g = np.zeros(mlp_param_count() - 28, dtype=np.float32)
    loss = C.c_float()
    err = C.create_string_buffer(1024)
    rc = lib().nh_mlp_loss_grad(wp, fp, tp, f.shape[0], device, C.byref(loss),
                                g.ctypes.data_as(_fp), err, 1024)
    if rc != 0:
        _raise(rc, err)
    return loss.value, g


def mlp_predict(params, features, device=0):
    """BF16-MFMA inference: n x (log sigma^2, log phi)."""
    w, wp = _f(params)
    f, fp = _f(features)
    out = np.zeros((f.shape[0], 2), dtype=np.float32)
    err = C.create_string_buffer(1024)
    rc = lib().nh_mlp_predict(wp, fp, f.shape[0], device, out.ctypes.data_as(_fp), err, 1024)
    if rc != 0:
        _raise(rc, err)
    return out


def load_weights(path=WEIGHTS_PATH):
    """The trained warm-start parameters shipped with the package."""
    w = np.load(path, allow_pickle=False).astype(np.float32)
    if w.size != mlp_param_count():
        raise ValueError(f"{path}: {w.size} parameters, expected {mlp_param_count()}")
    return w
