"""ctypes view of the CPU f64 oracle (oracle/lib/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline import this module, and only as the checker.  Build with
``make -C oracle`` (``__graft_entry__.build()`` does it).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liboracle.so")

_dp = C.POINTER(C.c_double)
_u64p = C.POINTER(C.c_uint64)
_u32p = C.POINTER(C.c_uint32)
_ip = C.POINTER(C.c_int)


class NebulaParams(C.Structure):
    """nebula_params_t (include/nebula_hip.h), NebulaParams in edge-rs."""
    _fields_ = [("min_subject", C.c_double), ("min_cell", C.c_double),
                ("max_subject", C.c_double), ("max_cell", C.c_double),
                ("method", C.c_int32), ("reml", C.c_int32),
                ("cutoff_cell", C.c_double), ("kappa", C.c_double),
                ("cpc", C.c_double), ("mincp", C.c_uint64),
                ("min_subjects", C.c_uint64), ("eps", C.c_double)]


class NebulaFitC(C.Structure):
    _fields_ = [("n_genes", C.c_size_t), ("n_coef", C.c_size_t),
                ("coefficients", _dp), ("covariance", _dp), ("se", _dp),
                ("subject_overdispersion", _dp), ("cell_overdispersion", _dp),
                ("convergence", C.POINTER(C.c_int32)),
                ("sigma_at_bound", C.POINTER(C.c_uint8)),
                ("gene_index", _u64p)]


def params(method="LN", **kw):
    p = NebulaParams(1e-4, 1e-4, 10.0, 1000.0, 0, 0, 20.0, 800.0, 0.005, 5, 0, 1e-6)
    p.method = 1 if method == "HL" else 0
    for k, v in kw.items():
        if k == "min":
            p.min_subject, p.min_cell = v
        elif k == "max":
            p.max_subject, p.max_cell = v
        else:
            setattr(p, k, v)
    return p


_lib = None


def default_threads():
    """Worker threads for the oracle: OMP_NUM_THREADS where the environment
    sets the process's CPU share, else the CPUs this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:
        return max(1, os.cpu_count() or 1)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing at {LIB_PATH}; run make -C oracle")
        L = C.CDLL(LIB_PATH)
        L.oracle_nebula_sparse.restype = C.c_int
        L.oracle_nebula_sparse.argtypes = [C.c_size_t, C.c_size_t, _u32p, _u32p, _dp, _u64p,
                                           _dp, C.c_size_t, _dp, C.POINTER(NebulaParams),
                                           C.c_int, C.POINTER(NebulaFitC), _dp, _dp,
                                           C.c_int, C.c_char_p, C.c_size_t]
        L.oracle_fit_free.argtypes = [C.POINTER(NebulaFitC)]
        L.oracle_opt_pml.restype = C.c_int
        L.oracle_opt_pml.argtypes = [_dp, _dp, C.c_size_t, C.c_size_t, _dp, _u64p, C.c_size_t,
                                     _u64p, C.c_size_t, _dp, _dp, _dp, C.c_double, C.c_double,
                                     C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                                     _dp, _dp, _dp, _dp, _ip, C.c_char_p, C.c_size_t]
        L.oracle_pml_ll_grad.restype = C.c_int
        L.oracle_pml_ll_grad.argtypes = [_dp, _dp, C.c_size_t, C.c_size_t, _dp, _u64p,
                                         C.c_size_t, _u64p, C.c_size_t, _dp, _dp, _dp,
                                         C.c_double, C.c_double, _dp, _dp]
        L.oracle_ptmg.restype = C.c_int
        L.oracle_ptmg.argtypes = [_dp, _dp, C.c_size_t, C.c_size_t, _dp, _u64p, C.c_size_t,
                                  _dp, _u64p, C.c_size_t, _dp, C.c_int, _dp, _dp, _dp]
        L.oracle_zero_group.restype = None
        L.oracle_zero_group.argtypes = [_dp, C.c_size_t, _dp, C.c_size_t, C.c_int, C.c_int, _dp,
                                        _dp, _dp]
        L.oracle_zero_cells.restype = C.c_int
        L.oracle_zero_cells.argtypes = [_dp, _dp, C.c_size_t, _u64p, C.c_size_t, C.c_size_t,
                                        _u64p, _u64p, _u64p, _u64p, _dp]
        L.oracle_special.argtypes = [C.c_int, _dp, _dp, C.c_size_t]
        L.oracle_min_eigenvalue.restype = C.c_double
        L.oracle_min_eigenvalue.argtypes = [_dp, C.c_size_t]
        L.oracle_ldlt_solve.argtypes = [_dp, _dp, C.c_size_t]
        L.oracle_determinant.restype = C.c_double
        L.oracle_determinant.argtypes = [_dp, C.c_size_t]
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _u64(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a, a.ctypes.data_as(_u64p)


def _u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(_u32p)


class OracleError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code


def opt_pml(design, offset, counts, cells, subject_start, subject_total, beta_init,
            subject, cell, reml=False, eps=1e-6, ord=1, log_w_init=None, max_iter=0,
            nbm=False, tabled=False):
    """oracle opt_pml (src/sc/pml.rs:468) on one gene; returns a dict."""
    design = np.asarray(design, dtype=np.float64)
    n_cells = len(offset)
    nb = design.size // n_cells
    k = len(subject_start) - 1
    keep = []
    args = []
    for a in (design, offset, counts):
        x, p = _d(a)
        keep.append(x)
        args.append(p)
    c, cp = _u64(cells)
    s, sp = _u64(subject_start)
    t, tp = _d(subject_total)
    b, bp = _d(beta_init)
    lw = None
    lwp = None
    if log_w_init is not None:
        lw, lwp = _d(log_w_init)
    beta = np.zeros(nb)
    log_w = np.zeros(k)
    info = np.zeros(nb * nb)
    sc = np.zeros(4)
    ints = np.zeros(3, dtype=np.int32)
    err = C.create_string_buffer(512)
    rc = lib().oracle_opt_pml(args[0], args[1], n_cells, nb, args[2], cp, len(c), sp, k, tp,
                              bp, lwp, subject, cell, int(reml), eps, ord, max_iter, int(nbm),
                              int(tabled), beta.ctypes.data_as(_dp), log_w.ctypes.data_as(_dp),
                              info.ctypes.data_as(_dp), sc.ctypes.data_as(_dp),
                              ints.ctypes.data_as(_ip), err, 512)
    if rc != 0:
        raise OracleError(rc, err.value.decode())
    return dict(beta=beta, log_w=log_w, information=info, log_likelihood=sc[0],
                log_likelihood_prev=sc[1], log_det=sc[2], second_order=sc[3],
                convergence=int(ints[0]), iterations=int(ints[1]), backtracks=int(ints[2]))


def pml_ll_grad(design, offset, counts, cells, subject_start, subject_total, beta, log_w,
                subject, cell):
    design = np.asarray(design, dtype=np.float64)
    n_cells = len(offset)
    nb = design.size // n_cells
    k = len(subject_start) - 1
    d, dp = _d(design)
    o, op = _d(offset)
    y, yp = _d(counts)
    c, cp = _u64(cells)
    s, sp = _u64(subject_start)
    t, tp = _d(subject_total)
    b, bp = _d(beta)
    w, wp = _d(log_w)
    obj = C.c_double()
    grad = np.zeros(nb + k)
    rc = lib().oracle_pml_ll_grad(dp, op, n_cells, nb, yp, cp, len(c), sp, k, tp, bp, wp,
                                  subject, cell, C.byref(obj), grad.ctypes.data_as(_dp))
    if rc:
        raise OracleError(rc, "pml_ll_grad")
    return obj.value, grad


def ptmg(design, log_offset, counts, cells, totals, fid, params, hessian=True, tabled=False):
    """Stage-one marginal NLL, gradient and Hessian (src/sc/ptmg.rs:585,634)."""
    design = np.asarray(design, dtype=np.float64)
    n = len(log_offset)
    nb = design.size // n
    d, dp = _d(design)
    o, op = _d(log_offset)
    y, yp = _d(counts)
    c, cp = _u64(cells)
    t, tp = _d(totals)
    f, fp = _u64(fid)
    x, xp = _d(params)
    v = C.c_double()
    g = np.zeros(nb + 2)
    h = np.zeros((nb + 2) ** 2) if hessian else None
    rc = lib().oracle_ptmg(dp, op, n, nb, yp, cp, len(c), tp, fp, len(f) - 1, xp, int(tabled),
                           C.byref(v), g.ctypes.data_as(_dp),
                           h.ctypes.data_as(_dp) if hessian else None)
    if rc:
        raise OracleError(rc, "zero-count tables do not apply to this design")
    return v.value, g, h


def zero_group(log_o, s, frm=0, m=6, direct=True):
    """One zero-count group (src/sc/zeros.rs:291) from log offsets: the table
    at points s (len(s) x m), the compensated direct sums (len(s) x 6) and
    (s_lo, s_hi, n_base, n_panels)."""
    lo, lop = _d(log_o)
    sv, sp = _d(np.atleast_1d(s))
    out = np.zeros(len(sv) * m)
    dirs = np.zeros(len(sv) * 6) if direct else None
    info = np.zeros(4)
    lib().oracle_zero_group(lop, len(lo), sp, len(sv), frm, m, out.ctypes.data_as(_dp),
                            dirs.ctypes.data_as(_dp) if direct else None,
                            info.ctypes.data_as(_dp))
    return (out.reshape(len(sv), m), dirs.reshape(len(sv), 6) if direct else None, info)


def zero_cells(design, log_offset, fid):
    """ZeroCells::build (src/sc/zeros.rs:219): None when the tables do not
    apply, else a dict of subject_groups, loose, subject_loose, group_row and
    group_n."""
    design = np.asarray(design, dtype=np.float64)
    n = len(log_offset)
    nb = design.size // n
    d, dp = _d(design)
    o, op = _d(log_offset)
    f, fp = _u64(fid)
    k = len(f) - 1
    sg = np.zeros(k + 1, dtype=np.uint64)
    sl = np.zeros(k + 1, dtype=np.uint64)
    lo = np.zeros(n, dtype=np.uint64)
    gr = np.zeros(n, dtype=np.uint64)
    gn = np.zeros(n)
    ok = lib().oracle_zero_cells(dp, op, n, fp, k, nb, sg.ctypes.data_as(_u64p),
                                 sl.ctypes.data_as(_u64p), lo.ctypes.data_as(_u64p),
                                 gr.ctypes.data_as(_u64p), gn.ctypes.data_as(_dp))
    if not ok:
        return None
    ng = int(sg[-1])
    return dict(subject_groups=sg.astype(np.int64), subject_loose=sl.astype(np.int64),
                loose=lo[: int(sl[-1])].astype(np.int64), group_row=gr[:ng].astype(np.int64),
                group_n=gn[:ng].copy())


_FG = C.CFUNCTYPE(C.c_double, _dp, _dp, C.c_void_p)


def lbfgsb(fg, x0, lower, upper, ftol=1e-12, pgtol=1e-10, max_iter=500):
    """L-BFGS-B (src/numeric/lbfgsb.rs:876, memory 10, 40 evaluations per line
    search) on fg(x) -> (f, g).  Returns (x, f, ok)."""
    n = len(x0)
    x0, x0p = _d(x0)
    lo, lop = _d(lower)
    hi, hip = _d(upper)

    def cb(xp, gp, _ctx):
        x = np.ctypeslib.as_array(xp, (n,)).copy()
        f, g = fg(x)
        gv = np.ctypeslib.as_array(gp, (n,))
        gv[:] = g
        return float(f)

    fn = _FG(cb)
    out = np.zeros(n)
    f = C.c_double()
    L = lib()
    L.oracle_lbfgsb.restype = C.c_int
    L.oracle_lbfgsb.argtypes = [C.c_size_t, _dp, _dp, _dp, C.c_double, C.c_double, C.c_int, _FG,
                                C.c_void_p, _dp, C.POINTER(C.c_double)]
    rc = L.oracle_lbfgsb(n, x0p, lop, hip, ftol, pgtol, max_iter, fn, None,
                         out.ctypes.data_as(_dp), C.byref(f))
    return out, f.value, rc == 0


def special(which, x):
    x, xp = _d(np.atleast_1d(x))
    y = np.zeros_like(x)
    lib().oracle_special({"lgamma": 0, "digamma": 1, "trigamma": 2}[which], xp,
                         y.ctypes.data_as(_dp), len(x))
    return y


def nebula_sparse(indptr, indices, data, n_cells, subject_id, design, n_coef, offset=None,
                  prm=None, threads=0, with_stage_two=False, tables=True):
    """oracle nebula_sparse (src/sc/nebula.rs:434) -> dict of numpy arrays."""
    n_genes = len(indptr) - 1
    ip, ipp = _u32(indptr)
    ix, ixp = _u32(indices)
    dv, dvp = _d(data)
    sid, sidp = _u64(subject_id)
    de, dep = _d(design)
    off = None
    offp = None
    if offset is not None:
        off, offp = _d(offset)
    threads = threads or default_threads()
    fit = NebulaFitC()
    info = np.zeros(8 * n_genes) if with_stage_two else None
    t2 = C.c_double(0.0)
    err = C.create_string_buffer(1024)
    rc = lib().oracle_nebula_sparse(n_genes, n_cells, ipp, ixp, dvp, sidp, dep, n_coef, offp,
                                    C.byref(prm) if prm is not None else None, threads,
                                    C.byref(fit),
                                    info.ctypes.data_as(_dp) if info is not None else None,
                                    C.byref(t2), int(tables), err, 1024)
    if rc != 0:
        raise OracleError(rc, err.value.decode())
    g = fit.n_genes
    p = fit.n_coef
    pk = p * (p + 1) // 2
    out = dict(
        coefficients=np.ctypeslib.as_array(fit.coefficients, (g * p,)).reshape(g, p).copy(),
        covariance=np.ctypeslib.as_array(fit.covariance, (g * pk,)).reshape(g, pk).copy(),
        se=np.ctypeslib.as_array(fit.se, (g * p,)).reshape(g, p).copy(),
        subject_overdispersion=np.ctypeslib.as_array(fit.subject_overdispersion, (g,)).copy(),
        cell_overdispersion=np.ctypeslib.as_array(fit.cell_overdispersion, (g,)).copy(),
        convergence=np.ctypeslib.as_array(fit.convergence, (g,)).copy(),
        sigma_at_bound=np.ctypeslib.as_array(fit.sigma_at_bound, (g,)).astype(bool),
        gene_index=np.ctypeslib.as_array(fit.gene_index, (g,)).copy(),
        t_stage_two=t2.value,
    )
    if info is not None:
        out["stage_two"] = info[: 8 * g].reshape(g, 8).copy()
    lib().oracle_fit_free(C.byref(fit))
    return out


def profile_objective(indptr, indices, data, n_cells, subject_id, design, n_coef, offset, prm,
                      genes, points, threads=0, tables=True):
    """The f64 profile objective (nebula's pql_ll, -loglik) at (subject, cell)
    points for kept-gene indices `genes`."""
    n_genes = len(indptr) - 1
    ip, ipp = _u32(indptr)
    ix, ixp = _u32(indices)
    dv, dvp = _d(data)
    sid, sidp = _u64(subject_id)
    de, dep = _d(design)
    off = None
    offp = None
    if offset is not None:
        off, offp = _d(offset)
    g, gp = _u64(genes)
    pts, ptsp = _d(np.asarray(points, dtype=np.float64).reshape(-1, 2))
    out = np.zeros(len(g))
    err = C.create_string_buffer(512)
    L = lib()
    L.oracle_profile_objective.restype = C.c_int
    L.oracle_profile_objective.argtypes = [C.c_size_t, C.c_size_t, _u32p, _u32p, _dp, _u64p, _dp,
                                           C.c_size_t, _dp, C.POINTER(NebulaParams), C.c_size_t,
                                           _u64p, _dp, _dp, C.c_int, C.c_int, C.c_char_p,
                                           C.c_size_t]
    rc = L.oracle_profile_objective(n_genes, n_cells, ipp, ixp, dvp, sidp, dep, n_coef, offp,
                                    C.byref(prm), len(g), gp, ptsp, out.ctypes.data_as(_dp),
                                    threads or default_threads(), int(tables), err, 512)
    if rc != 0:
        raise OracleError(rc, err.value.decode())
    return out


_SEARCH_FN = C.CFUNCTYPE(C.c_double, _dp, C.c_int, C.c_int, C.c_int, C.c_void_p)


def stage_two_search(f, start, lower, upper, ord=1):
    """StageTwoSearch on a Python objective f(x, order, call) (the reference's
    run_search harness, src/sc/nebula.rs:3516-3542): two components run
    BOBYQA, one the simplex search.  Returns (x or None, failed, order,
    evaluations)."""
    n = len(start)
    L = lib()
    L.oracle_stage_two_search.restype = C.c_int
    L.oracle_stage_two_search.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, _SEARCH_FN, C.c_void_p,
                                          _dp, _ip, _ip, _ip]

    def cb(x, nn, order, call, ctx):
        return float(f(np.array([x[i] for i in range(nn)]), order, call))

    fn = _SEARCH_FN(cb)
    s, sp = _d(start)
    lo, lop = _d(lower)
    hi, hip = _d(upper)
    x = np.zeros(n)
    failed, order, evals = C.c_int(0), C.c_int(0), C.c_int(0)
    found = L.oracle_stage_two_search(n, sp, lop, hip, ord, fn, None, x.ctypes.data_as(_dp),
                                      C.byref(failed), C.byref(order), C.byref(evals))
    return (x if found else None), bool(failed.value), order.value, evals.value
