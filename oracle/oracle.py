"""ctypes binding of the CPU oracle (oracle/ccd_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product library never loads it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libccd_oracle.so")
_lib = None


class OrcData(ctypes.Structure):
    _fields_ = [
        ("K", ctypes.c_int), ("J", ctypes.c_int),
        ("col_ptr", ctypes.c_void_p), ("rows", ctypes.c_void_p), ("vals", ctypes.c_void_p),
        ("fmt", ctypes.c_void_p), ("y", ctypes.c_void_p), ("offs", ctypes.c_void_p),
        ("strata", ctypes.c_void_p), ("model", ctypes.c_int),
    ]


class OrcControl(ctypes.Structure):
    _fields_ = [
        ("tolerance", ctypes.c_double), ("max_iterations", ctypes.c_int),
        ("convergence_type", ctypes.c_int), ("initial_bound", ctypes.c_double),
        ("max_bound_count", ctypes.c_int), ("has_offset", ctypes.c_int),
        ("algorithm", ctypes.c_int),
    ]


class OrcResult(ctypes.Structure):
    _fields_ = [
        ("log_likelihood", ctypes.c_double), ("log_prior", ctypes.c_double),
        ("iterations", ctypes.c_int), ("return_flag", ctypes.c_int),
        ("bound_count", ctypes.c_int), ("pred_log_likelihood", ctypes.c_double),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        vp = ctypes.c_void_p
        _lib.orc_fit.restype = ctypes.c_int
        _lib.orc_fit.argtypes = [vp] * 9
        _lib.orc_eval.restype = ctypes.c_int
        _lib.orc_eval.argtypes = [vp, vp, vp, ctypes.c_int, vp, vp, vp]
        _lib.orc_eval_cols.restype = ctypes.c_int
        _lib.orc_eval_cols.argtypes = [vp, vp, vp, vp, ctypes.c_int, vp, vp, vp]
        _lib.orc_prior_delta.restype = ctypes.c_double
        _lib.orc_prior_delta.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double]
        _lib.orc_prior_log_density.restype = ctypes.c_double
        _lib.orc_prior_log_density.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
        _lib.orc_grid_point.restype = ctypes.c_double
        _lib.orc_grid_point.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                        ctypes.c_int]
        _lib.orc_mm_eval.restype = ctypes.c_int
        _lib.orc_mm_eval.argtypes = [vp] * 5
        _lib.orc_mm_xbeta.restype = ctypes.c_int
        _lib.orc_mm_xbeta.argtypes = [vp] * 3
        _lib.orc_cox_pids.restype = ctypes.c_int
        _lib.orc_cox_pids.argtypes = [vp] * 5
        _lib.orc_search_step.restype = ctypes.c_int
        _lib.orc_search_step.argtypes = [vp, vp, vp, ctypes.c_int] + [ctypes.c_double] * 4 + [vp] * 3
        _lib.orc_search_quadratic.restype = ctypes.c_int
        _lib.orc_search_quadratic.argtypes = [vp, vp, vp, ctypes.c_int, vp]
        _lib.orc_bootstrap_weights.restype = ctypes.c_int
        _lib.orc_bootstrap_weights.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_long,
                                               ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                               ctypes.c_void_p]
        _lib.orc_cv_weights.restype = ctypes.c_int
        _lib.orc_cv_weights.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_long, ctypes.c_int,
                                        ctypes.c_void_p]
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def _data_struct(p, keep):
    arrs = dict(
        col_ptr=np.ascontiguousarray(p.col_ptr, dtype=np.int64),
        rows=np.ascontiguousarray(p.rows, dtype=np.int32),
        vals=np.ascontiguousarray(p.vals, dtype=np.float64),
        fmt=np.ascontiguousarray(p.fmt, dtype=np.int8),
        y=np.ascontiguousarray(p.y, dtype=np.float64),
        offs=np.ascontiguousarray(p.offs, dtype=np.float64),
        strata=np.ascontiguousarray(p.strata, dtype=np.int32),
    )
    keep.append(arrs)
    return OrcData(p.K, p.J, _ptr(arrs["col_ptr"]), _ptr(arrs["rows"]), _ptr(arrs["vals"]),
                   _ptr(arrs["fmt"]), _ptr(arrs["y"]), _ptr(arrs["offs"]), _ptr(arrs["strata"]),
                   int(p.model))


def fit(p, spec, heldout=None, algorithm="ccd"):
    """Run the oracle fit.  `p` is a ccd_amd.data.CcdProblem, `spec` a FitSpec,
    `algorithm` "ccd" or "mm".  Returns (beta, dict of diagnostics); the dict's
    "status" is 0, -1 (no MM for this model) or -2 (MM log posterior decreased)."""
    keep = []
    d = _data_struct(p, keep)
    pt = np.ascontiguousarray(spec.prior_type, dtype=np.int32)
    pv = np.ascontiguousarray(spec.prior_var, dtype=np.float64)
    fx = np.ascontiguousarray(spec.fixed, dtype=np.uint8)
    beta = np.array(spec.beta0, dtype=np.float64, copy=True)
    w = None if spec.weights is None else np.ascontiguousarray(spec.weights, dtype=np.float64)
    ho = None if heldout is None else np.ascontiguousarray(heldout, dtype=np.float64)
    ctl = OrcControl(spec.tolerance, spec.max_iterations, spec.convergence_type,
                     spec.initial_bound, spec.max_bound_count, int(p.has_offset),
                     {"ccd": 0, "mm": 1}[algorithm])
    res = OrcResult()
    status = lib().orc_fit(ctypes.byref(d), _ptr(pt), _ptr(pv), _ptr(fx), _ptr(beta), _ptr(w), _ptr(ho),
                  ctypes.byref(ctl), ctypes.byref(res))
    return beta, dict(log_likelihood=res.log_likelihood, log_prior=res.log_prior,
                      iterations=res.iterations, return_flag=res.return_flag,
                      bound_count=res.bound_count, pred_log_likelihood=res.pred_log_likelihood,
                      status=status)


def evaluate(p, beta, j=-1, weights=None):
    """(log likelihood, gradient_j, hessian_j) at `beta` with fresh statistics."""
    keep = []
    d = _data_struct(p, keep)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    ll, g, h = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().orc_eval(ctypes.byref(d), _ptr(b), _ptr(w), ctypes.c_int(j), ctypes.byref(ll),
                   ctypes.byref(g), ctypes.byref(h))
    return ll.value, g.value, h.value


def evaluate_cols(p, beta, cols, weights=None):
    """(log likelihood, gradients, Hessians) of the coordinates `cols` at `beta`,
    from one set of fresh statistics."""
    keep = []
    d = _data_struct(p, keep)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    c = np.ascontiguousarray(cols, dtype=np.int32)
    g = np.zeros(len(c))
    h = np.zeros(len(c))
    ll = ctypes.c_double()
    lib().orc_eval_cols(ctypes.byref(d), _ptr(b), _ptr(w), _ptr(c), len(c), ctypes.byref(ll),
                        _ptr(g), _ptr(h))
    return ll.value, g, h


def mm_evaluate(p, beta, weights=None):
    """MM gradients and Hessians of every column at `beta` (fresh statistics)."""
    keep = []
    d = _data_struct(p, keep)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    g = np.zeros(p.J)
    h = np.zeros(p.J)
    rc = lib().orc_mm_eval(ctypes.byref(d), _ptr(b), _ptr(w), _ptr(g), _ptr(h))
    if rc != 0:
        raise ValueError("no MM increments for this model")
    return g, h


def mm_xbeta(p, beta):
    """xBeta formed from the transpose, as the MM step forms it."""
    keep = []
    d = _data_struct(p, keep)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    out = np.zeros(p.K)
    lib().orc_mm_xbeta(ctypes.byref(d), _ptr(b), _ptr(out))
    return out


def prior_delta(ptype, var, g, h, beta):
    return lib().orc_prior_delta(ptype, var, g, h, beta)


def prior_log_density(ptype, var, beta):
    return lib().orc_prior_log_density(ptype, var, beta)


def grid_point(lower, upper, size, step):
    return lib().orc_grid_point(lower, upper, size, step)


def cox_pids(p, weights=None):
    keep = []
    d = _data_struct(p, keep)
    pid = np.zeros(p.K, dtype=np.int32)
    reset = np.zeros(p.K + 1, dtype=np.int32)
    nres = ctypes.c_int()
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    N = lib().orc_cox_pids(ctypes.byref(d), _ptr(w), _ptr(pid), _ptr(reset), ctypes.byref(nres))
    return N, pid, reset[: nres.value].copy()


def search_step(tried, stdstep=10.0, stop_by_y=0.01, stop_by_x=np.log(1.5), first_cut=1.0):
    t = np.asarray(tried, dtype=np.float64).reshape(-1, 3)
    xs, ys, sd = (np.ascontiguousarray(t[:, i]) for i in range(3))
    cont, nxt, exp_ = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
    lib().orc_search_step(_ptr(xs), _ptr(ys), _ptr(sd), len(t), stdstep, stop_by_y, stop_by_x,
                          first_cut, ctypes.byref(cont), ctypes.byref(nxt), ctypes.byref(exp_))
    return bool(cont.value), nxt.value, exp_.value


def search_quadratic(tried):
    t = np.asarray(tried, dtype=np.float64).reshape(-1, 3)
    xs, ys, sd = (np.ascontiguousarray(t[:, i]) for i in range(3))
    c = np.zeros(3)
    lib().orc_search_quadratic(_ptr(xs), _ptr(ys), _ptr(sd), len(t), _ptr(c))
    return c


def bootstrap_weights(ids, seed, replicate, exclude=None):
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    ex = None if exclude is None else np.ascontiguousarray(exclude, dtype=np.float64)
    out = np.zeros(len(ids), dtype=np.float64)
    lib().orc_bootstrap_weights(_ptr(ids), len(ids), seed, replicate, _ptr(ex),
                                0 if ex is None else len(ex), _ptr(out))
    return out


def cv_weights(ids, nfold, seed, task, by_pid=True):
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    out = np.zeros(len(ids), dtype=np.float64)
    lib().orc_cv_weights(_ptr(ids), len(ids), int(by_pid), nfold, seed, task, _ptr(out))
    return out


def grid_search_cv_chain(p, spec, ids, variances, n_fold=10, seed=123, by_pid=True, threads=1):
    """The reference's warm-started grid search (resetCoefficients = FALSE), fit by
    fit on the CPU: AbstractCrossValidationDriver::doCrossValidationStep
    (AbstractCrossValidationDriver.cpp:131-248) runs a grid point's tasks on the
    TaskScheduler's threads (Thread.h:64-80: thread t takes tasks [t*c, (t+1)*c),
    c = ceil(n_fold / threads)), each thread's CCD starting every task from the
    coefficients its previous task left (ccdTask->update without resetBeta), and
    setting them to zero after a task that did not succeed (:219-225); the grid
    loop (GridSearchCrossValidationDriver.cpp:96-140) keeps the threads' CCDs from
    grid point to grid point.  Returns (scores [grid x n_fold], results, chains'
    final coefficients)."""
    import dataclasses
    T = max(1, min(threads, n_fold))
    c = n_fold // T + (n_fold % T != 0)
    chains = [list(range(t * c, min(n_fold, (t + 1) * c))) for t in range(T) if t * c < n_fold]
    zero = np.zeros(p.J)
    if p.has_offset:
        zero[0] = spec.beta0[0]
    beta = [np.array(spec.beta0, np.float64) for _ in chains]
    pt = np.asarray(spec.prior_type)
    scores = np.full((len(variances), n_fold), np.nan)
    results = [[None] * n_fold for _ in variances]
    for g, var in enumerate(variances):
        pv = np.where(pt != 0, var, np.asarray(spec.prior_var))
        for t, chain in enumerate(chains):
            for f in chain:
                w = cv_weights(ids, n_fold, seed, f, by_pid=by_pid)
                s = dataclasses.replace(spec, prior_var=pv, weights=w, beta0=beta[t].copy())
                b, info = fit(p, s, heldout=1.0 - w)
                results[g][f] = info
                if info["return_flag"] == 0:
                    scores[g, f] = info["pred_log_likelihood"]
                    beta[t] = b
                else:
                    beta[t] = zero.copy()
    return scores, results, beta
