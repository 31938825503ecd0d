"""ctypes front end of libccdhip.so (the C ABI in include/ccdhip.h).

The library must have been built (python -m ccd_amd.build); importing this
module never falls back to any other implementation: a missing library raises.
"""
import ctypes
import dataclasses
import os

import numpy as np

from .data import CONDITIONAL_LOGISTIC, COX, CcdProblem, FitSpec

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libccdhip.so")

FP64, MIXED = 0, 1

# every symbol include/ccdhip.h declares
EXPORTS = [
    "ccdh_last_error", "ccdh_device_info", "ccdh_data_create", "ccdh_data_create_long",
    "ccdh_data_get_csc", "ccdh_data_get_tables", "ccdh_data_destroy",
    "ccdh_data_device_bytes", "ccdh_data_shape", "ccdh_data_precision", "ccdh_fit_create", "ccdh_fit_clone",
    "ccdh_fit_destroy", "ccdh_set_weights", "ccdh_fixed_terms_loglik",
    "ccdh_fixed_terms_grad_hess", "ccdh_zero_xbeta", "ccdh_axpy_xbeta", "ccdh_compute_xbeta",
    "ccdh_remaining_stats", "ccdh_numerator", "ccdh_grad_hess", "ccdh_update_xbeta",
    "ccdh_loglik", "ccdh_gradient_objective", "ccdh_pred_loglik", "ccdh_get_xbeta",
    "ccdh_mm_grad_hess", "ccdh_fit_plain", "ccdh_fit_mm", "ccdh_fit_mm_stats",
    "ccdh_mm_last_stats", "ccdh_batch_fit", "ccdh_batch_last_stats", "ccdh_cv_fold_weights",
    "ccdh_cv_grid_point", "ccdh_bootstrap_weights", "ccdh_search_step", "ccdh_search_quadratic",
    "ccdh_cox_groups", "ccdh_cox_groups_noalloc",
    "ccdh_server_create", "ccdh_server_submit",
    "ccdh_server_submit_resident", "ccdh_server_poll", "ccdh_server_wait",
    "ccdh_server_get_stats", "ccdh_server_destroy",
]


class CcdhError(RuntimeError):
    pass


class NonIncreasing(CcdhError):
    """An MM fit whose log posterior decreased; carries the coefficients reached
    and the iteration."""

    def __init__(self, msg, beta, iterations):
        super().__init__(msg)
        self.beta = beta
        self.iterations = iterations


class DataDesc(ctypes.Structure):
    _fields_ = [
        ("model", ctypes.c_int), ("precision", ctypes.c_int), ("K", ctypes.c_int),
        ("J", ctypes.c_int), ("col_ptr", ctypes.c_void_p), ("rows", ctypes.c_void_p),
        ("vals", ctypes.c_void_p), ("fmt", ctypes.c_void_p), ("y", ctypes.c_void_p),
        ("offs", ctypes.c_void_p), ("strata", ctypes.c_void_p), ("has_offset", ctypes.c_int),
    ]


class LongDesc(ctypes.Structure):
    _fields_ = [
        ("model", ctypes.c_int), ("precision", ctypes.c_int), ("n_rows", ctypes.c_int64),
        ("row_id", ctypes.c_void_p), ("stratum_id", ctypes.c_void_p), ("y", ctypes.c_void_p),
        ("time", ctypes.c_void_p), ("n_entries", ctypes.c_int64),
        ("covariate_id", ctypes.c_void_p), ("entry_row_id", ctypes.c_void_p),
        ("value", ctypes.c_void_p), ("add_intercept", ctypes.c_int),
    ]


class MmStats(ctypes.Structure):
    _fields_ = [
        ("chunk_launches", ctypes.c_int64), ("chunk_ms", ctypes.c_double),
        ("xbeta_launches", ctypes.c_int64), ("xbeta_ms", ctypes.c_double),
        ("nnz", ctypes.c_int64), ("entries", ctypes.c_int64), ("rows", ctypes.c_int64),
        ("values", ctypes.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class FitSpecC(ctypes.Structure):
    _fields_ = [
        ("prior_type", ctypes.c_void_p), ("prior_var", ctypes.c_void_p),
        ("fixed", ctypes.c_void_p), ("beta0", ctypes.c_void_p), ("weights", ctypes.c_void_p),
        ("heldout", ctypes.c_void_p), ("tolerance", ctypes.c_double),
        ("max_iterations", ctypes.c_int), ("convergence_type", ctypes.c_int),
        ("initial_bound", ctypes.c_double), ("max_bound_count", ctypes.c_int),
    ]


class FitResult(ctypes.Structure):
    _fields_ = [
        ("log_likelihood", ctypes.c_double), ("log_prior", ctypes.c_double),
        ("pred_log_likelihood", ctypes.c_double), ("iterations", ctypes.c_int),
        ("return_flag", ctypes.c_int), ("bound_count", ctypes.c_int), ("_pad", ctypes.c_int),
        ("seconds", ctypes.c_double),
    ]

    def as_dict(self):
        return dict(log_likelihood=self.log_likelihood, log_prior=self.log_prior,
                    pred_log_likelihood=self.pred_log_likelihood, iterations=self.iterations,
                    return_flag=self.return_flag, bound_count=self.bound_count,
                    seconds=self.seconds)


class BatchStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        "launches", "sweeps", "sweep_rounds", "gather_nnz", "gather_runs", "update_nnz",
        "update_runs", "scan_pids", "objective_rows", "ticks_single", "ticks_long",
        "ticks_objective", "long_columns")] + [
        ("kernel_ms", ctypes.c_double), ("wall_ms", ctypes.c_double),
        ("launch_ms", ctypes.c_float * 128), ("launch_fits", ctypes.c_int32 * 128),
        ("launch_gang", ctypes.c_int32 * 128), ("tuned_pass_items", ctypes.c_int32),
        ("tuned_max_gang", ctypes.c_int32), ("tune_ms", ctypes.c_double),
        ("tune_rates", ctypes.c_double * 4), ("bundles", ctypes.c_int32),
        ("tuned_threads", ctypes.c_int32), ("gang_restarts", ctypes.c_int32),
        ("shared_setups", ctypes.c_int32)]

    def as_dict(self):
        out = {n: getattr(self, n) for n, _ in self._fields_[:15]}
        out.update(tuned_pass_items=self.tuned_pass_items, tuned_max_gang=self.tuned_max_gang,
                   tune_ms=self.tune_ms, tune_rates=list(self.tune_rates), bundles=self.bundles,
                   tuned_threads=self.tuned_threads, gang_restarts=self.gang_restarts,
                   shared_setups=self.shared_setups)
        k = min(128, self.launches)
        out["timeline"] = [(round(self.launch_ms[i], 2), self.launch_fits[i], self.launch_gang[i])
                           for i in range(k)]
        return out


class BatchOptions(ctypes.Structure):
    _fields_ = [("threads", ctypes.c_int), ("sweeps_per_sync", ctypes.c_int),
                ("use_graph", ctypes.c_int), ("max_gang", ctypes.c_int),
                ("gang_min_nnz", ctypes.c_int64), ("pass_items", ctypes.c_int),
                ("autotune", ctypes.c_int), ("quantum_ms", ctypes.c_double),
                ("bundle", ctypes.c_int), ("solo", ctypes.c_int), ("spin_limit", ctypes.c_int64)]


class ServerOptions(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("max_slots", ctypes.c_int),
                ("threads", ctypes.c_int), ("max_gang", ctypes.c_int),
                ("quantum_ms", ctypes.c_double), ("hbm_budget_bytes", ctypes.c_int64),
                ("spin_limit", ctypes.c_int64)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("latency", ctypes.c_double), ("queued", ctypes.c_double),
                ("prepare", ctypes.c_double), ("footprint", ctypes.c_int64),
                ("launches", ctypes.c_int64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        "submitted", "completed", "failed", "in_flight", "queued", "launches")] + [
        ("kernel_ms", ctypes.c_double)] + [(n, ctypes.c_int64) for n in (
            "arena_bytes", "arena_used", "arena_peak", "gang_restarts")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


_lib = None


def load_library(path=LIB_PATH):
    """Load libccdhip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise CcdhError(f"{path} not found: build it with `python -m ccd_amd.build`")
    lib = ctypes.CDLL(path)
    vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    sig = {
        "ccdh_last_error": (ctypes.c_char_p, []),
        "ccdh_device_info": (ci, [vp, ctypes.c_size_t, vp, vp]),
        "ccdh_data_create": (ci, [vp, vp]),
        "ccdh_data_create_long": (ci, [vp, vp]),
        "ccdh_data_get_csc": (ci, [vp, vp, vp, vp, vp, vp, vp, vp]),
        "ccdh_data_get_tables": (ci, [vp] * 11),
        "ccdh_data_destroy": (ci, [vp]),
        "ccdh_data_device_bytes": (ctypes.c_int64, [vp]),
        "ccdh_data_shape": (ci, [vp, vp, vp, vp, vp]),
        "ccdh_data_precision": (ci, [vp]),
        "ccdh_fit_create": (ci, [vp, vp]),
        "ccdh_fit_clone": (ci, [vp, vp]),
        "ccdh_fit_destroy": (ci, [vp]),
        "ccdh_set_weights": (ci, [vp, vp, ci]),
        "ccdh_fixed_terms_loglik": (ci, [vp, ci]),
        "ccdh_fixed_terms_grad_hess": (ci, [vp, ci]),
        "ccdh_zero_xbeta": (ci, [vp]),
        "ccdh_axpy_xbeta": (ci, [vp, cd, ci]),
        "ccdh_compute_xbeta": (ci, [vp, vp]),
        "ccdh_remaining_stats": (ci, [vp, ci]),
        "ccdh_numerator": (ci, [vp, ci, ci]),
        "ccdh_grad_hess": (ci, [vp, ci, ci, vp, vp]),
        "ccdh_update_xbeta": (ci, [vp, cd, ci, ci]),
        "ccdh_loglik": (ci, [vp, ci, vp]),
        "ccdh_gradient_objective": (ci, [vp, ci, vp]),
        "ccdh_pred_loglik": (ci, [vp, vp, vp]),
        "ccdh_get_xbeta": (ci, [vp, vp]),
        "ccdh_mm_grad_hess": (ci, [vp, vp, ci, vp]),
        "ccdh_fit_plain": (ci, [vp, vp, vp, vp]),
        "ccdh_fit_mm": (ci, [vp, vp, vp, vp]),
        "ccdh_fit_mm_stats": (ci, [vp, vp]),
        "ccdh_mm_last_stats": (ci, [vp]),
        "ccdh_batch_fit": (ci, [vp, ci, vp, vp, vp, vp]),
        "ccdh_batch_last_stats": (ci, [vp]),
        "ccdh_cv_fold_weights": (ci, [vp, ci, ci, ci, ctypes.c_long, ci, vp]),
        "ccdh_cv_grid_point": (cd, [cd, cd, ci, ci]),
        "ccdh_bootstrap_weights": (ci, [vp, ci, ctypes.c_long, ci, vp, ci, vp]),
        "ccdh_search_step": (ci, [vp, vp, vp, ci, cd, cd, cd, cd, vp, vp, vp]),
        "ccdh_search_quadratic": (ci, [vp, vp, vp, ci, vp]),
        "ccdh_cox_groups": (ci, [vp, vp, vp, vp, vp, vp, vp]),
        "ccdh_cox_groups_noalloc": (ci, [vp, vp, vp, vp, vp, vp, vp]),
        "ccdh_server_create": (ci, [vp, vp]),
        "ccdh_server_submit": (ci, [vp, vp, vp, vp]),
        "ccdh_server_submit_resident": (ci, [vp, vp, vp, vp]),
        "ccdh_server_poll": (ci, [vp, ctypes.c_int64]),
        "ccdh_server_wait": (ci, [vp, ctypes.c_int64, cd, vp, vp, vp]),
        "ccdh_server_get_stats": (ci, [vp, vp]),
        "ccdh_server_destroy": (ci, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        msg = load_library().ccdh_last_error().decode()
        raise CcdhError(f"ccdhip error {rc}: {msg}")


def _p(a):
    return None if a is None else a.ctypes.data


def device_info():
    lib = load_library()
    buf = ctypes.create_string_buffer(256)
    cu, gb = ctypes.c_int(), ctypes.c_double()
    _check(lib.ccdh_device_info(buf, 256, ctypes.byref(cu), ctypes.byref(gb)))
    return buf.value.decode(), cu.value, gb.value


def _desc(p: CcdProblem, precision):
    """ccdh_data_desc over contiguous copies of the problem's arrays (kept in the
    returned dict, which must outlive every use of the descriptor)."""
    a = dict(
        col_ptr=np.ascontiguousarray(p.col_ptr, np.int64),
        rows=np.ascontiguousarray(p.rows, np.int32),
        vals=np.ascontiguousarray(p.vals, np.float64),
        fmt=np.ascontiguousarray(p.fmt, np.int8),
        y=np.ascontiguousarray(p.y, np.float64),
        offs=np.ascontiguousarray(p.offs, np.float64),
        strata=np.ascontiguousarray(p.strata, np.int32),
    )
    desc = DataDesc(p.model, precision, p.K, p.J, _p(a["col_ptr"]), _p(a["rows"]),
                    _p(a["vals"]), _p(a["fmt"]), _p(a["y"]), _p(a["offs"]), _p(a["strata"]),
                    int(p.has_offset))
    return desc, a


class Data:
    """A problem resident in GPU memory (ccdh_data_t)."""

    def __init__(self, p: CcdProblem, precision=FP64):
        lib = load_library()
        self.problem = p
        self.precision = precision
        desc, self._arrs = _desc(p, precision)
        h = ctypes.c_void_p()
        _check(lib.ccdh_data_create(ctypes.byref(desc), ctypes.byref(h)))
        self.handle = h
        self._arrs = None  # the library copied everything

    @classmethod
    def from_long(cls, model, row_id, y, covariate_id, entry_row_id, value=None, stratum_id=None,
                  time=None, add_intercept=True, precision=FP64):
        """A data set from long-format outcomes and covariate entries, sorted and
        mapped on the device (ccdh_data_create_long)."""
        lib = load_library()
        a = dict(row_id=np.ascontiguousarray(row_id, np.int64),
                 y=np.ascontiguousarray(y, np.float64),
                 cov=np.ascontiguousarray(covariate_id, np.int64),
                 erow=np.ascontiguousarray(entry_row_id, np.int64),
                 val=None if value is None else np.ascontiguousarray(value, np.float64),
                 strata=None if stratum_id is None else np.ascontiguousarray(stratum_id, np.int64),
                 time=None if time is None else np.ascontiguousarray(time, np.float64))
        desc = LongDesc(model, precision, len(a["row_id"]), _p(a["row_id"]), _p(a["strata"]),
                        _p(a["y"]), _p(a["time"]), len(a["cov"]), _p(a["cov"]), _p(a["erow"]),
                        _p(a["val"]), int(add_intercept))
        h = ctypes.c_void_p()
        _check(lib.ccdh_data_create_long(ctypes.byref(desc), ctypes.byref(h)))
        self = cls.__new__(cls)
        self.handle = h
        self.precision = precision
        self._arrs = None
        self.problem = self.to_problem()
        return self

    def to_problem(self):
        """The resident data set's compressed columns as a CcdProblem."""
        lib = load_library()
        m, K, J, off = (ctypes.c_int() for _ in range(4))
        _check(lib.ccdh_data_shape(self.handle, ctypes.byref(m), ctypes.byref(K), ctypes.byref(J),
                                   ctypes.byref(off)))
        col_ptr = np.zeros(J.value + 1, np.int64)
        _check(lib.ccdh_data_get_csc(self.handle, _p(col_ptr), None, None, None, None, None, None))
        nnz = int(col_ptr[-1])
        rows, vals = np.zeros(nnz, np.int32), np.zeros(nnz)
        fmt = np.zeros(J.value, np.int8)
        y, offs, strata = np.zeros(K.value), np.zeros(K.value), np.zeros(K.value, np.int32)
        _check(lib.ccdh_data_get_csc(self.handle, None, _p(rows), _p(vals), _p(fmt), _p(y),
                                     _p(offs), _p(strata)))
        return CcdProblem(model=m.value, K=K.value, J=J.value, col_ptr=col_ptr, rows=rows,
                          vals=vals, fmt=fmt, y=y, offs=offs, strata=strata,
                          has_offset=bool(off.value))

    @property
    def device_bytes(self):
        return load_library().ccdh_data_device_bytes(self.handle)

    def tables(self):
        """The data set's derived per-column tables (ccdh_data_get_tables): CLR
        stratum runs, nonzero strata and stratum starts; row shards and each
        column's shard offsets; column blocks (not Cox)."""
        lib = load_library()
        m, K, J, off = (ctypes.c_int() for _ in range(4))
        _check(lib.ccdh_data_shape(self.handle, ctypes.byref(m), ctypes.byref(K), ctypes.byref(J),
                                   ctypes.byref(off)))
        n_seg, n_strata = ctypes.c_int64(), ctypes.c_int()
        _check(lib.ccdh_data_get_tables(self.handle, ctypes.byref(n_seg), ctypes.byref(n_strata),
                                        *([None] * 8)))
        col_ptr = np.zeros(J.value + 1, np.int64)
        _check(lib.ccdh_data_get_csc(self.handle, _p(col_ptr), None, None, None, None, None, None))
        clr, cox = m.value == CONDITIONAL_LOGISTIC, m.value == COX
        t = dict(shard_row=np.zeros(65, np.int32), col_shard=np.zeros(J.value * 65, np.int32))
        if clr:
            t.update(seg_ptr=np.zeros(J.value + 1, np.int64), seg_off=np.zeros(n_seg.value, np.int32),
                     seg_key=np.zeros(n_seg.value, np.int32),
                     nz_key=np.zeros(int(col_ptr[-1]), np.int32),
                     strat_start=np.zeros(n_strata.value + 1, np.int32))
        if not cox:
            t["blk"] = np.zeros(J.value, np.int8)
        g = lambda k: _p(t[k]) if k in t else None  # noqa: E731
        _check(lib.ccdh_data_get_tables(self.handle, None, None, g("seg_ptr"), g("seg_off"),
                                        g("seg_key"), g("nz_key"), g("strat_start"),
                                        g("shard_row"), g("col_shard"), g("blk")))
        return t

    def close(self):
        if self.handle:
            load_library().ccdh_data_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _spec_struct(spec: FitSpec, keep, heldout=None):
    arrs = dict(
        pt=np.ascontiguousarray(spec.prior_type, np.int32),
        pv=np.ascontiguousarray(spec.prior_var, np.float64),
        fx=np.ascontiguousarray(spec.fixed, np.uint8),
        b0=np.ascontiguousarray(spec.beta0, np.float64),
        w=None if spec.weights is None else np.ascontiguousarray(spec.weights, np.float64),
        ho=None if heldout is None else np.ascontiguousarray(heldout, np.float64),
    )
    keep.append(arrs)
    return FitSpecC(_p(arrs["pt"]), _p(arrs["pv"]), _p(arrs["fx"]), _p(arrs["b0"]),
                    _p(arrs["w"]), _p(arrs["ho"]), spec.tolerance, spec.max_iterations,
                    spec.convergence_type, spec.initial_bound, spec.max_bound_count)


def fit_plain(data: Data, spec: FitSpec, heldout=None, algorithm="ccd"):
    """One fit with the host driver: "ccd" steps one coordinate at a time through
    per-coordinate kernels; "mm" steps every coordinate at once from the MM
    gradient (logistic and conditional logistic models) and raises NonIncreasing
    when the log posterior decreases."""
    lib = load_library()
    keep = []
    s = _spec_struct(spec, keep, heldout)
    beta = np.zeros(data.problem.J)
    res = FitResult()
    fn = {"ccd": lib.ccdh_fit_plain, "mm": lib.ccdh_fit_mm}[algorithm]
    rc = fn(data.handle, ctypes.byref(s), _p(beta), ctypes.byref(res))
    if rc == -5:
        raise NonIncreasing(lib.ccdh_last_error().decode(), beta, res.iterations)
    _check(rc)
    return beta, res.as_dict()


def mm_last_stats():
    """Event-timed device time of the last MM fit's chunk pass and xBeta kernels."""
    st = MmStats()
    _check(load_library().ccdh_mm_last_stats(ctypes.byref(st)))
    return st.as_dict()


def fit_batch(data: Data, specs, heldouts=None, threads=0, sweeps_per_sync=0, use_graph=1,
              max_gang=0, gang_min_nnz=0, pass_items=0, autotune=0, quantum_ms=0.0, bundle=0,
              solo=0, spin_limit=0):
    """Many fits over one resident data set with device-resident sweeps; a fit
    runs on up to `max_gang` workgroups when fewer fits than CUs remain (solo=-1:
    a gang sweeps every column together, none on its lead alone; spin_limit: polls a
    gang rank waits for its peers before it gives up and the fit restarts on one
    workgroup, 0 for the default)."""
    lib = load_library()
    n = len(specs)
    keep = []
    arr = (FitSpecC * n)()
    for i, sp in enumerate(specs):
        arr[i] = _spec_struct(sp, keep, None if heldouts is None else heldouts[i])
    betas = np.zeros((n, data.problem.J))
    res = (FitResult * n)()
    opt = BatchOptions(threads, sweeps_per_sync, use_graph, max_gang, gang_min_nnz, pass_items,
                       autotune, quantum_ms, bundle, solo, spin_limit)
    _check(lib.ccdh_batch_fit(data.handle, n, arr, _p(betas), res, ctypes.byref(opt)))
    return betas, [r.as_dict() for r in res]


def batch_last_stats():
    st = BatchStats()
    _check(load_library().ccdh_batch_last_stats(ctypes.byref(st)))
    return st.as_dict()


def cv_fold_weights(ids, n_fold, seed, task, by_pid=True):
    """Fold weights of CV task `task` (fold = task % n_fold), 0 for held-out units."""
    ids = np.ascontiguousarray(ids, np.int32)
    out = np.zeros(len(ids))
    _check(load_library().ccdh_cv_fold_weights(_p(ids), len(ids), int(by_pid), n_fold, seed,
                                               task, _p(out)))
    return out


def bootstrap_weights(ids, seed, replicate, exclude=None):
    """Weights of bootstrap replicate `replicate`: per row, how many times its
    unit (ids: stratum / pid of each row) was drawn."""
    ids = np.ascontiguousarray(ids, np.int32)
    ex = None if exclude is None else np.ascontiguousarray(exclude, np.float64)
    out = np.zeros(len(ids))
    _check(load_library().ccdh_bootstrap_weights(_p(ids), len(ids), seed, replicate, _p(ex),
                                                 0 if ex is None else len(ex), _p(out)))
    return out


def bootstrap_batch(data: Data, spec: FitSpec, ids, replicates, seed, warm_start=None, **opts):
    """Bootstrap replicates (BootstrapDriver::drive, drivers/BootstrapDriver.cpp:47-71)
    as one batch: replicate r refits `spec` under the weights of bootstrap sample r,
    every replicate starting from `warm_start` (default spec.beta0).  The reference
    chains its warm starts from replicate to replicate, so its estimates agree
    with these to the convergence tolerance rather than bit for bit.  Returns
    (betas [replicates x J], results)."""
    b0 = np.array(spec.beta0 if warm_start is None else warm_start, dtype=np.float64)
    specs = [dataclasses.replace(spec, weights=bootstrap_weights(ids, seed, r), beta0=b0.copy())
             for r in range(replicates)]
    return fit_batch(data, specs, **opts)


def profile_batch(data: Data, spec: FitSpec, index, values, mode_beta, **opts):
    """Profile log likelihood (OptimizationProfile::objective, CcdInterface.cpp:223-260)
    at many values of coordinate `index` as one batch: each fit holds beta[index]
    fixed at a value and starts from the mode with that entry replaced (the
    reference evaluates the values one after another from the previous one's
    estimate; the optimum per value is the same).  Returns (log likelihoods,
    betas, results)."""
    specs = []
    for x in values:
        b0 = np.array(mode_beta, dtype=np.float64)
        b0[index] = x
        fx = np.array(spec.fixed, dtype=np.uint8)
        fx[index] = 1
        specs.append(dataclasses.replace(spec, beta0=b0, fixed=fx))
    betas, res = fit_batch(data, specs, **opts)
    return np.array([r["log_likelihood"] for r in res]), betas, res


def cox_groups(data: Data, weights=None, noalloc=False):
    """Risk-set groups of a Cox data set under `weights`, built on the device:
    (pid [K], reset [N+1], first row [N+1], first group of each row shard [65], N).
    noalloc: through ccdh_cox_groups_noalloc (no device allocation, at most
    1,048,576 rows)."""
    K = data.problem.K
    w = None if weights is None else np.ascontiguousarray(weights, np.float64)
    pid = np.zeros(K, np.int32)
    reset = np.zeros(K + 1, np.uint8)
    first = np.zeros(K + 1, np.int32)
    shard = np.zeros(65, np.int32)
    n = ctypes.c_int()
    lib = load_library()
    fn = lib.ccdh_cox_groups_noalloc if noalloc else lib.ccdh_cox_groups
    _check(fn(data.handle, _p(w), _p(pid), _p(reset), _p(first), _p(shard), ctypes.byref(n)))
    N = n.value
    return pid, reset[:N + 1], first[:N + 1], shard, N


def search_step(tried, stdstep=10.0, stop_by_y=0.01, stop_by_x=np.log(1.5), first_cut=1.0):
    """Next point of the unimodal hyperparameter search over `tried` = [(x, mean,
    stddev), ...] in the order tried: (continue, next x, expected value)."""
    t = np.asarray(tried, dtype=np.float64).reshape(-1, 3)
    xs, ys, sd = (np.ascontiguousarray(t[:, i]) for i in range(3))
    cont, nxt, exp_ = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
    _check(load_library().ccdh_search_step(_p(xs), _p(ys), _p(sd), len(t), stdstep, stop_by_y,
                                           stop_by_x, first_cut, ctypes.byref(cont),
                                           ctypes.byref(nxt), ctypes.byref(exp_)))
    return bool(cont.value), nxt.value, exp_.value


def search_quadratic(tried):
    t = np.asarray(tried, dtype=np.float64).reshape(-1, 3)
    xs, ys, sd = (np.ascontiguousarray(t[:, i]) for i in range(3))
    c = np.zeros(3)
    _check(load_library().ccdh_search_quadratic(_p(xs), _p(ys), _p(sd), len(t), _p(c)))
    return c


def auto_search_cv(data: Data, spec: FitSpec, ids, n_fold=10, seed=123, start_variance=1.0,
                   max_steps=10, by_pid=True, threads=1, reset_coefficients=True, **opts):
    """Auto-search cross-validation over one prior variance
    (AutoSearchCrossValidationDriver::doCrossValidationLoop,
    drivers/AutoSearchCrossValidationDriver.cpp:63-165, nDim = 1): each step fits
    the n_fold folds at the current variance, scores each by its held-out
    predictive log likelihood (a fold that does not converge scores NaN and is
    left out), and the unimodal search (UniModalSearch(10, 0.01, log 1.5))
    proposes the next variance, until it converges or max_steps.  Starts as in
    grid_search_cv: reset_coefficients=True, every fold from zero (each step one
    batch); False, the reference's warm chains over `threads` threads.
    Returns (variance, expected score, trace [(variance, mean, stddev)])."""
    chain = _CvChains(data, spec, ids, n_fold, seed, by_pid, threads, reset_coefficients, None)
    var = float(start_variance)
    tried = []
    expected = float("nan")
    for _ in range(max_steps):
        pred, _ = chain.step(var, **opts)
        ok = pred[pred == pred]
        mean = float(ok.mean())
        sd = float(np.sqrt(np.mean(ok * ok) - mean * mean))
        tried.append((var, mean, sd))
        cont, var, expected = search_step(tried)
        if not cont:
            break
    return var, expected, tried


def cv_chains(n_fold, threads):
    """TaskScheduler's split of a grid point's n_fold tasks over `threads` threads
    (Thread.h:64-80): thread t runs tasks [t*c, (t+1)*c) in order, c = ceil(n_fold /
    threads) (fewer threads when the tasks run out)."""
    T = max(1, min(threads, n_fold))
    c = n_fold // T + (n_fold % T != 0)
    return [list(range(t * c, min(n_fold, (t + 1) * c))) for t in range(T) if t * c < n_fold]


class _CvChains:
    """The fold tasks of cross-validation steps (AbstractCrossValidationDriver::
    doCrossValidationStep, AbstractCrossValidationDriver.cpp:131-248) and the
    coefficients each thread's CCD carries from task to task."""

    def __init__(self, data, spec, ids, n_fold, seed, by_pid, threads, reset, start):
        self.data, self.spec, self.n_fold, self.reset = data, spec, n_fold, reset
        self.weights = [cv_fold_weights(ids, n_fold, seed, f, by_pid=by_pid) for f in range(n_fold)]
        self.zero = np.zeros(data.problem.J)
        if data.problem.has_offset:
            self.zero[0] = spec.beta0[0]
        self.chains = [list(range(n_fold))] if reset else cv_chains(n_fold, threads)
        self.beta = [np.array(spec.beta0 if start is None else start, np.float64)
                     for _ in self.chains]

    def step(self, var, **opts):
        """Fit every fold at prior variance `var`; (scores [n_fold], results)."""
        pt = np.asarray(self.spec.prior_type)
        pv = np.where(pt != 0, var, np.asarray(self.spec.prior_var))
        n = self.n_fold
        scores, results = np.full(n, np.nan), [None] * n
        if self.reset:
            rounds = [[(0, f) for f in range(n)]]
        else:
            depth = max(len(c) for c in self.chains)
            rounds = [[(t, c[k]) for t, c in enumerate(self.chains) if k < len(c)]
                      for k in range(depth)]
        for rnd in rounds:
            specs = [dataclasses.replace(self.spec, prior_var=pv, weights=self.weights[f],
                                         beta0=(self.zero if self.reset else self.beta[t]).copy())
                     for t, f in rnd]
            b, res = fit_batch(self.data, specs, heldouts=[1.0 - self.weights[f] for _, f in rnd],
                               **opts)
            for i, (t, f) in enumerate(rnd):
                results[f] = res[i]
                ok = res[i]["return_flag"] == 0
                if ok:
                    scores[f] = res[i]["pred_log_likelihood"]
                if not self.reset:
                    self.beta[t] = b[i].copy() if ok else self.zero.copy()
        return scores, results


def grid_search_cv(data: Data, spec: FitSpec, ids, variances, n_fold=10, seed=123, by_pid=True,
                   threads=1, reset_coefficients=False, start=None, **opts):
    """Grid-search cross-validation (GridSearchCrossValidationDriver::doCrossValidationLoop,
    GridSearchCrossValidationDriver.cpp:96-140, with AbstractCrossValidationDriver::
    doCrossValidationStep, AbstractCrossValidationDriver.cpp:131-248) over the prior
    variances `variances`, one repetition of n_fold folds (fold weights as
    cv_fold_weights; held-out scores on the complement).

    reset_coefficients=False is the reference's default (createControl(resetCoefficients
    = FALSE), R/ModelFit.R:506): the tasks run in the reference's thread chains
    (cv_chains): every task starts from the estimate its chain's previous task
    reached -- across grid points too -- and a task that does not end with SUCCESS
    leaves its chain at zero (resetBeta, AbstractCrossValidationDriver.cpp:219-225),
    its score NaN.  The tasks at the same position of every chain run as one batch.
    reset_coefficients=True starts every task from zero and runs each grid point's
    tasks as one batch.  start: the chains' first coefficients (default spec.beta0).

    Returns (point estimates [grid] (mean of the scores that are not NaN), scores
    [grid x n_fold], results [grid x n_fold], final coefficients of each chain)."""
    chain = _CvChains(data, spec, ids, n_fold, seed, by_pid, threads, reset_coefficients, start)
    scores = np.full((len(variances), n_fold), np.nan)
    results = []
    for g, var in enumerate(variances):
        scores[g], res = chain.step(var, **opts)
        results.append(res)
    with np.errstate(invalid="ignore"):
        point = np.array([np.nanmean(r) if np.any(r == r) else np.nan for r in scores])
    return point, scores, results, chain.beta


def cv_grid_point(lower, upper, size, step):
    return load_library().ccdh_cv_grid_point(lower, upper, size, step)


class FitHandle:
    """Direct access to the per-coordinate operations (AbstractModelSpecifics)."""

    def __init__(self, data: Data):
        self.lib = load_library()
        self.data = data
        h = ctypes.c_void_p()
        _check(self.lib.ccdh_fit_create(data.handle, ctypes.byref(h)))
        self.handle = h

    def set_weights(self, w=None, use_cv=False):
        wa = None if w is None else np.ascontiguousarray(w, np.float64)
        _check(self.lib.ccdh_set_weights(self.handle, _p(wa), int(use_cv)))

    def fixed_terms(self, use_cv=False):
        _check(self.lib.ccdh_fixed_terms_loglik(self.handle, int(use_cv)))
        _check(self.lib.ccdh_fixed_terms_grad_hess(self.handle, int(use_cv)))

    def compute_xbeta(self, beta):
        b = np.ascontiguousarray(beta, np.float64)
        _check(self.lib.ccdh_compute_xbeta(self.handle, _p(b)))

    def remaining_stats(self, use_w=False):
        _check(self.lib.ccdh_remaining_stats(self.handle, int(use_w)))

    def mm_grad_hess(self, fixed=None, use_w=False):
        """(gradients, Hessians) of every column from the MM increments."""
        J = self.data.problem.J
        fx = np.zeros(J, np.uint8) if fixed is None else np.ascontiguousarray(fixed, np.uint8)
        gh = np.zeros(2 * J)
        _check(self.lib.ccdh_mm_grad_hess(self.handle, _p(fx), int(use_w), _p(gh)))
        return gh[0::2].copy(), gh[1::2].copy()

    def grad_hess(self, j, use_w=False):
        _check(self.lib.ccdh_numerator(self.handle, j, int(use_w)))
        g, h = ctypes.c_double(), ctypes.c_double()
        _check(self.lib.ccdh_grad_hess(self.handle, j, int(use_w), ctypes.byref(g),
                                       ctypes.byref(h)))
        return g.value, h.value

    def update(self, delta, j, use_w=False):
        _check(self.lib.ccdh_update_xbeta(self.handle, float(delta), j, int(use_w)))

    def loglik(self, use_cv=False):
        v = ctypes.c_double()
        _check(self.lib.ccdh_loglik(self.handle, int(use_cv), ctypes.byref(v)))
        return v.value

    def gradient_objective(self, use_cv=False):
        v = ctypes.c_double()
        _check(self.lib.ccdh_gradient_objective(self.handle, int(use_cv), ctypes.byref(v)))
        return v.value

    def pred_loglik(self, w):
        wa = np.ascontiguousarray(w, np.float64)
        v = ctypes.c_double()
        _check(self.lib.ccdh_pred_loglik(self.handle, _p(wa), ctypes.byref(v)))
        return v.value

    def xbeta(self):
        out = np.zeros(self.data.problem.K)
        _check(self.lib.ccdh_get_xbeta(self.handle, _p(out)))
        return out

    def close(self):
        if self.handle:
            self.lib.ccdh_fit_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching fit server (ccdh_server_*).

    submit() queues a fit over a host problem (uploaded when admitted) or over a
    resident Data; wait() returns (beta, result, info) or raises the request's
    error.  The host arrays of a submitted problem are kept here until the
    request is collected."""

    def __init__(self, precision=FP64, max_slots=0, threads=0, quantum_ms=0.0,
                 hbm_budget_bytes=0, max_gang=0, spin_limit=0):
        self.lib = load_library()
        opt = ServerOptions(precision, max_slots, threads, max_gang, quantum_ms, hbm_budget_bytes,
                            spin_limit)
        h = ctypes.c_void_p()
        _check(self.lib.ccdh_server_create(ctypes.byref(opt), ctypes.byref(h)))
        self.handle = h
        self.precision = precision
        self._keep = {}

    def submit(self, problem, spec: FitSpec, heldout=None):
        keep = []
        s = _spec_struct(spec, keep, heldout)
        rid = ctypes.c_int64()
        if isinstance(problem, Data):
            _check(self.lib.ccdh_server_submit_resident(self.handle, problem.handle,
                                                        ctypes.byref(s), ctypes.byref(rid)))
            J = problem.problem.J
        else:
            desc, arrs = _desc(problem, self.precision)
            keep.append(arrs)
            _check(self.lib.ccdh_server_submit(self.handle, ctypes.byref(desc), ctypes.byref(s),
                                               ctypes.byref(rid)))
            J = problem.J
        self._keep[rid.value] = (keep, J)
        return rid.value

    def poll(self, rid):
        rc = self.lib.ccdh_server_poll(self.handle, rid)
        if rc < 0:
            _check(rc)
        return rc == 1

    def wait(self, rid, timeout_ms=-1.0):
        keep, J = self._keep[rid]
        beta = np.zeros(J)
        res, info = FitResult(), RequestInfo()
        rc = self.lib.ccdh_server_wait(self.handle, rid, float(timeout_ms), _p(beta),
                                       ctypes.byref(res), ctypes.byref(info))
        if rc == -4 and timeout_ms >= 0 and "timeout" in self.lib.ccdh_last_error().decode():
            _check(rc)  # still in flight: keep its arrays
        del self._keep[rid]
        _check(rc)
        return beta, res.as_dict(), info.as_dict()

    def stats(self):
        st = ServerStats()
        _check(self.lib.ccdh_server_get_stats(self.handle, ctypes.byref(st)))
        return st.as_dict()

    def close(self):
        if self.handle:
            self.lib.ccdh_server_destroy(self.handle)
            self.handle = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
