"""Host-side problem containers mirroring Cyclops' ModelData / CompressedDataMatrix.

A problem is a column-major compressed matrix X (one column per covariate,
each with a format tag) plus the outcome vectors:

* ``fmt[j]`` follows ``FormatType`` (src/cyclops/CompressedDataMatrix.h:36-38):
  DENSE=0, SPARSE=1, INDICATOR=2, INTERCEPT=3.  INDICATOR/INTERCEPT columns carry
  no values (x = 1).  DENSE/INTERCEPT columns span all K rows.
* ``y`` outcome, ``offs`` the time vector (event time for Cox), ``strata`` the
  stratum of each row (ModelData pid, src/cyclops/ModelData.h:748-751).
* Rows of grouped models are sorted by stratum; Cox rows by (stratum, -time, y)
  as R/NewDataConversion.R:162 sorts them.
* An offset enters as DENSE column 0 with its coefficient fixed at 1
  (src/RcppModelData.cpp:795-802, src/cyclops/CyclicCoordinateDescent.cpp:218-221).

Model ids follow ``ModelType`` (src/cyclops/Types.h:146-159).
"""
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

DENSE, SPARSE, INDICATOR, INTERCEPT = 0, 1, 2, 3

POISSON = 2
LOGISTIC = 3
CONDITIONAL_LOGISTIC = 4
COX = 8
MODEL_NAMES = {POISSON: "poisson", LOGISTIC: "logistic",
               CONDITIONAL_LOGISTIC: "conditional_logistic", COX: "cox"}

PRIOR_NONE, PRIOR_LAPLACE, PRIOR_NORMAL = 0, 1, 2
GRADIENT, LANGE, MITTAL, ONE_STEP = 0, 1, 2, 3
SUCCESS, FAIL, MAX_ITERATIONS, ILLCONDITIONED = 0, 1, 2, 3


@dataclass
class CcdProblem:
    model: int
    K: int
    J: int
    col_ptr: np.ndarray          # int64[J+1]
    rows: np.ndarray             # int32[nnz]
    vals: np.ndarray             # float64[nnz] (entries of INDICATOR columns unused)
    fmt: np.ndarray              # int8[J]
    y: np.ndarray                # float64[K]
    offs: np.ndarray             # float64[K]
    strata: np.ndarray           # int32[K]
    has_offset: bool = False
    labels: Optional[list] = None

    @property
    def nnz(self):
        return int(self.col_ptr[-1])

    def column(self, j):
        a, b = int(self.col_ptr[j]), int(self.col_ptr[j + 1])
        r = self.rows[a:b]
        if self.fmt[j] in (INDICATOR, INTERCEPT):
            v = np.ones(b - a)
        else:
            v = self.vals[a:b]
        return r, v

    def to_dense(self):
        X = np.zeros((self.K, self.J))
        for j in range(self.J):
            r, v = self.column(j)
            X[r, j] = v
        return X

    def validate(self):
        assert self.col_ptr.shape == (self.J + 1,)
        assert self.rows.shape[0] == self.nnz and self.vals.shape[0] == self.nnz
        for j in range(self.J):
            a, b = int(self.col_ptr[j]), int(self.col_ptr[j + 1])
            r = self.rows[a:b]
            if self.fmt[j] in (DENSE, INTERCEPT):
                assert b - a == self.K and np.array_equal(r, np.arange(self.K))
            elif b - a > 1:
                assert np.all(np.diff(r) > 0), "row indices must be strictly increasing"
        if self.model in (CONDITIONAL_LOGISTIC, COX):
            assert np.all(np.diff(self.strata) >= 0), "rows must be sorted by stratum"
        return self


@dataclass
class FitSpec:
    prior_type: np.ndarray
    prior_var: np.ndarray
    fixed: np.ndarray
    beta0: np.ndarray
    weights: Optional[np.ndarray] = None
    tolerance: float = 1e-6
    max_iterations: int = 1000
    convergence_type: int = GRADIENT
    initial_bound: float = 2.0
    max_bound_count: int = 5


def build_problem(model, columns, y, offs=None, strata=None, offset=None, intercept=False,
                  labels=None):
    """Build a problem from a list of columns.

    Each column is ``(fmt, rows, values)``; for DENSE pass ``values`` of length K and
    ``rows=None``; for INDICATOR pass ``values=None``.  ``offset`` (length K) becomes
    DENSE column 0 with a fixed coefficient; ``intercept`` adds an INTERCEPT column
    after it.
    """
    y = np.asarray(y, dtype=np.float64)
    K = y.shape[0]
    cols = []
    if offset is not None:
        cols.append((DENSE, None, np.asarray(offset, dtype=np.float64)))
    if intercept:
        cols.append((INTERCEPT, None, None))
    cols.extend(columns)
    J = len(cols)
    col_ptr = np.zeros(J + 1, dtype=np.int64)
    rows_l, vals_l, fmt = [], [], np.zeros(J, dtype=np.int8)
    for j, (f, r, v) in enumerate(cols):
        fmt[j] = f
        if f in (DENSE, INTERCEPT):
            r = np.arange(K, dtype=np.int32)
            v = np.ones(K) if f == INTERCEPT else np.asarray(v, dtype=np.float64)
        else:
            r = np.asarray(r, dtype=np.int32)
            order = np.argsort(r, kind="stable")
            r = r[order]
            v = np.ones(len(r)) if (f == INDICATOR or v is None) else np.asarray(v, np.float64)[order]
        rows_l.append(r)
        vals_l.append(v)
        col_ptr[j + 1] = col_ptr[j] + len(r)
    rows = np.concatenate(rows_l).astype(np.int32) if rows_l else np.zeros(0, np.int32)
    vals = np.concatenate(vals_l).astype(np.float64) if vals_l else np.zeros(0)
    offs_a = np.zeros(K) if offs is None else np.asarray(offs, dtype=np.float64)
    strata_a = np.zeros(K, dtype=np.int32) if strata is None else np.asarray(strata, np.int32)
    return CcdProblem(model, K, J, col_ptr, rows, vals, fmt, y, offs_a, strata_a,
                      has_offset=offset is not None, labels=labels).validate()


def default_spec(p, prior=PRIOR_NONE, variance=1.0, exclude=(), weights=None, **kw):
    """Prior on every covariate except the offset, the intercept and ``exclude``
    (those get no prior: src/RcppCyclopsInterface.cpp:834-847, R/ModelFit.R:120-138)."""
    J = p.J
    pt = np.full(J, prior, dtype=np.int32)
    pv = np.full(J, variance, dtype=np.float64)
    fixed = np.zeros(J, dtype=np.uint8)
    beta0 = np.zeros(J)
    for j in range(J):
        if p.fmt[j] == INTERCEPT or j in exclude:
            pt[j] = PRIOR_NONE
    if p.has_offset:
        pt[0] = PRIOR_NONE
        fixed[0] = 1
        beta0[0] = 1.0
    return FitSpec(pt, pv, fixed, beta0, weights=weights, **kw)


def sort_cox(time, y, strata, row_id=None):
    """Row order of R/NewDataConversion.R:162: stratum, -time, y, rowId."""
    n = len(time)
    rid = np.arange(n) if row_id is None else row_id
    return np.lexsort((rid, y, -np.asarray(time), strata))
