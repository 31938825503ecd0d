"""Pins of the oracle's long-format ingestion (oracle/ingest.py) -- CPU only.

  * round trip: a problem with known compressed columns, written out as long
    input (distinct row ids rising with the row, so every model's row order is
    the problem's own; covariate ids rising with the column; outcome rows and
    entries shuffled; zero-valued entries and entries of unknown rows added),
    comes back exactly -- columns, rows, values, formats, outcomes, times,
    strata, the intercept and the Poisson log-time offset;
  * the loaders' rules on hand-made inputs: the INDICATOR -> SPARSE upcast,
    zero values left out, the Cox row order (stratum, time descending, y, row
    id) against the sort the problems are built with, stratum numbering, and
    the errors the reference raises.
"""
import numpy as np
import pytest

from ccd_amd import synth
from ccd_amd.data import (CONDITIONAL_LOGISTIC, COX, DENSE, INDICATOR, INTERCEPT, LOGISTIC,
                          POISSON, SPARSE, build_problem, sort_cox)
from oracle.ingest import IngestError, ingest


def to_long(p, rng, time=None):
    """Long-format input for problem p (see the module docstring)."""
    K = p.K
    row_id = 1000 + np.cumsum(rng.integers(1, 5, K))
    cols = [j for j in range(p.J) if p.fmt[j] in (INDICATOR, SPARSE)]
    cov_id = 10 * np.cumsum(rng.integers(1, 7, len(cols)))
    cov, erow, val = [], [], []
    for c, j in zip(cov_id, cols):
        r, v = p.column(j)
        cov.append(np.full(len(r), c))
        erow.append(row_id[r])
        val.append(v)
        # zero values at rows the column does not hold: left out
        free = np.setdiff1d(np.arange(K), r)[:2]
        cov.append(np.full(len(free), c))
        erow.append(row_id[free])
        val.append(np.zeros(len(free)))
    # entries whose row id is not an outcome row: left out
    cov.append(cov_id[:3])
    erow.append(np.array([7, 8, 9]))
    val.append(np.array([2.0, 3.0, 4.0]))
    cov, erow, val = (np.concatenate(a) for a in (cov, erow, val))
    e = rng.permutation(len(cov))
    o = rng.permutation(K)
    stratum = None
    if p.model in (CONDITIONAL_LOGISTIC, COX):
        sid = 5 + 3 * np.arange(int(p.strata.max()) + 1)  # rising with the stratum number
        stratum = sid[p.strata][o]
    t = time if time is not None else (p.offs if p.model == COX else None)
    return dict(model=p.model, row_id=row_id[o], y=p.y[o], covariate_id=cov[e],
                entry_row_id=erow[e], value=val[e], stratum_id=stratum,
                time=None if t is None else np.asarray(t)[o])


def assert_same_problem(q, p):
    assert (q.model, q.K, q.J, q.has_offset) == (p.model, p.K, p.J, p.has_offset)
    np.testing.assert_array_equal(q.col_ptr, p.col_ptr)
    np.testing.assert_array_equal(q.rows, p.rows)
    np.testing.assert_array_equal(q.fmt, p.fmt)
    for j in range(p.J):
        if p.fmt[j] in (SPARSE, DENSE):
            a, b = p.col_ptr[j], p.col_ptr[j + 1]
            np.testing.assert_array_equal(q.vals[a:b], p.vals[a:b])
    np.testing.assert_array_equal(q.y, p.y)
    if p.model == COX:
        np.testing.assert_array_equal(q.offs, p.offs)
    if p.model in (CONDITIONAL_LOGISTIC, COX):
        np.testing.assert_array_equal(q.strata, p.strata)


def round_trip_problems():
    rng = np.random.default_rng(31)
    K = 400
    ind = [np.nonzero(rng.random(K) < q)[0] for q in (0.3, 0.05, 0.5)]
    sp_rows = np.nonzero(rng.random(K) < 0.4)[0]
    sp_vals = rng.normal(0, 1, len(sp_rows))
    cols = [(INDICATOR, ind[0], None), (SPARSE, sp_rows, sp_vals), (INDICATOR, ind[1], None),
            (INDICATOR, ind[2], None)]
    out = []
    y_lr = (rng.random(K) < 0.3).astype(float)
    out.append(("lr", build_problem(LOGISTIC, cols, y_lr, intercept=True), None))
    t = rng.uniform(30, 3650, K)
    y_po = rng.poisson(2.0, K).astype(float)
    out.append(("poisson", build_problem(POISSON, cols, y_po, intercept=True, offset=np.log(t)), t))
    strata = np.sort(rng.integers(0, 60, K)).astype(np.int32)
    strata = np.unique(strata, return_inverse=True)[1].astype(np.int32)
    y_clr = (rng.random(K) < 0.25).astype(float)
    out.append(("clr", build_problem(CONDITIONAL_LOGISTIC, cols, y_clr, strata=strata), None))
    time = rng.integers(1, 40, K).astype(float)
    ev = (rng.random(K) < 0.6).astype(float)
    cst = (rng.random(K) < 0.5).astype(np.int32)
    order = sort_cox(time, ev, cst)
    inv = np.empty(K, int)
    inv[order] = np.arange(K)
    ccols = [(f, None if r is None else inv[r], v) for f, r, v in cols]
    out.append(("cox", build_problem(COX, ccols, ev[order], offs=time[order], strata=cst[order]),
                None))
    out.append(("lr_synth", synth.lr_indicator(K=3000, P=300, p=0.02, n_true=20, seed=1)[0], None))
    out.append(("clr_synth", synth.clr_strata(n_strata=800, P=200, per_row=5, seed=3)[0], None))
    out.append(("cox_synth", synth.cox_zipf(K=3000, P=200, per_row=6, n_true=20, seed=4)[0], None))
    return out


@pytest.mark.parametrize("name,p,t", round_trip_problems(),
                         ids=lambda v: v if isinstance(v, str) else "")
def test_ingest_round_trip(name, p, t):
    rng = np.random.default_rng(32)
    long = to_long(p, rng, time=t)
    q = ingest(**long, add_intercept=bool(np.any(p.fmt == INTERCEPT)))
    assert_same_problem(q, p)


def test_ingest_format_upcast_and_zero_values():
    # column 10: 1, 1, 2.5 -> SPARSE; column 20: 0, 1 -> INDICATOR (the zero left
    # out); column 30: only zeros -> an empty INDICATOR column
    q = ingest(LOGISTIC, row_id=[3, 1, 2], y=[0, 1, 0],
               covariate_id=[10, 10, 10, 20, 20, 30],
               entry_row_id=[1, 2, 3, 1, 2, 3], value=[1, 1, 2.5, 0, 1, 0], add_intercept=False)
    np.testing.assert_array_equal(q.fmt, [SPARSE, INDICATOR, INDICATOR])
    np.testing.assert_array_equal(q.col_ptr, [0, 3, 4, 4])
    np.testing.assert_array_equal(q.rows, [0, 1, 2, 1])       # rows ordered by row id
    np.testing.assert_array_equal(q.vals[:3], [1, 1, 2.5])
    np.testing.assert_array_equal(q.y, [1, 0, 0])


def test_ingest_cox_order_and_strata():
    # stratum ids 9 and 4; within a stratum time descending, then y, then row id
    q = ingest(COX, row_id=[1, 2, 3, 4, 5, 6], y=[1, 0, 1, 1, 0, 1],
               time=[5, 5, 7, 5, 2, 5], stratum_id=[9, 9, 9, 4, 4, 4],
               covariate_id=[1, 1], entry_row_id=[6, 3])
    # stratum 4: (t5,y1,r4) (t5,y1,r6) (t2,y0,r5); stratum 9: (t7,y1,r3) (t5,y0,r2) (t5,y1,r1)
    np.testing.assert_array_equal(q.offs, [5, 5, 2, 7, 5, 5])
    np.testing.assert_array_equal(q.y, [1, 1, 0, 1, 0, 1])
    np.testing.assert_array_equal(q.strata, [0, 0, 0, 1, 1, 1])
    np.testing.assert_array_equal(q.rows, [1, 3])


ERROR_CASES = [
    (dict(model=LOGISTIC, row_id=[1, 2], y=[0, 1], covariate_id=[5, 5], entry_row_id=[1, 1],
          value=[1.0, 2.0]), "Repeated row-column entry"),
    (dict(model=LOGISTIC, row_id=[1, 2], y=[0, 1], covariate_id=[5, 5], entry_row_id=[2, 2],
          value=[0.0, 0.0]), "Repeated row-column entry"),
    (dict(model=LOGISTIC, row_id=[1, 1], y=[0, 1], covariate_id=[5], entry_row_id=[1]),
     "repeated outcome row id"),
    (dict(model=POISSON, row_id=[1, 2], y=[0, 1], time=[1.0, 0.0], covariate_id=[5],
          entry_row_id=[1]), "time cannot be non-positive"),
    (dict(model=COX, row_id=[1, 2, 3], y=[0, 1, 2], time=[1.0, 2.0, 3.0], covariate_id=[5],
          entry_row_id=[1]), "only accepts one outcome type"),
    (dict(model=CONDITIONAL_LOGISTIC, row_id=[1, 2], y=[0, 1], covariate_id=[5],
          entry_row_id=[1]), "stratum ids required"),
]


@pytest.mark.parametrize("case,msg", ERROR_CASES)
def test_ingest_errors(case, msg):
    with pytest.raises(IngestError, match=msg):
        ingest(**case)


def test_ingest_no_intercept_for_conditional_models():
    q = ingest(CONDITIONAL_LOGISTIC, row_id=[1, 2], y=[0, 1], stratum_id=[3, 3],
               covariate_id=[5], entry_row_id=[2], add_intercept=True)
    assert q.J == 1 and q.fmt[0] == INDICATOR


EDGE_CASES = [
    # no covariate entries: the intercept only
    ("no_entries", dict(model=LOGISTIC, row_id=[5, 3], y=[1, 0], covariate_id=[],
                        entry_row_id=[])),
    # every entry names an unknown row: no covariate column
    ("unknown_rows", dict(model=LOGISTIC, row_id=[5, 3], y=[1, 0], covariate_id=[7, 8],
                          entry_row_id=[1, 2], value=[1.0, 2.0])),
    # Poisson without covariates: log-time offset and intercept
    ("poisson_offset_only", dict(model=POISSON, row_id=[5, 3], y=[1, 0], time=[2.0, 3.0],
                                 covariate_id=[], entry_row_id=[])),
]


def test_ingest_edge_cases():
    q = ingest(**EDGE_CASES[0][1])
    assert q.J == 1 and list(q.fmt) == [INTERCEPT] and list(q.y) == [0.0, 1.0]
    q = ingest(**EDGE_CASES[1][1])
    assert q.J == 1 and list(q.col_ptr) == [0, 2]
    q = ingest(**EDGE_CASES[2][1])
    assert list(q.fmt) == [DENSE, INTERCEPT] and q.has_offset
    np.testing.assert_array_equal(q.vals[:2], np.log([3.0, 2.0]))
