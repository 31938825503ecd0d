"""The per-column tables a long-format ingestion builds on the device
(ccd_tables.hip) against the tables ccdh_data_create builds on the host from the
same compressed columns: conditional-logistic stratum runs, nonzero strata and
stratum starts, the 64 row shards and every column's shard offsets, and the column
blocks -- identical, for every model, on the golden fixtures and on synthetic
designs that reach the tables' edge cases (one stratum holding every row, fewer
rows than shards, long tie groups of Cox events, columns over the block limit,
thousands of short columns).  Also the device sorts behind the ingestion on keys
over their whole ranges (negative and extreme ids, negative, signed-zero and tied
times), against the oracle's restatement of the loaders.
"""
import numpy as np
import pytest

from ccd_amd import api, synth
from ccd_amd.data import (CONDITIONAL_LOGISTIC, COX, DENSE, INTERCEPT, LOGISTIC, POISSON,
                          SPARSE, CcdProblem)
from oracle.ingest import ingest
from tests import fixtures
from tests.test_ingest import assert_same_problem

pytestmark = pytest.mark.gpu


def _both(p, precision=api.FP64):
    """Tables of p ingested on the device, and of the ingested columns on the host."""
    long = synth.to_long(p)
    if p.model == POISSON and "time" not in long:
        long["time"] = np.ones(p.K)  # an exposure of 1: a zero offset column
    d = api.Data.from_long(**long, precision=precision)
    try:
        q = d.problem
        dev = d.tables()
    finally:
        d.close()
    h = api.Data(q, precision=precision)
    try:
        host = h.tables()
    finally:
        h.close()
    return q, dev, host


def _same(dev, host):
    assert sorted(dev) == sorted(host)
    for k in host:
        np.testing.assert_array_equal(dev[k], host[k], err_msg=k)


def _clr(K_strata, P, per, seed, single=False):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(2, 9, K_strata) if not single else np.array([K_strata * 5])
    K = int(sizes.sum())
    strata = np.repeat(np.arange(len(sizes)), sizes).astype(np.int32)
    cols = []
    for j in range(P):
        n = int(min(K, max(1, rng.zipf(1.6) * per)))
        cols.append(np.sort(rng.choice(K, n, replace=False)))
    col_ptr = np.concatenate([[0], np.cumsum([len(c) for c in cols])]).astype(np.int64)
    rows = np.concatenate(cols).astype(np.int32)
    y = np.zeros(K)
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    y[starts] = 1.0
    return CcdProblem(CONDITIONAL_LOGISTIC, K, P, col_ptr, rows, np.ones(len(rows)),
                      np.full(P, 2, np.int8), y, np.zeros(K), strata).validate()


def _dense_as_sparse(p):
    """p with its DENSE covariate columns (not the Poisson offset) stored as SPARSE
    columns of their nonzero entries: long-format input has no dense columns."""
    cols, rows, vals, fmt = [0], [], [], p.fmt.copy()
    for j in range(p.J):
        a, b = p.col_ptr[j], p.col_ptr[j + 1]
        r, v = p.rows[a:b], p.vals[a:b]
        if p.fmt[j] == DENSE and not (p.has_offset and j == 0):
            r, v = np.arange(p.K)[v != 0], v[v != 0]
            fmt[j] = SPARSE
        rows.append(r)
        vals.append(v)
        cols.append(cols[-1] + len(r))
    return CcdProblem(p.model, p.K, p.J, np.array(cols, np.int64),
                      np.concatenate(rows).astype(np.int32), np.concatenate(vals), fmt, p.y,
                      p.offs, p.strata, has_offset=p.has_offset).validate()


@pytest.mark.parametrize("name", fixtures.names())
def test_tables_golden_fixtures(name):
    p, _ = fixtures.load(name)
    q, dev, host = _both(_dense_as_sparse(p))
    _same(dev, host)


SYNTH = [
    ("clr_many_short_columns", lambda: _clr(3000, 2500, 3, 1)),
    ("clr_heavy_columns", lambda: _clr(2000, 300, 40, 2)),
    ("clr_one_stratum", lambda: _clr(4000, 50, 30, 3, single=True)),
    ("clr_few_rows", lambda: _clr(5, 12, 2, 4)),
    ("lr", lambda: synth.lr_indicator(K=5000, P=3000, p=0.002, n_true=20, seed=5)[0]),
    ("lr_tiny", lambda: synth.lr_indicator(K=40, P=30, p=0.1, n_true=3, seed=6)[0]),
    ("poisson", lambda: synth.poisson_zipf(K=20000, P=2000, seed=7)[0]),
    ("cox_days", lambda: synth.cox_zipf(K=20000, P=500, per_row=10, seed=8)[0]),
    # times of 1..5 days: tie groups of thousands of events, which shard bounds may not split
    ("cox_long_ties", lambda: synth.cox_zipf(K=20000, P=300, per_row=10, n_true=50,
                                             rate=0.5, censor_max=5.0, seed=10)[0]),
]


@pytest.mark.parametrize("name,make", SYNTH, ids=[s[0] for s in SYNTH])
def test_tables_synthetic(name, make):
    p = make()
    q, dev, host = _both(p)
    _same(dev, host)
    assert_same_problem(q, ingest(**synth.to_long(p)))
    if "blk" in host and name.startswith("clr_many"):
        assert (host["blk"] >= 2).any()  # blocks occur


def test_tables_mixed_precision():
    p = _clr(1000, 400, 4, 9)
    q, dev, host = _both(p, precision=api.MIXED)
    _same(dev, host)


def test_ingest_key_ranges_against_oracle():
    rng = np.random.default_rng(77)
    K = 5000
    # row ids spread over the whole int64 range, negative ones included
    row_id = rng.choice(np.array([-(2**62), -(2**40), -7, 0, 3, 2**33, 2**62 - 1, 2**63 - 1],
                                 np.int64), 8, replace=False)
    row_id = np.concatenate([row_id, rng.integers(-(2**60), 2**60, K - 8, dtype=np.int64)])
    row_id = np.unique(row_id)
    rng.shuffle(row_id)
    K = len(row_id)
    # times with ties, negatives and both signed zeros; strata with negative ids
    time = rng.choice(np.array([-3.5, -0.0, 0.0, 1.0, 2.5, 1e300, -1e-300]), K)
    y = rng.integers(0, 2, K).astype(np.float64)
    stratum = rng.choice(np.array([-(2**50), -1, 0, 7, 2**45]), K)
    m = 30000
    cov = rng.choice(np.array([-(2**61), -5, 0, 1, 2**61, 12345], np.int64), m)
    erow = rng.choice(row_id, m)
    pairs = np.unique(np.stack([cov, erow], 1), axis=0)
    cov, erow = pairs[:, 0], pairs[:, 1]
    o = rng.permutation(len(cov))
    cov, erow = cov[o], erow[o]
    val = rng.choice(np.array([1.0, 1.0, 0.0, -2.0]), len(cov))
    case = dict(model=COX, row_id=row_id, y=y, time=time, stratum_id=stratum, covariate_id=cov,
                entry_row_id=erow, value=val)
    d = api.Data.from_long(**case)
    try:
        q = d.problem
        dev = d.tables()
    finally:
        d.close()
    assert_same_problem(q, ingest(**case))
    h = api.Data(q)
    try:
        _same(dev, h.tables())
    finally:
        h.close()
    for model in (LOGISTIC, POISSON, CONDITIONAL_LOGISTIC):
        c = dict(case, model=model)
        if model == POISSON:
            c["time"] = np.abs(time) + 0.5
        if model != CONDITIONAL_LOGISTIC:
            c.pop("stratum_id")
        if model == LOGISTIC:
            c.pop("time")
        d = api.Data.from_long(**c)
        try:
            q = d.problem
        finally:
            d.close()
        assert_same_problem(q, ingest(**c))
        assert q.J == len(np.unique(cov)) + (2 if model == POISSON else
                                                       1 if model == LOGISTIC else 0)
        assert np.any(q.fmt == SPARSE) and (model == CONDITIONAL_LOGISTIC or q.fmt[
            1 if model == POISSON else 0] == INTERCEPT)
