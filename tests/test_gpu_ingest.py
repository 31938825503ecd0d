"""GPU parity of the long-format ingestion (ccdh_data_create_long) against the
oracle's restatement: the device-built data set (columns, rows, values, formats,
outcomes, times, strata) is identical to the oracle's and, on the round-trip
problems, to the problem written out; the reference's errors are raised with the
same messages; a fit over ingested data matches the oracle fit of the original
problem to 1e-8 with the same iterations.
"""
import numpy as np
import pytest

from ccd_amd import api
from ccd_amd.data import COX, INTERCEPT, LOGISTIC, PRIOR_LAPLACE, default_spec
from oracle import oracle as orc
from oracle.ingest import ingest
from tests.test_gpu_plain import check_fit
from tests.test_ingest import (EDGE_CASES, ERROR_CASES, assert_same_problem, round_trip_problems,
                               to_long)

pytestmark = pytest.mark.gpu

_IDS = lambda v: v if isinstance(v, str) else ""  # noqa: E731


@pytest.mark.parametrize("name,p,t", round_trip_problems(), ids=_IDS)
def test_device_ingest_round_trip(name, p, t):
    long = to_long(p, np.random.default_rng(32), time=t)
    intercept = bool(np.any(p.fmt == INTERCEPT))
    d = api.Data.from_long(**long, add_intercept=intercept)
    try:
        q = d.problem
    finally:
        d.close()
    assert_same_problem(q, p)
    assert_same_problem(q, ingest(**long, add_intercept=intercept))


def test_device_ingest_rules():
    d = api.Data.from_long(LOGISTIC, row_id=[3, 1, 2], y=[0, 1, 0],
                           covariate_id=[10, 10, 10, 20, 20, 30],
                           entry_row_id=[1, 2, 3, 1, 2, 3], value=[1, 1, 2.5, 0, 1, 0],
                           add_intercept=False)
    q = d.problem
    d.close()
    ref = ingest(LOGISTIC, row_id=[3, 1, 2], y=[0, 1, 0], covariate_id=[10, 10, 10, 20, 20, 30],
                 entry_row_id=[1, 2, 3, 1, 2, 3], value=[1, 1, 2.5, 0, 1, 0], add_intercept=False)
    assert_same_problem(q, ref)
    case = dict(row_id=[1, 2, 3, 4, 5, 6], y=[1, 0, 1, 1, 0, 1], time=[5, 5, 7, 5, 2, 5],
                stratum_id=[9, 9, 9, 4, 4, 4], covariate_id=[1, 1], entry_row_id=[6, 3])
    d = api.Data.from_long(COX, **case)
    q = d.problem
    d.close()
    assert_same_problem(q, ingest(COX, **case))


@pytest.mark.parametrize("case,msg", ERROR_CASES)
def test_device_ingest_errors(case, msg):
    with pytest.raises(api.CcdhError, match=msg):
        api.Data.from_long(**case)


def test_fit_over_ingested_data():
    name, p, t = [c for c in round_trip_problems() if c[0] == "clr_synth"][0]
    long = to_long(p, np.random.default_rng(33), time=t)
    d = api.Data.from_long(**long)
    sp = default_spec(p, prior=PRIOR_LAPLACE, variance=0.1)
    bg, ig = api.fit_plain(d, sp)
    bb, ib = api.fit_batch(d, [sp])
    d.close()
    bc, ic = orc.fit(p, sp)
    check_fit(bg, ig, bc, ic)
    check_fit(bb[0], ib[0], bc, ic)


@pytest.mark.parametrize("name,case", EDGE_CASES, ids=[c[0] for c in EDGE_CASES])
def test_device_ingest_edge_cases(name, case):
    d = api.Data.from_long(**case)
    q = d.problem
    d.close()
    assert_same_problem(q, ingest(**case))


def test_fit_over_ingested_mixed_precision_data():
    # FP32 xBeta / exp streams over an ingested data set: the mixed tolerance of
    # the FP64 oracle fit of the original problem
    name, p, t = [c for c in round_trip_problems() if c[0] == "lr_synth"][0]
    d = api.Data.from_long(**to_long(p, np.random.default_rng(34), time=t), precision=api.MIXED)
    sp = default_spec(p, prior=PRIOR_LAPLACE, variance=1.0)
    bg, ig = api.fit_plain(d, sp)
    d.close()
    bc, ic = orc.fit(p, sp)
    check_fit(bg, ig, bc, ic, rtol=2e-4, atol=2e-5, ll_rtol=1e-6, same_path=False)
