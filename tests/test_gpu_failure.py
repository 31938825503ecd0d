"""GPU parity on the failure path: the ILLCONDITIONED retry (bound / 10, cold
restart; CyclicCoordinateDescent.cpp:586-615, 849-858) and MAX_ITERATIONS, on
designs with a separable covariate (tests/testthat/test-largeBernoulli.R:25-101).

Every engine — plain, batched (one workgroup, a gang, two fits bundled in one
slot, many fits in one launch) and served — ends with the oracle's return flag,
bound count and iteration count, NaN exactly where the oracle has NaN, and the
other coefficients and the log likelihood within 1e-8 relative (FP64)."""
import numpy as np
import pytest

from ccd_amd import api
from ccd_amd.data import PRIOR_LAPLACE, default_spec
from oracle import oracle as orc
from tests.test_oracle_failure import failure_cases

pytestmark = pytest.mark.gpu

CASES = failure_cases()
IDS = [c[0] for c in CASES]


def check_flagged(bg, ig, bc, ic, name=""):
    assert ig["return_flag"] == ic["return_flag"], name
    assert ig["bound_count"] == ic["bound_count"], name
    assert ig["iterations"] == ic["iterations"], name
    assert np.array_equal(np.isnan(bg), np.isnan(bc)), name
    ok = ~np.isnan(bc)
    np.testing.assert_allclose(bg[ok], bc[ok], rtol=1e-8, atol=1e-10, err_msg=name)
    if np.isnan(ic["log_likelihood"]):
        assert np.isnan(ig["log_likelihood"]), name
    else:
        np.testing.assert_allclose(ig["log_likelihood"], ic["log_likelihood"], rtol=1e-8)


@pytest.mark.parametrize("case", range(len(CASES)), ids=IDS)
def test_plain_failure_path(case):
    name, p, spec = CASES[case]
    print("plain", name, flush=True)
    bc, ic = orc.fit(p, spec)
    d = api.Data(p)
    bg, ig = api.fit_plain(d, spec)
    d.close()
    check_flagged(bg, ig, bc, ic, name)


@pytest.mark.parametrize("gang", [1, 4])
@pytest.mark.parametrize("case", range(len(CASES)), ids=IDS)
def test_batch_failure_path(case, gang):
    name, p, spec = CASES[case]
    print("batch", name, "gang", gang, flush=True)
    bc, ic = orc.fit(p, spec)
    d = api.Data(p)
    bg, ig = api.fit_batch(d, [spec], max_gang=gang, gang_min_nnz=1)
    d.close()
    check_flagged(bg[0], ig[0], bc, ic, name)


def test_batch_failure_path_bundled_and_in_the_mix():
    """Two fits share a slot (lock step over interleaved state): the one that
    needs a restart waits for its partner and restarts alone; other fits of the
    launch are unaffected."""
    name, p, spec = CASES[0]
    lap = default_spec(p, prior=PRIOR_LAPLACE, variance=1.0)
    specs = [spec, lap, default_spec(p, max_bound_count=1), lap]
    refs = [orc.fit(p, s) for s in specs]
    d = api.Data(p)
    for bundle in (2, 1):
        print("bundle", bundle, flush=True)
        bg, ig = api.fit_batch(d, specs, bundle=bundle, max_gang=1)
        for i, (bc, ic) in enumerate(refs):
            check_flagged(bg[i], ig[i], bc, ic, f"{name} fit {i} bundle {bundle}")
    d.close()


def test_served_failure_path():
    refs = [orc.fit(p, s) for _, p, s in CASES]
    with api.Server(quantum_ms=1.0) as srv:
        ids = [srv.submit(p, s) for _, p, s in CASES]
        got = [srv.wait(i) for i in ids]
    for (name, _, _), (bc, ic), (bg, ig, _) in zip(CASES, refs, got):
        check_flagged(bg, ig, bc, ic, name)
