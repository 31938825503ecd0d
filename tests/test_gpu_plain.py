"""GPU parity of the plain (per-coordinate) engine against the CPU oracle.

Tolerances: the FP64 path matches the oracle's coefficients and log likelihood
to 1e-8 relative (atol 1e-10 for coefficients at zero), with the same iteration
count and return flag.  The mixed path (FP32 xBeta/exp streams, FP64
reductions) is held to 2e-4 relative on coefficients (atol 2e-5) and 1e-6
relative on the log likelihood.
"""
import numpy as np
import pytest

from ccd_amd import api, synth
from ccd_amd.data import (CONDITIONAL_LOGISTIC, COX, PRIOR_LAPLACE, PRIOR_NORMAL, default_spec)
from oracle import oracle as orc
from tests import fixtures
from tests.test_oracle import _small_problems

pytestmark = pytest.mark.gpu

RTOL64, ATOL64 = 1e-8, 1e-10


def check_fit(bg, ig, bc, ic, rtol=RTOL64, atol=ATOL64, ll_rtol=RTOL64, same_path=True):
    np.testing.assert_allclose(bg, bc, rtol=rtol, atol=atol)
    np.testing.assert_allclose(ig["log_likelihood"], ic["log_likelihood"], rtol=ll_rtol)
    if same_path:
        assert ig["iterations"] == ic["iterations"]
        assert ig["return_flag"] == ic["return_flag"]


@pytest.mark.parametrize("name", fixtures.names())
@pytest.mark.parametrize("prior", ["none", "laplace"])
def test_plain_fixture_parity(name, prior):
    p, extra = fixtures.load(name)
    spec = default_spec(p, prior=PRIOR_LAPLACE if prior == "laplace" else 0, variance=1.0)
    if "weights" in extra:
        spec.weights = extra["weights"]
    bc, ic = orc.fit(p, spec)
    d = api.Data(p)
    bg, ig = api.fit_plain(d, spec)
    check_fit(bg, ig, bc, ic)
    if prior == "none":  # and the golden MLE within the fit's own tolerance
        np.testing.assert_allclose(bg, extra["beta"], rtol=2e-3, atol=2e-3)


@pytest.mark.parametrize("name,p", _small_problems(), ids=lambda v: v if isinstance(v, str) else "")
@pytest.mark.parametrize("weighted", [False, True])
def test_per_coordinate_ops_match_oracle(name, p, weighted):
    rng = np.random.default_rng(5)
    beta = rng.normal(0, 0.4, p.J)
    if p.has_offset:
        beta[0] = 1.0
    w = None
    if weighted:
        if p.model == CONDITIONAL_LOGISTIC:
            w = np.repeat((rng.random(p.strata.max() + 1) < 0.7).astype(float), 5)
        else:
            w = rng.choice([0.0, 0.5, 1.0, 2.0], p.K)
    d = api.Data(p)
    fh = api.FitHandle(d)
    fh.set_weights(w, use_cv=w is not None)
    fh.fixed_terms(use_cv=w is not None)
    fh.compute_xbeta(beta)
    fh.remaining_stats(use_w=w is not None)
    for j in range(p.J):
        ll_c, g_c, h_c = orc.evaluate(p, beta, j, weights=w)
        g, h = fh.grad_hess(j, use_w=w is not None)
        np.testing.assert_allclose([g, h], [g_c, h_c], rtol=1e-11, atol=1e-12)
    ll = fh.loglik(use_cv=w is not None)
    np.testing.assert_allclose(ll, ll_c, rtol=1e-12)
    # one update then re-evaluate every coordinate against fresh oracle statistics
    j = p.J - 1
    fh.update(0.37, j, use_w=w is not None)
    beta2 = beta.copy()
    beta2[j] += 0.37
    for jj in range(p.J):
        _, g_c, h_c = orc.evaluate(p, beta2, jj, weights=w)
        g, h = fh.grad_hess(jj, use_w=w is not None)
        np.testing.assert_allclose([g, h], [g_c, h_c], rtol=1e-10, atol=1e-11)
    np.testing.assert_allclose(fh.xbeta(), beta2 @ p.to_dense().T, rtol=1e-12, atol=1e-12)


# (problem, prior, variance, max_iterations); the Poisson design has heavy Zipf
# columns collinear with the intercept, so CCD needs hundreds of sweeps there and
# the test compares the first 40 (same trajectory, MAX_ITERATIONS flag on both).
SYNTH = {
    "lr": lambda: (synth.lr_indicator(K=3000, P=300, p=0.02, n_true=20, seed=1)[0],
                   PRIOR_LAPLACE, 1.0, 1000),
    "poisson": lambda: (synth.poisson_zipf(K=20000, P=2000, per_row=10, n_true=50, seed=2)[0],
                        PRIOR_NORMAL, 0.1, 40),
    "cox": lambda: (synth.cox_zipf(K=6000, P=400, per_row=8, n_true=40, seed=3)[0],
                    PRIOR_LAPLACE, 0.05, 1000),
    "clr": lambda: (synth.clr_strata(n_strata=1500, P=400, per_row=8, n_true=40, seed=4)[0],
                    PRIOR_LAPLACE, 0.1, 1000),
}


@pytest.mark.parametrize("name", sorted(SYNTH))
def test_plain_synthetic_parity(name):
    p, prior, var, maxit = SYNTH[name]()
    spec = default_spec(p, prior=prior, variance=var, max_iterations=maxit)
    bc, ic = orc.fit(p, spec)
    bg, ig = api.fit_plain(api.Data(p), spec)
    check_fit(bg, ig, bc, ic)


@pytest.mark.parametrize("name", sorted(SYNTH))
def test_plain_cv_weights_and_predictive_ll(name):
    p, prior, var, maxit = SYNTH[name]()
    by_pid = p.model in (CONDITIONAL_LOGISTIC,)
    ids = p.strata if by_pid else np.zeros(p.K, np.int32)
    w = orc.cv_weights(ids, 10, 123, 3, by_pid=by_pid)
    spec = default_spec(p, prior=prior, variance=var, weights=w, max_iterations=maxit)
    heldout = 1.0 - w
    bc, ic = orc.fit(p, spec, heldout=heldout)
    bg, ig = api.fit_plain(api.Data(p), spec, heldout=heldout)
    check_fit(bg, ig, bc, ic)
    np.testing.assert_allclose(ig["pred_log_likelihood"], ic["pred_log_likelihood"], rtol=1e-8)


@pytest.mark.parametrize("name", sorted(SYNTH))
def test_plain_mixed_precision_tolerance(name):
    p, prior, var, maxit = SYNTH[name]()
    spec = default_spec(p, prior=prior, variance=var, max_iterations=maxit)
    bc, ic = orc.fit(p, spec)
    bg, ig = api.fit_plain(api.Data(p, precision=api.MIXED), spec)
    check_fit(bg, ig, bc, ic, rtol=2e-4, atol=2e-5, ll_rtol=1e-6, same_path=False)
