"""The oracle's failure path (CPU): the ILLCONDITIONED retry of
CyclicCoordinateDescent::update (CyclicCoordinateDescent.cpp:586-615: a NaN
objective ends findMode with ILLCONDITIONED, :849-858, and the fit restarts from
zero coefficients with the bounding box divided by 10, at most maxBoundCount
times) and MAX_ITERATIONS, on designs with a separable covariate shaped like the
reference's separability tests (tests/testthat/test-largeBernoulli.R:25-101).

Pins that do not restate the retry loop: the attempt that ends a retried fit is
an ordinary fit started from zero with the bounding box 2 / 10^(count - 1), so it
must equal a fresh one-attempt fit with that initial bound bit for bit; a fit
that ends ILLCONDITIONED has a NaN coefficient exactly where the separable
covariate is and a NaN log likelihood."""
import numpy as np

from ccd_amd import synth
from ccd_amd.data import COX, ILLCONDITIONED, LOGISTIC, MAX_ITERATIONS, SUCCESS, default_spec
from oracle import oracle as orc

SEPARABLE = 6  # intercept, five covariates, then the separable copy of covariate 3


def failure_cases():
    """(name, problem, spec) of the failure-path designs, each checked below to
    take the path it is named for."""
    lr666 = synth.separable(LOGISTIC, seed=666)
    lr3 = synth.separable(LOGISTIC, seed=3)
    cox = synth.separable(COX, seed=666)
    return [
        ("lr_retry_then_success", lr666, default_spec(lr666)),
        ("lr_retries_exhausted", lr3, default_spec(lr3)),
        ("lr_one_attempt", lr666, default_spec(lr666, max_bound_count=1)),
        ("cox_max_iterations", cox, default_spec(cox, max_iterations=100)),
    ]


def _fresh(p, spec, count):
    s = default_spec(p, max_bound_count=1, initial_bound=spec.initial_bound / 10.0 ** (count - 1),
                     max_iterations=spec.max_iterations)
    return orc.fit(p, s)


def test_separable_designs_take_the_failure_paths():
    cases = {n: orc.fit(p, s) for n, p, s in failure_cases()}
    b, i = cases["lr_retry_then_success"]
    assert i["return_flag"] == SUCCESS and i["bound_count"] == 3 and np.all(np.isfinite(b))
    assert b[SEPARABLE] > 20  # the separable coefficient ran away before the box shrank
    b, i = cases["lr_retries_exhausted"]
    assert i["return_flag"] == ILLCONDITIONED and i["bound_count"] == 5
    b, i = cases["lr_one_attempt"]
    assert i["return_flag"] == ILLCONDITIONED and i["bound_count"] == 1
    b, i = cases["cox_max_iterations"]
    assert i["return_flag"] == MAX_ITERATIONS and i["iterations"] == 100


def test_retried_fit_equals_a_fresh_fit_with_the_shrunken_box():
    for name, p, spec in failure_cases():
        b, i = orc.fit(p, spec)
        bf, fi = _fresh(p, spec, i["bound_count"])
        assert fi["bound_count"] == 1
        assert np.array_equal(b, bf, equal_nan=True), name
        assert i["iterations"] == fi["iterations"] and i["return_flag"] == fi["return_flag"], name
        if i["bound_count"] > 1:  # and every earlier box alone ends ill-conditioned
            for c in range(1, i["bound_count"]):
                _, ei = _fresh(p, spec, c)
                assert ei["return_flag"] == ILLCONDITIONED, (name, c)


def test_ill_conditioned_fits_are_nan_exactly_at_the_separable_covariate():
    for name, p, spec in failure_cases():
        b, i = orc.fit(p, spec)
        if i["return_flag"] != ILLCONDITIONED:
            continue
        assert np.flatnonzero(np.isnan(b)).tolist() == [SEPARABLE], name
        assert np.isnan(i["log_likelihood"]), name


def test_separable_covariate_only_on_events():
    for model in (LOGISTIC, COX):
        p = synth.separable(model, seed=666)
        r, _ = p.column(SEPARABLE if model == LOGISTIC else SEPARABLE - 1)
        assert len(r) > 10 and np.all(p.y[r] == 1)
