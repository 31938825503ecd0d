"""The oracle's failure path (CPU): the ILLCONDITIONED retry of
CyclicCoordinateDescent::update (CyclicCoordinateDescent.cpp:586-615: a NaN
objective ends findMode with ILLCONDITIONED, :849-858, and the fit restarts from
zero coefficients with the bounding box divided by 10, at most maxBoundCount
attempts) and MAX_ITERATIONS, on designs shaped like the reference's separability
tests (tests/testthat/test-largeBernoulli.R:25-101) and on an exp overflow.

Pins that do not restate the retry loop:
* each case takes the path its construction forces (a separated logistic
  coefficient marches until its step is 0 / 0; a clipped first Poisson step of 2
  on values of 500..1000 overflows exp, one of 0.2 does not);
* the attempt that ends a retried fit is an ordinary fit from zero with the
  bounding box 2 / 10^(count - 1): it equals a fresh one-attempt fit with that
  initial bound bit for bit, and every earlier box alone ends ILLCONDITIONED;
* the outcome does not depend on summation order: the same design with its rows
  permuted takes the same path (flag, bound count, iterations, NaN positions)
  with coefficients within 1e-10, so these cases can be compared across engines
  that sum in other orders.  (The reference's own separability design, where two
  covariates are collinear on the events, is not such a case: its late sweeps
  drift along a flat direction at rounding level, and whether a NaN appears
  depends on the order of the sums.  The separated design's sweep count does
  depend on the last bits of exp() near 1e16, which is why the GPU tests compare
  its flag, bound count and NaN positions but not its sweeps.)"""
import numpy as np

from ccd_amd import synth
from ccd_amd.data import (COX, ILLCONDITIONED, INTERCEPT, MAX_ITERATIONS, SPARSE, SUCCESS,
                          build_problem, default_spec)
from oracle import oracle as orc


def failure_cases():
    """(name, problem, spec, expected) of the failure-path designs; expected =
    (return flag, bound count, NaN coefficient indices or None to skip)."""
    lr = synth.lr_separated()
    pois = synth.poisson_overflow()
    cox = synth.separable(COX, seed=666)
    sep = lr.J - 1
    return [
        ("lr_separated", lr, default_spec(lr), (ILLCONDITIONED, 5, [sep])),
        ("lr_separated_two_attempts", lr, default_spec(lr, max_bound_count=2),
         (ILLCONDITIONED, 2, [sep])),
        ("lr_separated_max_iterations", lr, default_spec(lr, max_iterations=10),
         (MAX_ITERATIONS, 1, [])),
        ("poisson_overflow_retry", pois, default_spec(pois), (SUCCESS, 2, [])),
        ("poisson_overflow_one_attempt", pois, default_spec(pois, max_bound_count=1),
         (ILLCONDITIONED, 1, [1, 2, 3])),
        ("cox_max_iterations", cox, default_spec(cox, max_iterations=100),
         (MAX_ITERATIONS, 1, [])),
    ]


def _fresh(p, spec, count):
    s = default_spec(p, max_bound_count=1, initial_bound=spec.initial_bound / 10.0 ** (count - 1),
                     max_iterations=spec.max_iterations)
    return orc.fit(p, s)


def _permuted(p, seed=1):
    perm = np.random.default_rng(seed).permutation(p.K)
    inv = np.argsort(perm)
    cols = [(p.fmt[j], inv[p.column(j)[0]], p.column(j)[1] if p.fmt[j] == SPARSE else None)
            for j in range(p.J) if p.fmt[j] != INTERCEPT]
    return build_problem(p.model, cols, p.y[perm], intercept=bool(np.any(p.fmt == INTERCEPT)))


def test_each_case_takes_its_path():
    for name, p, spec, (flag, count, nan_cols) in failure_cases():
        b, i = orc.fit(p, spec)
        assert (i["return_flag"], i["bound_count"]) == (flag, count), name
        assert np.flatnonzero(np.isnan(b)).tolist() == nan_cols, name
        assert np.isnan(i["log_likelihood"]) == (flag == ILLCONDITIONED), name
    # the separated coefficient ran up to where exp(x beta) / (1 + exp(x beta)) is 1
    lr = synth.lr_separated()
    b, i = orc.fit(lr, default_spec(lr, max_iterations=10))
    assert b[-1] > 5 and i["iterations"] == 10
    # the overflowing Poisson attempt stopped after one sweep at the clipped step
    pois = synth.poisson_overflow()
    b, i = orc.fit(pois, default_spec(pois, max_bound_count=1))
    assert i["iterations"] == 1 and b[0] == 2.0


def test_retried_fit_equals_a_fresh_fit_with_the_shrunken_box():
    for name, p, spec, _ in failure_cases():
        b, i = orc.fit(p, spec)
        bf, fi = _fresh(p, spec, i["bound_count"])
        assert np.array_equal(b, bf, equal_nan=True), name
        assert (i["iterations"], i["return_flag"]) == (fi["iterations"], fi["return_flag"]), name
        for c in range(1, i["bound_count"]):
            assert _fresh(p, spec, c)[1]["return_flag"] == ILLCONDITIONED, (name, c)


def test_outcome_does_not_depend_on_summation_order():
    for name, p, spec, _ in failure_cases():
        if p.model == COX:  # rows are in risk-set order
            continue
        b, i = orc.fit(p, spec)
        q = _permuted(p)
        bq, iq = orc.fit(q, default_spec(q, max_iterations=spec.max_iterations,
                                         max_bound_count=spec.max_bound_count))
        for k in ("return_flag", "bound_count", "iterations"):
            assert i[k] == iq[k], (name, k)
        assert np.array_equal(np.isnan(b), np.isnan(bq)), name
        ok = ~np.isnan(b)
        np.testing.assert_allclose(bq[ok], b[ok], rtol=1e-10, atol=1e-12, err_msg=name)


def test_separated_covariate_only_on_events():
    lr = synth.lr_separated()
    r, _ = lr.column(lr.J - 1)
    assert len(r) == 60 and np.all(lr.y[r] == 1)
    for j in range(1, lr.J - 1):
        assert not np.intersect1d(lr.column(j)[0], r).size
