"""Warm-started cross-validation chains (CPU): the TaskScheduler's task split
(Thread.h:64-80) and the oracle's warm-started grid search
(AbstractCrossValidationDriver.cpp:131-248 with resetCoefficients = FALSE).

Pins: the chains are the closed form of TaskScheduler's chunking; every warm
task reaches the optimum of its own fold and variance -- the cold-started fit's,
within the convergence tolerance -- in no more sweeps in total; and a chain that
meets a failed fit restarts from zero."""
import numpy as np
import pytest

from ccd_amd import api, synth
from ccd_amd.data import CONDITIONAL_LOGISTIC, PRIOR_LAPLACE, default_spec
from oracle import oracle as orc


def test_chains_follow_the_task_scheduler_split():
    assert api.cv_chains(10, 1) == [list(range(10))]
    assert api.cv_chains(10, 4) == [[0, 1, 2], [3, 4, 5], [6, 7, 8], [9]]
    assert api.cv_chains(10, 3) == [[0, 1, 2, 3], [4, 5, 6, 7], [8, 9]]
    assert api.cv_chains(10, 10) == [[i] for i in range(10)]
    assert api.cv_chains(4, 16) == [[0], [1], [2], [3]]
    # chunk = ceil(10 / 6) = 2: five threads get tasks, as TaskScheduler::execute stops
    assert api.cv_chains(10, 6) == [[0, 1], [2, 3], [4, 5], [6, 7], [8, 9]]


def _problem():
    p, _ = synth.clr_strata(n_strata=600, P=120, per_row=6, n_true=20, seed=8)
    return p


@pytest.mark.parametrize("threads", [1, 4])
def test_warm_chain_reaches_each_folds_optimum(threads):
    p = _problem()
    spec = default_spec(p, prior=PRIOR_LAPLACE, variance=1.0)
    variances = [0.01, 0.1, 1.0]
    scores, res, _ = orc.grid_search_cv_chain(p, spec, p.strata, variances, n_fold=4, seed=5,
                                              threads=threads)
    warm_iters = cold_iters = 0
    for g, var in enumerate(variances):
        for f in range(4):
            w = orc.cv_weights(p.strata, 4, 5, f)
            cold = default_spec(p, prior=PRIOR_LAPLACE, variance=var, weights=w)
            bc, ic = orc.fit(p, cold, heldout=1.0 - w)
            assert res[g][f]["return_flag"] == 0
            np.testing.assert_allclose(scores[g, f], ic["pred_log_likelihood"], rtol=1e-5)
            warm_iters += res[g][f]["iterations"]
            cold_iters += ic["iterations"]
    assert warm_iters <= cold_iters


def test_failed_task_restarts_its_chain_from_zero():
    """A fold whose fit stops at MAX_ITERATIONS scores NaN and the next task of its
    chain starts from zero: with max_iterations 1 every task fails, so each is a
    one-sweep fit from zero."""
    p = _problem()
    spec = default_spec(p, prior=PRIOR_LAPLACE, variance=1.0, max_iterations=1)
    scores, res, beta = orc.grid_search_cv_chain(p, spec, p.strata, [0.1, 1.0], n_fold=4, seed=5)
    assert np.all(np.isnan(scores))
    w = orc.cv_weights(p.strata, 4, 5, 2)
    b1, i1 = orc.fit(p, default_spec(p, prior=PRIOR_LAPLACE, variance=1.0, weights=w,
                                     max_iterations=1))
    assert res[1][2]["log_likelihood"] == i1["log_likelihood"]
    assert np.all(beta[0] == 0.0)
