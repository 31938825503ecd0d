"""GPU parity of warm-started cross-validation (resetCoefficients = FALSE, the
reference default): api.grid_search_cv runs the reference's thread chains
(every task starts from its chain's previous estimate, across grid points) with
the tasks at one chain position batched, and matches the oracle's task-by-task
chain: per task the same iterations and return flag, held-out scores within
1e-8, and the chains' final coefficients within 1e-8 (FP64)."""
import numpy as np
import pytest

from ccd_amd import api, synth
from ccd_amd.data import LOGISTIC, PRIOR_LAPLACE, default_spec
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _problems():
    clr, _ = synth.clr_strata(n_strata=600, P=120, per_row=6, n_true=20, seed=8)
    lr, _ = synth.lr_indicator(K=3000, P=200, p=0.03, n_true=15, seed=9)
    return {"clr": (clr, clr.strata, True), "lr": (lr, np.zeros(lr.K, np.int32), False)}


@pytest.mark.parametrize("name", ["clr", "lr"])
@pytest.mark.parametrize("threads", [1, 4])
def test_warm_grid_search_matches_the_oracle_chain(name, threads):
    p, ids, by_pid = _problems()[name]
    spec = default_spec(p, prior=PRIOR_LAPLACE, variance=1.0)
    variances = [0.01, 0.1, 1.0]
    print("warm cv", name, "threads", threads, flush=True)
    sc, rc, bc = orc.grid_search_cv_chain(p, spec, ids, variances, n_fold=4, seed=5,
                                          by_pid=by_pid, threads=threads)
    d = api.Data(p)
    point, sg, rg, bg = api.grid_search_cv(d, spec, ids, variances, n_fold=4, seed=5,
                                           by_pid=by_pid, threads=threads)
    d.close()
    for g in range(len(variances)):
        for f in range(4):
            assert rg[g][f]["iterations"] == rc[g][f]["iterations"], (g, f)
            assert rg[g][f]["return_flag"] == rc[g][f]["return_flag"], (g, f)
    np.testing.assert_allclose(sg, sc, rtol=1e-8)
    for a, b in zip(bg, bc):
        np.testing.assert_allclose(a, b, rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(point, np.nanmean(sc, axis=1), rtol=1e-8)


def test_cold_grid_search_matches_cold_oracle_fits():
    p, ids, by_pid = _problems()["clr"]
    spec = default_spec(p, prior=PRIOR_LAPLACE, variance=1.0)
    variances = [0.05, 0.5]
    _, sg, rg, _ = api.grid_search_cv(api.Data(p), spec, ids, variances, n_fold=4, seed=5,
                                      reset_coefficients=True)
    for g, var in enumerate(variances):
        for f in range(4):
            w = orc.cv_weights(ids, 4, 5, f)
            _, ic = orc.fit(p, default_spec(p, prior=PRIOR_LAPLACE, variance=var, weights=w),
                            heldout=1.0 - w)
            assert rg[g][f]["iterations"] == ic["iterations"]
            np.testing.assert_allclose(sg[g, f], ic["pred_log_likelihood"], rtol=1e-8)
