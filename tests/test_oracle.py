"""Pins of the CPU oracle (oracle/ccd_oracle.cpp) -- CPU only.

Each check is chosen so that a plausible mistake in the restatement fails it:
  * golden fixtures: maximum-likelihood estimates from an independent full-vector
    Newton solver on the reference's own test datasets, plus the SAS -2 logL
    printed in tests/testthat/test-largeBernoulli.R:20;
  * per-coordinate gradient/Hessian against the closed-form score and
    information matrix of the independent likelihoods, at random coefficients,
    for every model, format and weighting (Cox by explicit O(K^2) risk sets);
  * prior steps against brute-force one-dimensional minimisation;
  * penalised fits against their KKT conditions;
  * Cox tie grouping, CV fold weights and the variance grid against properties.
"""
import importlib.util
import os

import numpy as np
import pytest

from ccd_amd import synth
from ccd_amd.data import (COX, CONDITIONAL_LOGISTIC, DENSE, INDICATOR, LOGISTIC, POISSON,
                          PRIOR_LAPLACE, PRIOR_NONE, PRIOR_NORMAL, SPARSE, build_problem,
                          default_spec, sort_cox)
from oracle import oracle as orc
from tests import fixtures

_spec = importlib.util.spec_from_file_location(
    "make_golden", os.path.join(fixtures.GOLDEN, "make_golden.py"))
mg = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(mg)


def independent_terms(p, beta, weights=None):
    """(ll, score, information) from the dense, independent likelihoods."""
    X = p.to_dense()
    w = np.ones(p.K) if weights is None else weights
    if p.model == LOGISTIC:
        ll, g, H = mg.lr_terms(X, p.y, w, beta)
    elif p.model == POISSON:
        off = X[:, 0] if p.has_offset else np.zeros(p.K)
        Xf = X[:, 1:] if p.has_offset else X
        ll, g, H = mg.poisson_terms(Xf, p.y, w, beta[1:] if p.has_offset else beta, off)
        if p.has_offset:
            g = np.concatenate([[np.nan], g])
            H = np.pad(H, ((1, 0), (1, 0)), constant_values=np.nan)
    elif p.model == CONDITIONAL_LOGISTIC:
        assert weights is None or np.all((weights == 0) | (weights == 1))
        keep = w > 0
        strata = p.strata
        ll, g, H = mg.clr_terms(X[keep], p.y[keep], strata[keep], beta)
    else:
        ll, g, H = mg.cox_breslow_terms(X, p.offs, p.y, p.strata, w, beta)
    return ll, g, H


@pytest.mark.parametrize("name", fixtures.names())
def test_oracle_matches_newton_mle(name):
    p, extra = fixtures.load(name)
    spec = default_spec(p, tolerance=1e-13, max_iterations=200000)
    if "weights" in extra:
        spec.weights = extra["weights"]
    beta, info = orc.fit(p, spec)
    gold = extra["beta"]
    assert info["return_flag"] == 0
    np.testing.assert_allclose(beta, gold, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(info["log_likelihood"], float(extra["loglik"]), rtol=1e-9)


def test_oracle_sas_minus2ll():
    p, extra = fixtures.load("lr_sas17")
    beta, info = orc.fit(p, default_spec(p, tolerance=1e-12, max_iterations=100000))
    assert abs(2 * info["log_likelihood"] - float(extra["sas_minus2ll"])) < 5e-4


def _small_problems():
    rng = np.random.default_rng(7)
    out = []
    K = 60
    x_ind = np.nonzero(rng.random(K) < 0.3)[0]
    x_sp_rows = np.nonzero(rng.random(K) < 0.4)[0]
    x_sp_vals = rng.normal(0, 1, len(x_sp_rows))
    x_dense = rng.normal(0, 0.7, K)
    cols = [(INDICATOR, x_ind, None), (SPARSE, x_sp_rows, x_sp_vals), (DENSE, None, x_dense)]
    y_lr = (rng.random(K) < 0.4).astype(float)
    out.append(("lr", build_problem(LOGISTIC, cols, y_lr, intercept=True)))
    y_po = rng.poisson(2.0, K).astype(float)
    out.append(("poisson", build_problem(POISSON, cols, y_po, intercept=True)))
    out.append(("poisson_off", build_problem(POISSON, cols, y_po, intercept=True,
                                             offset=rng.normal(0, 0.3, K))))
    strata = np.repeat(np.arange(12), 5).astype(np.int32)
    y_clr = np.zeros(K)
    y_clr[np.arange(12) * 5 + rng.integers(0, 5, 12)] = 1
    out.append(("clr", build_problem(CONDITIONAL_LOGISTIC, cols, y_clr, strata=strata)))
    time = rng.integers(1, 15, K).astype(float)
    ev = (rng.random(K) < 0.6).astype(float)
    cstrata = (rng.random(K) < 0.5).astype(np.int32)
    order = sort_cox(time, ev, cstrata)
    inv = np.empty(K, int)
    inv[order] = np.arange(K)
    ccols = [(INDICATOR, inv[x_ind], None), (SPARSE, inv[x_sp_rows], x_sp_vals),
             (DENSE, None, x_dense[order])]
    out.append(("cox", build_problem(COX, ccols, ev[order], offs=time[order],
                                     strata=cstrata[order])))
    return out


@pytest.mark.parametrize("name,p", _small_problems(), ids=lambda v: v if isinstance(v, str) else "")
@pytest.mark.parametrize("weighted", [False, True])
def test_oracle_gradient_hessian_closed_form(name, p, weighted):
    rng = np.random.default_rng(11)
    beta = rng.normal(0, 0.4, p.J)
    if p.has_offset:
        beta[0] = 1.0
    w = None
    if weighted:
        if p.model == CONDITIONAL_LOGISTIC:
            w = np.repeat((rng.random(p.strata.max() + 1) < 0.7).astype(float), 5)
        else:
            w = rng.choice([0.0, 0.5, 1.0, 2.0], p.K)
    ll_i, score, H = independent_terms(p, beta, w)
    for j in range(p.J):
        if p.has_offset and j == 0:
            continue
        ll, g, h = orc.evaluate(p, beta, j, weights=w)
        np.testing.assert_allclose(ll, ll_i, rtol=1e-11, atol=1e-11)
        np.testing.assert_allclose(g, -score[j], rtol=1e-10, atol=1e-10)
        np.testing.assert_allclose(h, -H[j, j], rtol=1e-10, atol=1e-10)
    # every column from one set of statistics (the full-size checks use this form)
    cols = np.arange(1 if p.has_offset else 0, p.J)[::-1]
    ll, gs, hs = orc.evaluate_cols(p, beta, cols, weights=w)
    np.testing.assert_allclose(ll, ll_i, rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(gs, -score[cols], rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(hs, -np.diag(H)[cols], rtol=1e-10, atol=1e-10)


def _brute_min(f, lo=-6, hi=6):
    xs = np.linspace(lo, hi, 240001)
    v = f(xs)
    i = int(np.argmin(v))
    a, b = xs[max(i - 1, 0)], xs[min(i + 1, len(xs) - 1)]
    for _ in range(100):  # golden-section refine
        m1, m2 = a + (b - a) * 0.382, a + (b - a) * 0.618
        if f(np.array([m1]))[0] < f(np.array([m2]))[0]:
            b = m2
        else:
            a = m1
    return 0.5 * (a + b)


@pytest.mark.parametrize("g,h,beta,var", [
    (0.5, 2.0, 0.0, 1.0), (3.0, 2.0, 0.0, 1.0), (-3.0, 2.0, 0.0, 1.0), (1.2, 0.7, 0.4, 0.5),
    (-0.3, 1.5, -0.8, 2.0), (0.2, 3.0, 0.9, 0.1)])
def test_prior_delta_minimises_penalised_quadratic(g, h, beta, var):
    lam = np.sqrt(2.0 / var)
    # Laplace: argmin g d + h d^2 / 2 + lam |beta + d| unless the step flips the sign
    d = orc.prior_delta(PRIOR_LAPLACE, var, g, h, beta)
    f = lambda x: g * x + 0.5 * h * x * x + lam * np.abs(beta + x)  # noqa: E731
    dstar = _brute_min(f)
    if beta != 0 and np.sign(beta + dstar) != np.sign(beta) and np.sign(beta + dstar) != 0:
        assert d == -beta
    else:
        assert abs(d - dstar) < 1e-6
    # Normal: argmin g d + h d^2 / 2 + (beta + d)^2 / (2 var)
    dn = orc.prior_delta(PRIOR_NORMAL, var, g, h, beta)
    fn = lambda x: g * x + 0.5 * h * x * x + (beta + x) ** 2 / (2 * var)  # noqa: E731
    assert abs(dn - _brute_min(fn)) < 1e-6
    # no prior: Newton step
    assert abs(orc.prior_delta(PRIOR_NONE, var, g, h, beta) + g / h) < 1e-15


def test_prior_delta_sign_flip_clamps_to_zero():
    # beta=0.1, strong push across zero: the reference stops at zero
    assert orc.prior_delta(PRIOR_LAPLACE, 2.0, 5.0, 1.0, 0.1) == -0.1
    assert orc.prior_delta(PRIOR_LAPLACE, 2.0, 0.5, 1.0, 0.0) == 0.0


def test_prior_log_density():
    assert abs(orc.prior_log_density(PRIOR_LAPLACE, 2.0, 0.3) - (np.log(0.5) - 0.3)) < 1e-15
    v = 0.1
    expect = -0.5 * np.log(2 * np.pi * v) - 0.5 * 0.09 / v
    assert abs(orc.prior_log_density(PRIOR_NORMAL, v, 0.3) - expect) < 1e-14


@pytest.mark.parametrize("prior,var", [(PRIOR_LAPLACE, 0.05), (PRIOR_LAPLACE, 1.0),
                                       (PRIOR_NORMAL, 0.1)])
@pytest.mark.parametrize("name,p", _small_problems(), ids=lambda v: v if isinstance(v, str) else "")
def test_oracle_penalised_fit_kkt(name, p, prior, var):
    spec = default_spec(p, prior=prior, variance=var, tolerance=1e-13, max_iterations=100000)
    beta, info = orc.fit(p, spec)
    assert info["return_flag"] == 0
    _, score, _ = independent_terms(p, beta)
    lam = np.sqrt(2.0 / var)
    for j in range(p.J):
        if spec.fixed[j]:
            continue
        s = score[j]
        if spec.prior_type[j] == PRIOR_NONE:
            assert abs(s) < 1e-5
        elif prior == PRIOR_NORMAL:
            assert abs(s - beta[j] / var) < 1e-5
        elif beta[j] == 0.0:
            assert abs(s) <= lam + 1e-6
        else:
            assert abs(s - lam * np.sign(beta[j])) < 1e-5


def test_cox_pids_tie_grouping():
    # rows sorted by (-time, y): censored rows precede events at equal times
    time = np.array([5, 4, 4, 4, 3, 3, 2, 1.0])
    ev = np.array([1, 0, 1, 1, 1, 1, 0, 1.0])
    p = build_problem(COX, [(INDICATOR, [0, 2, 5], None)], ev, offs=time)
    N, pid, reset = orc.cox_pids(p)
    # tied events at t=4 share a pid; at t=3 too; censored rows are their own pid
    assert list(pid) == [0, 1, 2, 2, 3, 3, 4, 5]
    assert N == 6 and list(reset) == [6]
    w = np.array([1, 1, 0, 1, 1, 1, 1, 1.0])
    N, pid, reset = orc.cox_pids(p, w)
    assert list(pid) == [0, 1, 6, 2, 3, 3, 4, 5] and N == 6  # ignored row -> pid N


def test_cox_pids_strata_resets():
    time = np.array([5, 3, 1, 6, 2.0])
    ev = np.array([1, 1, 1, 1, 1.0])
    strata = np.array([0, 0, 0, 1, 1], np.int32)
    p = build_problem(COX, [(INDICATOR, [0, 3], None)], ev, offs=time, strata=strata)
    N, pid, reset = orc.cox_pids(p)
    assert list(pid) == [0, 1, 2, 3, 4] and list(reset) == [3, 5]


@pytest.mark.parametrize("by_pid", [True, False])
def test_cv_fold_weights_partition(by_pid):
    ids = np.repeat(np.arange(23), 3).astype(np.int32) if by_pid else np.zeros(50, np.int32)
    n_units = 23 if by_pid else 50
    fold = 10
    held = np.zeros(len(ids))
    sizes = []
    for task in range(fold):
        w = orc.cv_weights(ids, fold, 123, task, by_pid=by_pid)
        assert set(np.unique(w)) <= {0.0, 1.0}
        held += 1 - w
        units = np.unique(ids[w == 0]) if by_pid else np.nonzero(w == 0)[0]
        sizes.append(len(units))
        if by_pid:  # whole strata are held out together
            for u in units:
                assert np.all(w[ids == u] == 0)
    assert np.all(held == 1)  # every row held out exactly once over the folds
    base = n_units // fold
    assert sizes == [base + 1 if i < n_units - base * fold else base for i in range(fold)]
    # a different seed permutes differently
    w1 = orc.cv_weights(ids, fold, 123, 0, by_pid=by_pid)
    w2 = orc.cv_weights(ids, fold, 124, 0, by_pid=by_pid)
    assert not np.array_equal(w1, w2)


def test_grid_points_log_uniform():
    pts = np.array([orc.grid_point(1e-4, 1.0, 16, s) for s in range(16)])
    assert abs(pts[0] - 1e-4) < 1e-18 and abs(pts[-1] - 1.0) < 1e-12
    r = pts[1:] / pts[:-1]
    assert np.allclose(r, r[0], rtol=1e-12)
    assert orc.grid_point(1e-4, 1.0, 1, 0) == 1.0


def test_oracle_on_synthetic_lr_small():
    p, truth = synth.lr_indicator(K=2000, P=100, p=0.05, n_true=10, seed=3)
    spec = default_spec(p, prior=PRIOR_LAPLACE, variance=1.0)
    beta, info = orc.fit(p, spec)
    assert info["return_flag"] == 0 and info["iterations"] > 1
    # strong true effects are recovered with the right sign
    big = np.abs(truth["beta_vals"]) > 0.6
    est = beta[1 + truth["beta_cols"][big]]
    assert np.mean(np.sign(est) == np.sign(truth["beta_vals"][big])) >= 0.8
