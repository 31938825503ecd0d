"""GPU parity of the MM algorithm (ccdh_fit_mm, ccdh_mm_grad_hess, ccdh_compute_xbeta)
against the oracle's restatement of CyclicCoordinateDescent's MM path.

Tolerances (FP64): MM gradients and Hessians within 1e-11 relative of the oracle at
the same coefficients (chunked sums against sequential ones); xBeta from the
transpose within 1e-14 relative (the same per-row order); MM fits within 1e-8
relative on coefficients and log likelihood with the same iteration count and
return flag, and on overshooting designs the same "Non-increasing!" iteration.
Cox (Breslow ties) runs the reference's Survival increments; Poisson is refused.
"""
import numpy as np
import pytest

from ccd_amd import api
from ccd_amd.data import (CONDITIONAL_LOGISTIC, INDICATOR, LOGISTIC, PRIOR_LAPLACE, PRIOR_NONE,
                          PRIOR_NORMAL, SPARSE, build_problem, default_spec)
from oracle import oracle as orc
from tests.test_gpu_plain import check_fit
from tests.test_oracle import _small_problems
from tests.test_oracle_mm import _cox_mm_problems, _mm_problems, _weights

pytestmark = pytest.mark.gpu

_PROBLEMS = _mm_problems()
_IDS = lambda v: v if isinstance(v, str) else ""  # noqa: E731


@pytest.mark.parametrize("name,p", _PROBLEMS, ids=_IDS)
@pytest.mark.parametrize("weighted", [False, True])
def test_mm_grad_hess_matches_oracle(name, p, weighted):
    rng = np.random.default_rng(21)
    beta = rng.normal(0, 0.3, p.J)
    w = _weights(p, rng) if weighted else None
    fixed = np.zeros(p.J, np.uint8)
    fixed[::3] = 1
    d = api.Data(p)
    fh = api.FitHandle(d)
    try:
        fh.set_weights(w, use_cv=weighted)
        fh.fixed_terms(use_cv=weighted)
        fh.compute_xbeta(beta)
        fh.remaining_stats(use_w=weighted)
        g, h = fh.mm_grad_hess(fixed, use_w=weighted)
    finally:
        fh.close()
        d.close()
    g_c, h_c = orc.mm_evaluate(p, beta, weights=w)
    free = fixed == 0
    assert np.all(g[~free] == 0.0) and np.all(h[~free] == 0.0)
    scale = np.abs(g_c).max() + 1.0
    np.testing.assert_allclose(g[free], g_c[free], rtol=1e-11, atol=1e-12 * scale)
    np.testing.assert_allclose(h[free], h_c[free], rtol=1e-11)


def _format_problems():
    rng = np.random.default_rng(22)
    K = 300
    ind = np.nonzero(rng.random(K) < 0.4)[0]
    sp_rows = np.nonzero(rng.random(K) < 0.5)[0]
    sp_vals = rng.normal(0, 2, len(sp_rows))
    y = (rng.random(K) < 0.5).astype(float)
    return [
        ("dense_mix", dict(_small_problems())["lr"]),
        ("sparse_indicator", build_problem(LOGISTIC, [(INDICATOR, ind, None),
                                                      (SPARSE, sp_rows, sp_vals)], y,
                                           intercept=True)),
        ("sparse_only", build_problem(LOGISTIC, [(SPARSE, sp_rows, sp_vals)], y)),
    ] + [(n, p) for n, p in _PROBLEMS if n.endswith("synth")]


@pytest.mark.parametrize("name,p", _format_problems(), ids=_IDS)
def test_compute_xbeta_from_transpose(name, p):
    beta = np.random.default_rng(23).normal(0, 1, p.J)
    d = api.Data(p)
    fh = api.FitHandle(d)
    try:
        fh.compute_xbeta(beta)
        xb = fh.xbeta()
        fh.compute_xbeta(2.0 * beta)  # the copy is built once and reused
        xb2 = fh.xbeta()
    finally:
        fh.close()
        d.close()
    ref = orc.mm_xbeta(p, beta)
    np.testing.assert_allclose(xb, ref, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(xb2, orc.mm_xbeta(p, 2.0 * beta), rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("name,p", _PROBLEMS, ids=_IDS)
@pytest.mark.parametrize("prior,var", [(PRIOR_NONE, 1.0), (PRIOR_LAPLACE, 1.0),
                                       (PRIOR_LAPLACE, 0.01), (PRIOR_NORMAL, 0.1)])
@pytest.mark.parametrize("weighted", [False, True])
def test_mm_fit_matches_oracle(name, p, prior, var, weighted):
    w = _weights(p, np.random.default_rng(24)) if weighted else None
    sp = default_spec(p, prior=prior, variance=var, weights=w)
    _, ic_ccd = orc.fit(p, sp)
    if ic_ccd["return_flag"] != 0:
        pytest.skip("no mode: the unpenalised CLR likelihood is monotone on this design")
    bc, ic = orc.fit(p, sp, algorithm="mm")
    d = api.Data(p)
    try:
        if ic["status"] == -2:
            with pytest.raises(api.NonIncreasing) as err:
                api.fit_plain(d, sp, algorithm="mm")
            assert err.value.iterations == ic["iterations"]
            np.testing.assert_allclose(err.value.beta, bc, rtol=1e-8, atol=1e-10)
        else:
            assert ic["status"] == 0
            bg, ig = api.fit_plain(d, sp, algorithm="mm")
            check_fit(bg, ig, bc, ic)
    finally:
        d.close()


def test_mm_refused_for_poisson():
    p = dict(_small_problems())["poisson"]
    d = api.Data(p)
    fh = api.FitHandle(d)
    try:
        with pytest.raises(api.CcdhError, match="not model-specific"):
            api.fit_plain(d, default_spec(p), algorithm="mm")
        fh.set_weights(None)
        fh.fixed_terms()
        fh.compute_xbeta(np.zeros(p.J))
        fh.remaining_stats()
        with pytest.raises(api.CcdhError, match="not model-specific"):
            fh.mm_grad_hess()
    finally:
        fh.close()
        d.close()


_COX = _cox_mm_problems()


@pytest.mark.parametrize("name,p", _COX, ids=_IDS)
@pytest.mark.parametrize("weights", ["none", "positive", "zeros"])
def test_cox_mm_grad_hess_matches_oracle(name, p, weights):
    """Cox MM increments over each pid's own denominator; with zero weights a fresh
    fit reads the unwritten slot of the zero-weight rows (0/0) in the columns that
    touch them, as the oracle's fresh engine does."""
    rng = np.random.default_rng(25)
    beta = rng.normal(0, 0.4, p.J)
    w = {"none": None, "positive": rng.choice([0.5, 1.0, 2.0], p.K),
         "zeros": (rng.random(p.K) < 0.8).astype(float)}[weights]
    fixed = np.zeros(p.J, np.uint8)
    d = api.Data(p)
    fh = api.FitHandle(d)
    try:
        fh.set_weights(w, use_cv=w is not None)
        fh.fixed_terms(use_cv=w is not None)
        fh.compute_xbeta(beta)
        fh.remaining_stats(use_w=True)
        g, h = fh.mm_grad_hess(fixed, use_w=True)
    finally:
        fh.close()
        d.close()
    g_c, h_c = orc.mm_evaluate(p, beta, weights=w)
    np.testing.assert_array_equal(np.isnan(g), np.isnan(g_c))
    np.testing.assert_array_equal(np.isnan(h), np.isnan(h_c))
    ok = ~np.isnan(g_c)
    scale = np.abs(g_c[ok]).max() + np.abs(h_c[ok]).max() + 1.0 if ok.any() else 1.0
    np.testing.assert_allclose(g[ok], g_c[ok], rtol=1e-11, atol=1e-12 * scale)
    np.testing.assert_allclose(h[ok], h_c[ok], rtol=1e-11, atol=1e-12 * scale)


@pytest.mark.parametrize("name,p", _COX, ids=_IDS)
@pytest.mark.parametrize("prior,var", [(PRIOR_NORMAL, 1.0), (PRIOR_LAPLACE, 0.1)])
def test_cox_mm_fit_matches_oracle(name, p, prior, var):
    """From zero coefficients every Cox MM increment cancels, so the fit stays put;
    from a warm start the tie groups move it.  Either way the same trajectory,
    iterations and return flag, or the same "Non-increasing!" iteration."""
    for start in (np.zeros(p.J), np.random.default_rng(26).normal(0, 0.3, p.J)):
        sp = default_spec(p, prior=prior, variance=var)
        sp.beta0 = start.copy()
        bc, ic = orc.fit(p, sp, algorithm="mm")
        d = api.Data(p)
        try:
            if ic["status"] == -2:
                with pytest.raises(api.NonIncreasing) as err:
                    api.fit_plain(d, sp, algorithm="mm")
                assert err.value.iterations == ic["iterations"]
                np.testing.assert_allclose(err.value.beta, bc, rtol=1e-8, atol=1e-10)
            else:
                assert ic["status"] == 0
                bg, ig = api.fit_plain(d, sp, algorithm="mm")
                check_fit(bg, ig, bc, ic)
                if not start.any():
                    np.testing.assert_array_equal(bg, 0.0)
        finally:
            d.close()


def test_mm_grad_hess_needs_statistics():
    p = dict(_small_problems())["clr"]
    d = api.Data(p)
    fh = api.FitHandle(d)
    try:
        with pytest.raises(api.CcdhError, match="statistics"):
            fh.mm_grad_hess()
    finally:
        fh.close()
        d.close()


def test_mm_refused_on_mixed_precision_data():
    # FP32 row state puts ~1e-7 relative noise into the log posterior, which trips
    # the strict non-increasing check long before the fit converges
    p = dict(_PROBLEMS)["clr_synth"]
    d = api.Data(p, precision=api.MIXED)
    try:
        with pytest.raises(api.CcdhError, match="FP64 data only"):
            api.fit_plain(d, default_spec(p, prior=PRIOR_LAPLACE, variance=0.01), algorithm="mm")
    finally:
        d.close()


def test_mm_kernel_stats():
    p = dict(_PROBLEMS)["clr_synth"]
    sp = default_spec(p, prior=PRIOR_LAPLACE, variance=0.01)
    d = api.Data(p)
    try:
        _, ig = api.fit_plain(d, sp, algorithm="mm")
    finally:
        d.close()
    st = api.mm_last_stats()
    # one chunk pass and one xBeta launch per MM iteration
    assert st["chunk_launches"] == ig["iterations"] == st["xbeta_launches"]
    assert st["chunk_ms"] > 0 and st["xbeta_ms"] > 0
    assert st["nnz"] == p.nnz and st["entries"] == p.nnz and st["rows"] == p.K
    assert st["values"] == 0


def test_transpose_built_once_under_concurrent_clones():
    # fit handles over one data set on two threads reach the lazy build together
    import threading
    p = dict(_PROBLEMS)["clr_synth"]
    rng = np.random.default_rng(25)
    betas = [rng.normal(0, 0.3, p.J) for _ in range(2)]
    d = api.Data(p)
    out, errs = [None, None], []

    def run(i):
        try:
            fh = api.FitHandle(d)
            for _ in range(3):
                fh.compute_xbeta(betas[i])
            out[i] = fh.xbeta()
            fh.close()
        except Exception as e:  # reported below
            errs.append(e)

    ts = [threading.Thread(target=run, args=(i,)) for i in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    d.close()
    assert not errs
    for i in range(2):
        np.testing.assert_allclose(out[i], orc.mm_xbeta(p, betas[i]), rtol=1e-14, atol=1e-14)
