"""GPU tests of the continuous-batching fit server.

Requests of different models and sizes share launches; each result matches the
CPU oracle run alone (FP64: 1e-8 relative on coefficients and log likelihood,
same iteration count and return flag) and is bit-identical to the same request
served alone.  A short time quantum makes fits pause and resume between
launches many times.  Admission respects the memory budget.
"""
import numpy as np
import pytest

from ccd_amd import api, synth
from ccd_amd.data import CONDITIONAL_LOGISTIC, default_spec
from oracle import oracle as orc
from tests.test_gpu_plain import SYNTH, check_fit

pytestmark = pytest.mark.gpu


def _requests():
    """(problem, spec) pairs: every model, two prior strengths each."""
    out = []
    for name in ("lr", "poisson", "cox", "clr"):
        p, prior, var, maxit = SYNTH[name]()
        for scale in (1.0, 0.1):
            out.append((p, default_spec(p, prior=prior, variance=var * scale, max_iterations=maxit)))
    return out


@pytest.mark.parametrize("threads", [256, 512])
def test_serve_mixed_requests_match_oracle(threads):
    # 512 threads run 4 items per lane (two waves per SIMD)
    reqs = _requests()
    with api.Server(quantum_ms=2.0, threads=threads) as srv:
        ids = [srv.submit(p, sp) for p, sp in reqs]
        got = [srv.wait(i) for i in ids]
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    for (p, sp), (b, r, info) in zip(reqs, got):
        bc, ic = orc.fit(p, sp)
        check_fit(b, r, bc, ic)
        assert info["latency"] >= info["queued"] >= 0.0
        assert info["launches"] >= 1


def test_serve_same_alone_and_in_the_mix():
    """Same bits alone and in the mix; gangs are capped at 4 workgroups, which
    every fit large enough for a gang gets both ways (a gang-size change only
    reorders the per-coordinate sums)."""
    reqs = _requests()
    with api.Server(quantum_ms=0.05, max_gang=4) as srv:  # fits pause and resume many times
        ids = [srv.submit(p, sp) for p, sp in reqs]
        mix = [srv.wait(i) for i in ids]
    assert max(m[2]["launches"] for m in mix) > 3
    for k in (0, 3, 4, 7):
        p, sp = reqs[k]
        with api.Server(quantum_ms=50.0, max_gang=4) as srv:
            b, r, _ = srv.wait(srv.submit(p, sp))
        assert np.array_equal(b, mix[k][0])
        assert r["iterations"] == mix[k][1]["iterations"]
        assert r["log_likelihood"] == mix[k][1]["log_likelihood"]


def test_serve_matches_batched_engine():
    p, prior, var, maxit = SYNTH["clr"]()
    specs = [default_spec(p, prior=prior, variance=var * s, max_iterations=maxit)
             for s in (0.5, 1.0, 2.0)]
    bb, ib = api.fit_batch(api.Data(p), specs, bundle=1)  # one fit per slot, as served
    with api.Server() as srv:
        got = [srv.wait(srv.submit(p, sp)) for sp in specs]
    for i, (b, r, _) in enumerate(got):
        assert np.array_equal(b, bb[i])
        assert r["iterations"] == ib[i]["iterations"]


def test_serve_resident_data_with_cv_weights():
    p, prior, var, maxit = SYNTH["clr"]()
    d = api.Data(p)
    ids_ = p.strata
    with api.Server() as srv:
        rids, cases = [], []
        for fold in range(3):
            w = orc.cv_weights(ids_, 5, 99, fold, by_pid=p.model == CONDITIONAL_LOGISTIC)
            sp = default_spec(p, prior=prior, variance=var, weights=w, max_iterations=maxit)
            rids.append(srv.submit(d, sp, heldout=1.0 - w))
            cases.append((sp, 1.0 - w))
        got = [srv.wait(i) for i in rids]
    for (sp, ho), (b, r, _) in zip(cases, got):
        bc, ic = orc.fit(p, sp, heldout=ho)
        check_fit(b, r, bc, ic)
        np.testing.assert_allclose(r["pred_log_likelihood"], ic["pred_log_likelihood"], rtol=1e-8)


def test_serve_admission_by_footprint():
    p, prior, var, maxit = SYNTH["lr"]()
    sp = default_spec(p, prior=prior, variance=var)
    with api.Server() as srv:
        _, _, info = srv.wait(srv.submit(p, sp))
    fp = info["footprint"]
    budget = int(2.5 * fp)  # two requests in flight at most
    with api.Server(hbm_budget_bytes=budget) as srv:
        ids = [srv.submit(p, sp) for _ in range(6)]
        got = [srv.wait(i) for i in ids]
        st = srv.stats()
    assert st["arena_peak"] <= st["arena_bytes"] <= budget
    assert st["arena_peak"] >= fp
    assert max(g[2]["queued"] for g in got) > 0.0
    for b, _, _ in got[1:]:
        assert np.array_equal(b, got[0][0])
    # a request larger than the whole budget fails cleanly; the server keeps serving
    with api.Server(hbm_budget_bytes=fp // 2) as srv:
        rid = srv.submit(p, sp)
        with pytest.raises(api.CcdhError, match="memory budget"):
            srv.wait(rid)
        assert srv.stats()["failed"] == 1


def test_serve_mixed_precision():
    reqs = _requests()[::2]
    with api.Server(precision=api.MIXED) as srv:
        got = [srv.wait(srv.submit(p, sp)) for p, sp in reqs]
    for (p, sp), (b, r, _) in zip(reqs, got):
        bc, ic = orc.fit(p, sp)
        check_fit(b, r, bc, ic, rtol=2e-4, atol=2e-5, ll_rtol=1e-6, same_path=False)


@pytest.mark.parametrize("name", ["poisson", "cox"])
def test_serve_gangs_match_oracle(name):
    """A request large enough for a gang (Poisson: about 240k nonzeros, 4
    workgroups; Cox: a 6,000-pid risk-set scan per coordinate, up to 32) matches
    the oracle when served with a short quantum, alone and beside others."""
    p, prior, var, maxit = SYNTH[name]()
    sp = default_spec(p, prior=prior, variance=var, max_iterations=maxit)
    q, qprior, qvar, qmaxit = SYNTH["lr"]()
    qs = default_spec(q, prior=qprior, variance=qvar, max_iterations=qmaxit)
    with api.Server(quantum_ms=1.0) as srv:
        b, r, info = srv.wait(srv.submit(p, sp))
        ids = [srv.submit(p, sp), srv.submit(q, qs), srv.submit(p, sp)]
        more = [srv.wait(i) for i in ids]
    bc, ic = orc.fit(p, sp)
    check_fit(b, r, bc, ic)
    for bm, rm, _ in (more[0], more[2]):
        check_fit(bm, rm, bc, ic)


def _full_range_requests():
    """A design per request at the serving configuration's ranges: logistic 1e4 rows,
    logistic 2e5 x 5e4, Poisson 1e6 x 1e5 with offset, Cox 1e6 x 2e4, 20-60
    nonzeros per row, Zipf(1.1) columns; the large fits capped at a few sweeps
    so the oracle (Poisson) and the repeated serves stay within the test's time."""
    from ccd_amd.data import PRIOR_LAPLACE, PRIOR_NORMAL
    lr_small, _ = synth.lr_zipf(K=10_000, P=1_000, per_row=20, n_true=100, seed=71)
    lr_mid, _ = synth.lr_zipf(K=200_000, P=50_000, per_row=40, n_true=100, seed=72)
    pois, _ = synth.poisson_zipf(K=1_000_000, P=100_000, per_row=30, n_true=100, base=-5.0,
                                 seed=73)
    cox, _ = synth.cox_zipf(K=1_000_000, P=20_000, per_row=60, n_true=100, seed=74)
    return [
        (lr_small, default_spec(lr_small, prior=PRIOR_LAPLACE, variance=0.5), True),
        (lr_mid, default_spec(lr_mid, prior=PRIOR_LAPLACE, variance=0.1, max_iterations=3), False),
        (pois, default_spec(pois, prior=PRIOR_NORMAL, variance=0.1, max_iterations=2), True),
        (cox, default_spec(cox, prior=PRIOR_LAPLACE, variance=0.05, max_iterations=2), False),
    ]


@pytest.mark.timeout(600)
def test_serve_full_range_requests_alone_and_in_the_mix():
    """The serving configuration's row and column ranges, a design per request:
    every result is bit-identical served alone and in the mix (gangs capped at 4
    workgroups both ways), and the requests the oracle finishes in time (the
    small logistic fit to convergence, the Poisson fit's first two sweeps) match
    it to 1e-8."""
    reqs = _full_range_requests()
    print("full-range serve: mix", flush=True)
    with api.Server(quantum_ms=5.0, max_gang=4) as srv:
        ids = [srv.submit(p, sp) for p, sp, _ in reqs]
        mix = [srv.wait(i) for i in ids]
    for k, (p, sp, check) in enumerate(reqs):
        print("full-range serve: alone", k, p.K, p.J, flush=True)
        with api.Server(quantum_ms=50.0, max_gang=4) as srv:
            b, r, _ = srv.wait(srv.submit(p, sp))
        assert np.array_equal(b, mix[k][0]), k
        assert r["iterations"] == mix[k][1]["iterations"]
        if check:
            bc, ic = orc.fit(p, sp)
            check_fit(mix[k][0], mix[k][1], bc, ic)
