"""GPU parity of the batched (device-resident sweep) engine.

Every fit of a batch matches the CPU oracle run alone to 1e-8 relative
(coefficients, log likelihood, predictive log likelihood) with the same
iteration count and return flag; a fit's coefficients are bit-identical alone
and inside a batch of other fits at the same gang and bundle size (either only
reorders the per-coordinate sums); graph replay and direct launches agree.
"""
import numpy as np
import pytest

from ccd_amd import api, synth
from ccd_amd.data import CONDITIONAL_LOGISTIC, PRIOR_LAPLACE, PRIOR_NORMAL, default_spec
from oracle import oracle as orc
from tests import fixtures
from tests.test_gpu_plain import SYNTH, check_fit

pytestmark = pytest.mark.gpu


def cv_specs(p, prior, var, maxit, n_fold=4, grid=(0.05, 0.5)):
    by_pid = p.model == CONDITIONAL_LOGISTIC
    ids = p.strata if by_pid else np.zeros(p.K, np.int32)
    specs, held = [], []
    for v in grid:
        for fold in range(n_fold):
            w = orc.cv_weights(ids, n_fold, 17, fold, by_pid=by_pid)
            specs.append(default_spec(p, prior=prior, variance=v * var, weights=w,
                                      max_iterations=maxit))
            held.append(1.0 - w)
    return specs, held


@pytest.mark.parametrize("name", sorted(SYNTH))
def test_batch_cv_fits_match_oracle(name):
    p, prior, var, maxit = SYNTH[name]()
    specs, held = cv_specs(p, prior, var, maxit)
    d = api.Data(p)
    bg, ig = api.fit_batch(d, specs, heldouts=held)
    for i, sp in enumerate(specs):
        bc, ic = orc.fit(p, sp, heldout=held[i])
        check_fit(bg[i], ig[i], bc, ic)
        np.testing.assert_allclose(ig[i]["pred_log_likelihood"], ic["pred_log_likelihood"],
                                   rtol=1e-8)


@pytest.mark.parametrize("name", fixtures.names())
def test_batch_fixtures_match_oracle(name):
    p, extra = fixtures.load(name)
    specs = [default_spec(p), default_spec(p, prior=PRIOR_LAPLACE, variance=1.0),
             default_spec(p, prior=PRIOR_NORMAL, variance=0.5)]
    if "weights" in extra:
        for s in specs:
            s.weights = extra["weights"]
    bg, ig = api.fit_batch(api.Data(p), specs)
    for i, sp in enumerate(specs):
        bc, ic = orc.fit(p, sp)
        check_fit(bg[i], ig[i], bc, ic)


@pytest.mark.parametrize("name", sorted(SYNTH))
@pytest.mark.parametrize("gang", [1, 4])
def test_batch_same_alone_and_in_the_mix(name, gang):
    """Same bits alone and among other fits at the same gang size, one fit per
    slot (a gang-size or bundle change reorders the per-coordinate sums)."""
    p, prior, var, maxit = SYNTH[name]()
    specs, _ = cv_specs(p, prior, var, maxit, n_fold=3, grid=(0.1, 1.0, 10.0))
    d = api.Data(p)
    b_mix, i_mix = api.fit_batch(d, specs, max_gang=gang, gang_min_nnz=1, bundle=1)
    for i in (0, 4, len(specs) - 1):
        b_one, i_one = api.fit_batch(d, [specs[i]], max_gang=gang, gang_min_nnz=1, bundle=1)
        assert np.array_equal(b_one[0], b_mix[i])
        assert i_one[0]["iterations"] == i_mix[i]["iterations"]
        assert i_one[0]["log_likelihood"] == i_mix[i]["log_likelihood"]


def test_batch_graph_and_direct_launch_agree():
    p, prior, var, maxit = SYNTH["clr"]()
    specs, _ = cv_specs(p, prior, var, maxit, n_fold=2, grid=(1.0,))
    d = api.Data(p)
    b1, _ = api.fit_batch(d, specs, use_graph=1, sweeps_per_sync=3, max_gang=1)
    b2, _ = api.fit_batch(d, specs, use_graph=0, sweeps_per_sync=1, max_gang=1)
    assert np.array_equal(b1, b2)


@pytest.mark.parametrize("threads", [256, 512])
def test_batch_workgroup_sizes(threads):
    p, prior, var, maxit = SYNTH["lr"]()
    spec = default_spec(p, prior=prior, variance=var)
    bc, ic = orc.fit(p, spec)
    bg, ig = api.fit_batch(api.Data(p), [spec], threads=threads)
    check_fit(bg[0], ig[0], bc, ic)


@pytest.mark.parametrize("name", sorted(SYNTH))
def test_batch_mixed_precision(name):
    p, prior, var, maxit = SYNTH[name]()
    spec = default_spec(p, prior=prior, variance=var, max_iterations=maxit)
    bc, ic = orc.fit(p, spec)
    bg, ig = api.fit_batch(api.Data(p, precision=api.MIXED), [spec])
    check_fit(bg[0], ig[0], bc, ic, rtol=2e-4, atol=2e-5, ll_rtol=1e-6, same_path=False)


@pytest.mark.parametrize("solo", [0, -1])
@pytest.mark.parametrize("name", ["lr", "poisson", "clr", "cox"])
@pytest.mark.parametrize("gang", [2, 8, 64])
def test_batch_gangs_match_oracle(name, gang, solo):
    """Fits run by gangs of workgroups (row shards, one exchange per coordinate;
    solo 0: runs of short columns on the lead alone) match the oracle to the FP64
    tolerance with the same iterations."""
    p, prior, var, maxit = SYNTH[name]()
    specs, held = cv_specs(p, prior, var, maxit, n_fold=2, grid=(1.0,))
    d = api.Data(p)
    bg, ig = api.fit_batch(d, specs, heldouts=held, max_gang=gang, gang_min_nnz=1, solo=solo)
    for i, sp in enumerate(specs):
        bc, ic = orc.fit(p, sp, heldout=held[i])
        check_fit(bg[i], ig[i], bc, ic)
        np.testing.assert_allclose(ig[i]["pred_log_likelihood"], ic["pred_log_likelihood"],
                                   rtol=1e-8)


@pytest.mark.parametrize("name", ["lr", "poisson", "clr", "cox"])
def test_batch_shared_fold_weights(name):
    """A CV grid given one weights array per fold: each array is uploaded once and
    the fold's other fits take X'Y and the group weight sums of its first fit.
    The fits equal those of the same batch given a copy of the weights per fit
    (bit for bit on one workgroup per fit) and match the oracle (also on gangs)."""
    import dataclasses
    p, prior, var, maxit = SYNTH[name]()
    by_pid = p.model == CONDITIONAL_LOGISTIC
    ids = p.strata if by_pid else np.zeros(p.K, np.int32)
    fold_w = [orc.cv_weights(ids, 3, 17, f, by_pid=by_pid) for f in range(3)]
    shared = [default_spec(p, prior=prior, variance=v * var, weights=fold_w[f],
                           max_iterations=maxit) for v in (0.05, 0.5, 1.0) for f in range(3)]
    copies = [dataclasses.replace(sp, weights=sp.weights.copy()) for sp in shared]
    d = api.Data(p)
    bs, infs = api.fit_batch(d, shared, max_gang=1)
    assert api.batch_last_stats()["shared_setups"] == 6
    bc, infc = api.fit_batch(d, copies, max_gang=1)
    assert api.batch_last_stats()["shared_setups"] == 0
    np.testing.assert_array_equal(bs, bc)
    assert [r["iterations"] for r in infs] == [r["iterations"] for r in infc]
    bg, infg = api.fit_batch(d, shared, max_gang=8, gang_min_nnz=1)
    assert api.batch_last_stats()["shared_setups"] == 6
    for i, sp in enumerate(shared):
        br, ir = orc.fit(p, sp)
        check_fit(bs[i], infs[i], br, ir)
        check_fit(bg[i], infg[i], br, ir)


def _stratified_cox(n_strata=5):
    """The synthetic Cox design split into contiguous row blocks as strata (rows
    stay sorted by stratum, -time, y), so risk sets restart inside rank ranges."""
    import dataclasses
    p, _, _, _ = SYNTH["cox"]()
    strata = np.minimum(np.arange(p.K) * n_strata // p.K, n_strata - 1).astype(np.int32)
    return dataclasses.replace(p, strata=strata)


@pytest.mark.parametrize("solo", [0, -1])
@pytest.mark.parametrize("gang", [4, 32])
def test_batch_cox_gangs_strata_and_weights(gang, solo):
    """Cox gangs: the segmented risk-set scan carried across ranks restarts at
    strata inside and at the edges of rank ranges; zero-weight rows (pid N)
    under CV weights; the log-likelihood convergence criterion."""
    p = _stratified_cox()
    _, prior, var, maxit = SYNTH["cox"]()
    specs = []
    for fold in range(2):
        w = orc.cv_weights(np.zeros(p.K, np.int32), 4, 23, fold, by_pid=False)
        specs.append(default_spec(p, prior=prior, variance=var, weights=w, max_iterations=maxit))
    specs.append(default_spec(p, prior=prior, variance=var, max_iterations=maxit,
                              convergence_type=1))  # log-likelihood criterion: gang LL scan
    d = api.Data(p)
    bg, ig = api.fit_batch(d, specs, max_gang=gang, gang_min_nnz=1, solo=solo)
    st = api.batch_last_stats()
    assert max(t[2] for t in st["timeline"]) == gang
    for i, sp in enumerate(specs):
        bc, ic = orc.fit(p, sp)
        check_fit(bg[i], ig[i], bc, ic)


@pytest.mark.parametrize("name", [n for n in fixtures.names() if n.startswith("cox")])
def test_batch_cox_gangs_tiny_tables(name):
    """Gangs of 8 over 7-row tables: most ranks own no rows and no pids."""
    p, extra = fixtures.load(name)
    specs = [default_spec(p), default_spec(p, prior=PRIOR_LAPLACE, variance=1.0)]
    if "weights" in extra:
        for s in specs:
            s.weights = extra["weights"]
    bg, ig = api.fit_batch(api.Data(p), specs, max_gang=8, gang_min_nnz=1)
    for i, sp in enumerate(specs):
        bc, ic = orc.fit(p, sp)
        check_fit(bg[i], ig[i], bc, ic)


@pytest.mark.parametrize("name", ["lr", "poisson", "clr"])
@pytest.mark.parametrize("bundle,gang", [(2, 1), (4, 1), (8, 1), (8, 4), (4, 16)])
def test_batch_bundles_match_oracle(name, bundle, gang):
    """Fits bundled into slots (lanes carry (nonzero, fit) pairs over interleaved
    state, the fits step in lock step, a converged fit drops out while the others
    go on) match the oracle, alone and on gangs."""
    p, prior, var, maxit = SYNTH[name]()
    specs, held = cv_specs(p, prior, var, maxit, n_fold=4, grid=(0.05, 0.5, 5.0))
    d = api.Data(p)
    bg, ig = api.fit_batch(d, specs, heldouts=held, max_gang=gang, gang_min_nnz=1, bundle=bundle)
    st = api.batch_last_stats()
    assert st["bundles"] < len(specs)
    if name != "poisson":  # (the Poisson design runs every fit to its 40-sweep cap)
        assert len({r["iterations"] for r in ig}) > 1  # fits of a bundle finish apart
    for i, sp in enumerate(specs):
        bc, ic = orc.fit(p, sp, heldout=held[i])
        check_fit(bg[i], ig[i], bc, ic)
        np.testing.assert_allclose(ig[i]["pred_log_likelihood"], ic["pred_log_likelihood"],
                                   rtol=1e-8)


def test_batch_bundle_size_changes_only_rounding():
    p, prior, var, maxit = SYNTH["clr"]()
    specs, _ = cv_specs(p, prior, var, maxit, n_fold=4, grid=(0.5, 1.0))
    d = api.Data(p)
    b1, i1 = api.fit_batch(d, specs, bundle=1)
    b8, i8 = api.fit_batch(d, specs, bundle=8)
    np.testing.assert_allclose(b8, b1, rtol=1e-10, atol=1e-13)
    assert [r["iterations"] for r in i8] == [r["iterations"] for r in i1]


def test_batch_gang_size_changes_only_rounding():
    """The gang size changes only the order of the per-coordinate sums."""
    p, prior, var, maxit = SYNTH["clr"]()
    specs, _ = cv_specs(p, prior, var, maxit, n_fold=3, grid=(0.5,))
    d = api.Data(p)
    b1, i1 = api.fit_batch(d, specs, max_gang=1)
    b16, i16 = api.fit_batch(d, specs, max_gang=16, gang_min_nnz=1)
    np.testing.assert_allclose(b16, b1, rtol=1e-10, atol=1e-13)
    assert [r["iterations"] for r in i16] == [r["iterations"] for r in i1]


def test_batch_gangs_with_concurrent_kernels():
    """Gang fits stay exact while another host thread launches kernels on other
    streams (every gang runs inside one XCD, so no cache line is shared between
    XCDs)."""
    import threading
    p, prior, var, maxit = SYNTH["poisson"]()
    specs = [default_spec(p, prior=prior, variance=var * s, max_iterations=maxit)
             for s in (1.0, 0.1)]
    refs = [orc.fit(p, sp) for sp in specs]
    q, qprior, qvar, _ = SYNTH["lr"]()
    qspec = default_spec(q, prior=qprior, variance=qvar)
    dp, dq = api.Data(p), api.Data(q)
    stop = threading.Event()

    def other_work():
        while not stop.is_set():
            api.fit_plain(dq, qspec)

    th = threading.Thread(target=other_work)
    th.start()
    try:
        for _ in range(3):
            bg, ig = api.fit_batch(dp, specs, max_gang=4, gang_min_nnz=1)
            for i, (bc, ic) in enumerate(refs):
                check_fit(bg[i], ig[i], bc, ic)
    finally:
        stop.set()
        th.join()


def test_batch_gangs_long_columns_many_fits():
    """Gangs on heavy Zipf columns (several passes per rank, rank positions that
    move from column to column) with many fits in flight agree with one
    workgroup per fit."""
    from ccd_amd import synth
    p, _ = synth.lr_zipf(K=120_000, P=600, per_row=20, seed=9)
    specs = [default_spec(p, prior=PRIOR_LAPLACE, variance=v, max_iterations=3)
             for v in np.geomspace(0.01, 1.0, 48)]
    d = api.Data(p)
    b1, i1 = api.fit_batch(d, specs, max_gang=1)
    for _ in range(2):
        b2, i2 = api.fit_batch(d, specs, max_gang=2, gang_min_nnz=1)
        np.testing.assert_allclose(b2, b1, rtol=1e-10, atol=1e-13)
        assert [r["iterations"] for r in i2] == [r["iterations"] for r in i1]


_ZIPF = {}


def _zipf_case(name):
    """Zipf designs (about 10^6 nonzeros) whose head columns take several passes
    of a workgroup and whose tail is short columns (blocks, lead-alone runs), CV
    specs over two folds and two variances, and the oracle's fits (computed once).
    The Poisson fits stop at 8 sweeps (coordinate descent converges slowly on
    this design): both engines return MAX_ITERATIONS there."""
    if name not in _ZIPF:
        if name == "clr":
            p, _ = synth.clr_strata(n_strata=6000, P=600, per_row=25, s=1.1, max_controls=10,
                                    n_true=40, sd=0.5, seed=12)
            maxit = 200
        else:
            p, _ = synth.poisson_zipf(K=40_000, P=600, per_row=25, s=1.1, n_true=40, sd=0.3,
                                      seed=12)
            maxit = 8
        specs, _ = cv_specs(p, PRIOR_LAPLACE, 1.0, maxit, n_fold=2, grid=(0.02, 1.0))
        _ZIPF[name] = (p, specs, [orc.fit(p, sp) for sp in specs])
    return _ZIPF[name]


@pytest.mark.parametrize("gangs", [False, True])
@pytest.mark.parametrize("name", ["clr", "poisson"])
def test_batch_short_quantum_matches_oracle(name, gangs):
    """A 0.3 ms launch quantum: the fits pause hundreds of times, at the checks
    before multi-pass columns (priced by the workgroup's time per nonzero), at the
    cycle's end, inside lead-alone runs and at a gang's exchanges, and resume in
    the next launch; every fit still matches the oracle with the same iterations.
    Without gangs every slot is one workgroup (gangs allowed, none grown)."""
    p, specs, ref = _zipf_case(name)
    d = api.Data(p)
    b, info = api.fit_batch(d, specs, max_gang=8, gang_min_nnz=1 if gangs else 10**12,
                            quantum_ms=0.3)
    st = api.batch_last_stats()
    assert st["launches"] > 4 * max(r["iterations"] for r in info)
    for i, (bc, ic) in enumerate(ref):
        check_fit(b[i], info[i], bc, ic)


def test_batch_pass_width_changes_only_rounding():
    """Long-column passes of 4 or 8 nonzeros per lane give the same fits up to
    the order of the sums."""
    p, prior, var, maxit = SYNTH["clr"]()
    specs, _ = cv_specs(p, prior, var, maxit, n_fold=2, grid=(0.5,))
    d = api.Data(p)
    b8, i8 = api.fit_batch(d, specs, pass_items=8)
    b4, i4 = api.fit_batch(d, specs, pass_items=4)
    np.testing.assert_allclose(b4, b8, rtol=1e-10, atol=1e-13)
    assert [r["iterations"] for r in i4] == [r["iterations"] for r in i8]


@pytest.mark.parametrize("name", ["clr", "poisson"])
def test_batch_512_threads_4_items_matches_oracle(name):
    # two waves per SIMD: the same pass width as 256 x 8, other reduction order
    p, prior, var, maxit = SYNTH[name]()
    specs, held = cv_specs(p, prior, var, maxit, n_fold=2, grid=(1.0, 0.1))
    d = api.Data(p)
    b5, i5 = api.fit_batch(d, specs, heldouts=held, threads=512, pass_items=4, gang_min_nnz=1)
    b2, i2 = api.fit_batch(d, specs, heldouts=held, gang_min_nnz=1)
    d.close()
    for i, sp in enumerate(specs):
        bc, ic = orc.fit(p, sp, heldout=held[i])
        check_fit(b5[i], i5[i], bc, ic)
    np.testing.assert_allclose(b5, b2, rtol=1e-10, atol=1e-13)


def test_batch_autotune_matches_oracle_and_caches():
    p, prior, var, maxit = SYNTH["poisson"]()
    specs, held = cv_specs(p, prior, var, maxit, n_fold=2, grid=(1.0, 0.1))
    d = api.Data(p)
    bg, ig = api.fit_batch(d, specs, heldouts=held, autotune=1, gang_min_nnz=1)
    st = api.batch_last_stats()
    assert st["tuned_pass_items"] in (4, 8) and st["tuned_max_gang"] >= 1
    assert (st["tuned_threads"], st["tuned_pass_items"]) in ((256, 8), (512, 4))
    assert st["tune_ms"] > 0 and max(st["tune_rates"]) > 0
    for i, sp in enumerate(specs):
        bc, ic = orc.fit(p, sp, heldout=held[i])
        check_fit(bg[i], ig[i], bc, ic)
    b2, _ = api.fit_batch(d, specs, heldouts=held, autotune=1, gang_min_nnz=1)
    assert api.batch_last_stats()["tune_ms"] == 0.0  # the cached choice is reused
    np.testing.assert_allclose(b2, bg, rtol=1e-10, atol=1e-13)


@pytest.mark.parametrize("name", ["lr", "clr", "cox"])
def test_batch_bootstrap_replicates_match_oracle(name):
    """Bootstrap replicates (resampling weights per replicate, all warm-started from
    the point estimate) run as one batch; each matches the oracle fit of the same
    weights and start."""
    p, prior, var, maxit = SYNTH[name]()
    spec = default_spec(p, prior=prior, variance=var, max_iterations=maxit)
    b_hat, _ = orc.fit(p, spec)
    ids = p.strata if p.model == CONDITIONAL_LOGISTIC else np.arange(p.K, dtype=np.int32)
    d = api.Data(p)
    bg, ig = api.bootstrap_batch(d, spec, ids, replicates=6, seed=2024, warm_start=b_hat)
    for r in range(6):
        w = orc.bootstrap_weights(ids, 2024, r)
        sp = default_spec(p, prior=prior, variance=var, max_iterations=maxit, weights=w)
        sp.beta0 = b_hat.copy()
        bc, ic = orc.fit(p, sp)
        check_fit(bg[r], ig[r], bc, ic)


def test_batch_profile_likelihood_matches_oracle():
    """Profile log likelihood of one coordinate at seven values as one batch of
    fits with that coordinate fixed (warm start at the mode); each matches the
    oracle, and the profile peaks at the unpenalized estimate."""
    p, _, _, _ = SYNTH["lr"]()
    spec = default_spec(p, max_iterations=1000)  # no prior: the profile of the MLE
    b_hat, i_hat = orc.fit(p, spec)
    j = int(np.argmax(np.abs(b_hat[1:]))) + 1
    values = b_hat[j] + np.linspace(-0.6, 0.6, 7)
    d = api.Data(p)
    ll, bg, ig = api.profile_batch(d, spec, j, values, b_hat)
    for k, x in enumerate(values):
        sp = default_spec(p, max_iterations=1000)
        sp.fixed = sp.fixed.copy()
        sp.fixed[j] = 1
        sp.beta0 = b_hat.copy()
        sp.beta0[j] = x
        bc, ic = orc.fit(p, sp)
        check_fit(bg[k], ig[k], bc, ic)
        assert bg[k][j] == x
    assert np.argmax(ll) == 3
    np.testing.assert_allclose(ll[3], i_hat["log_likelihood"], rtol=1e-8)
    assert np.all(np.diff(ll[:4]) > 0) and np.all(np.diff(ll[3:]) < 0)


def test_batch_auto_search_cv_matches_oracle_search():
    """Auto-search cross-validation with each step's folds batched proposes the
    same variances, scores and end point as the same search driven by oracle fits."""
    p, prior, var, maxit = SYNTH["clr"]()
    spec = default_spec(p, prior=prior, variance=var, max_iterations=maxit)
    d = api.Data(p)
    v_gpu, e_gpu, trace = api.auto_search_cv(d, spec, p.strata, n_fold=5, seed=77,
                                             start_variance=0.05, max_steps=8)
    assert len(trace) >= 3
    # the same loop over oracle fits
    weights = [orc.cv_weights(p.strata, 5, 77, t, by_pid=True) for t in range(5)]
    vv, tried = 0.05, []
    for k in range(len(trace)):
        preds = []
        for w in weights:
            sp = default_spec(p, prior=prior, variance=vv, weights=w, max_iterations=maxit)
            _, ic = orc.fit(p, sp, heldout=1.0 - w)
            preds.append(ic["pred_log_likelihood"] if ic["return_flag"] == 0 else np.nan)
        pr = np.array(preds)
        pr = pr[pr == pr]
        mean = pr.mean()
        tried.append((vv, mean, np.sqrt(np.mean(pr * pr) - mean * mean)))
        np.testing.assert_allclose(trace[k][:2], tried[-1][:2], rtol=1e-7)
        # (the spread is sqrt(E[x^2] - mean^2) as the reference computes it: cancellation)
        np.testing.assert_allclose(trace[k][2], tried[-1][2], rtol=1e-4)
        cont, vv, ev = orc.search_step(tried)
        if not cont:
            break
    assert len(tried) == len(trace)
    np.testing.assert_allclose([v_gpu, e_gpu], [vv, ev], rtol=1e-6)


def test_device_cox_groups_match_oracle():
    """Risk-set groups built on the device (as every batched and served Cox fit
    builds them) equal the oracle's grouping exactly: tied and untied times,
    strata, zero-weight rows (leading ones included), CV and bootstrap weights."""
    probs = [fixtures.load(n)[0] for n in fixtures.names() if n.startswith("cox")]
    probs += [SYNTH["cox"]()[0], _stratified_cox()]
    rng = np.random.default_rng(12)
    for p in probs:
        d = api.Data(p)
        lead = np.ones(p.K)
        lead[: max(1, p.K // 10)] = 0.0
        ws = [None, lead, rng.choice([0.0, 1.0, 2.0], p.K),
              orc.cv_weights(np.zeros(p.K, np.int32), 4, 5, 1, by_pid=False),
              orc.bootstrap_weights(np.arange(p.K, dtype=np.int32), 9, 0)]
        for w in ws:
            pid, reset, first, shard, N = api.cox_groups(d, w)
            No, pido, acc = orc.cox_pids(p, w)
            assert N == No
            assert np.array_equal(pid, pido)
            want = np.zeros(N + 1, np.uint8)
            want[[r for r in acc if r < N]] = 1
            assert np.array_equal(reset, want)
            rows = np.arange(p.K)
            exp_first = np.full(N + 1, p.K)
            np.minimum.at(exp_first, pid[pid < N], rows[pid < N])
            assert np.array_equal(first, exp_first)
            assert shard[0] == 0 and shard[-1] == N and np.all(np.diff(shard) >= 0)
        d.close()


@pytest.mark.parametrize("gang", [4, 16])
def test_batch_cox_long_lead_alone_runs(gang):
    """Cox gangs over more than 2,048 consecutive short columns: the lead sweeps
    them alone and the ranks meet every 2,048 coordinates inside the run; tied
    days make few risk-set pieces.  Matches the oracle, and a gang that sweeps
    every column together (solo=-1)."""
    p, _ = synth.cox_zipf(K=8000, P=5000, per_row=6, n_true=40, seed=31)
    spec = default_spec(p, prior=PRIOR_LAPLACE, variance=0.05, max_iterations=1000)
    d = api.Data(p)
    print("cox lead-alone runs, gang", gang, flush=True)
    bg, ig = api.fit_batch(d, [spec], max_gang=gang, gang_min_nnz=1)
    st = api.batch_last_stats()
    assert max(t[2] for t in st["timeline"]) == gang
    bt, it = api.fit_batch(d, [spec], max_gang=gang, gang_min_nnz=1, solo=-1)
    d.close()
    bc, ic = orc.fit(p, spec)
    check_fit(bg[0], ig[0], bc, ic)
    check_fit(bt[0], it[0], bc, ic)
