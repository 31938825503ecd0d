"""Seeded synthetic problems with the shapes of the benchmark workloads.

  lr_indicator ...... logistic, iid Bernoulli(p) indicator columns (10,000 x 1,000, p=0.01)
  poisson_zipf ...... Poisson with log-exposure offset, Zipf(s) column frequencies
                      (1,000,000 x 100,000, ~40 nnz/row)
  cox_zipf .......... Breslow Cox, integer-day event times with ties, uniform censoring
                      (2,000,000 x 50,000, 30 nnz/row)
  clr_strata ........ conditional logistic, strata of 1 case + U{1..10} controls
                      (200,000 strata x 20,000 columns, 25 nnz/row)
  lr_zipf ........... logistic over Zipf(s) columns (served requests)
  request_stream .... Poisson arrivals of logistic / Poisson / Cox fit requests with
                      log-uniform rows and columns and U{20..60} nonzeros per row

All generators are vectorised numpy and deterministic in their seed.
"""
import numpy as np

from .data import (COX, CONDITIONAL_LOGISTIC, INDICATOR, LOGISTIC, POISSON, CcdProblem,
                   INTERCEPT, DENSE)

DEFAULT_SEED = 20261203


def _csc_from_pairs(rows, cols, K, P):
    """Sorted, de-duplicated (row, col) pairs -> col_ptr, rows (indicator columns)."""
    key = cols.astype(np.int64) * K + rows.astype(np.int64)
    key = np.unique(key)
    c = (key // K).astype(np.int64)
    r = (key % K).astype(np.int32)
    col_ptr = np.zeros(P + 1, dtype=np.int64)
    np.cumsum(np.bincount(c, minlength=P), out=col_ptr[1:])
    return col_ptr, r


def zipf_pairs(rng, K, P, per_row, s):
    """Each row draws about `per_row` distinct columns with P(col i) ~ 1/(i+1)^s."""
    w = 1.0 / np.arange(1, P + 1, dtype=np.float64) ** s
    cdf = np.cumsum(w / w.sum())
    cdf[-1] = 1.0
    # heavy columns repeat within a row; oversample so the distinct count is ~per_row
    draws = int(np.ceil(per_row * 1.25)) + 2
    out_r, out_c = [], []
    chunk = max(1, 4_000_000 // draws)
    for r0 in range(0, K, chunk):
        r1 = min(K, r0 + chunk)
        n = r1 - r0
        u = rng.random((n, draws))
        c = np.searchsorted(cdf, u).astype(np.int64)
        c.sort(axis=1)
        dup = np.zeros_like(c, dtype=bool)
        dup[:, 1:] = c[:, 1:] == c[:, :-1]
        # keep the first `per_row` distinct columns of each row
        keep = ~dup
        rank = np.cumsum(keep, axis=1)
        keep &= rank <= per_row
        rr = np.broadcast_to(np.arange(r0, r1)[:, None], c.shape)
        out_r.append(rr[keep].astype(np.int32))
        out_c.append(c[keep])
    return np.concatenate(out_r), np.concatenate(out_c)


def _assemble(model, K, P, col_ptr, rows, y, offs=None, strata=None, offset=None,
              intercept=False):
    """Prepend optional offset / intercept columns to an indicator CSC."""
    extra_fmt, extra_rows, extra_vals, extra_len = [], [], [], []
    if offset is not None:
        extra_fmt.append(DENSE)
        extra_rows.append(np.arange(K, dtype=np.int32))
        extra_vals.append(np.asarray(offset, np.float64))
        extra_len.append(K)
    if intercept:
        extra_fmt.append(INTERCEPT)
        extra_rows.append(np.arange(K, dtype=np.int32))
        extra_vals.append(np.ones(K))
        extra_len.append(K)
    nnz = int(col_ptr[-1])
    J = P + len(extra_fmt)
    cp = np.zeros(J + 1, dtype=np.int64)
    off = 0
    for i, n in enumerate(extra_len):
        off += n
        cp[i + 1] = off
    cp[len(extra_len):] = col_ptr + off
    all_rows = np.concatenate(extra_rows + [rows]).astype(np.int32) if extra_rows else rows
    vals = np.concatenate(extra_vals + [np.ones(nnz)]) if extra_vals else np.ones(nnz)
    fmt = np.array(extra_fmt + [INDICATOR] * P, dtype=np.int8)
    return CcdProblem(model, K, J, cp, all_rows, vals, fmt, np.asarray(y, np.float64),
                      np.zeros(K) if offs is None else np.asarray(offs, np.float64),
                      np.zeros(K, np.int32) if strata is None else np.asarray(strata, np.int32),
                      has_offset=offset is not None)


def _xbeta(col_ptr, rows, K, beta_cols, beta_vals):
    eta = np.zeros(K)
    for c, b in zip(beta_cols, beta_vals):
        eta[rows[col_ptr[c]:col_ptr[c + 1]]] += b
    return eta


def lr_indicator(K=10_000, P=1_000, p=0.01, n_true=50, sd=0.5, intercept=-2.0,
                 seed=DEFAULT_SEED):
    rng = np.random.default_rng(seed)
    mask = rng.random((K, P)) < p
    r, c = np.nonzero(mask)
    col_ptr, rows = _csc_from_pairs(r, c, K, P)
    bc = rng.choice(P, n_true, replace=False)
    bv = rng.normal(0.0, sd, n_true)
    eta = _xbeta(col_ptr, rows, K, bc, bv) + intercept
    y = (rng.random(K) < 1.0 / (1.0 + np.exp(-eta))).astype(np.float64)
    prob = _assemble(LOGISTIC, K, P, col_ptr, rows, y, intercept=True)
    return prob, dict(beta_cols=bc, beta_vals=bv)


def poisson_zipf(K=1_000_000, P=100_000, per_row=40, s=1.1, n_true=500, sd=0.3,
                 base=-7.0, seed=DEFAULT_SEED):
    rng = np.random.default_rng(seed)
    r, c = zipf_pairs(rng, K, P, per_row, s)
    col_ptr, rows = _csc_from_pairs(r, c, K, P)
    bc = rng.choice(P, n_true, replace=False)
    bv = rng.normal(0.0, sd, n_true)
    offset = np.log(rng.uniform(30.0, 3650.0, K))
    eta = offset + _xbeta(col_ptr, rows, K, bc, bv) + base
    y = rng.poisson(np.exp(eta)).astype(np.float64)
    prob = _assemble(POISSON, K, P, col_ptr, rows, y, offset=offset, intercept=True)
    return prob, dict(beta_cols=bc, beta_vals=bv)


def cox_zipf(K=2_000_000, P=50_000, per_row=30, s=1.1, n_true=500, sd=0.3, rate=1e-3,
             censor_max=3650.0, seed=DEFAULT_SEED):
    rng = np.random.default_rng(seed)
    r, c = zipf_pairs(rng, K, P, per_row, s)
    col_ptr, rows = _csc_from_pairs(r, c, K, P)
    bc = rng.choice(P, n_true, replace=False)
    bv = rng.normal(0.0, sd, n_true)
    eta = _xbeta(col_ptr, rows, K, bc, bv)
    t_event = np.ceil(rng.exponential(1.0 / (np.exp(eta) * rate)))
    t_cens = np.ceil(rng.uniform(0.0, censor_max, K))
    y = (t_event <= t_cens).astype(np.float64)
    time = np.where(y == 1, t_event, t_cens)
    # sort rows by (stratum, -time, y, rowId) and renumber the covariate rows
    order = np.lexsort((np.arange(K), y, -time))
    inv = np.empty(K, dtype=np.int64)
    inv[order] = np.arange(K)
    new_rows = inv[rows]
    cols = np.repeat(np.arange(P), np.diff(col_ptr))
    col_ptr, rows = _csc_from_pairs(new_rows, cols, K, P)
    prob = _assemble(COX, K, P, col_ptr, rows, y[order], offs=time[order])
    return prob, dict(beta_cols=bc, beta_vals=bv)


def clr_strata(n_strata=200_000, P=20_000, per_row=25, s=1.1, max_controls=10, n_true=200,
               sd=0.5, seed=DEFAULT_SEED):
    rng = np.random.default_rng(seed)
    sizes = 1 + rng.integers(1, max_controls + 1, n_strata)
    K = int(sizes.sum())
    strata = np.repeat(np.arange(n_strata, dtype=np.int32), sizes)
    r, c = zipf_pairs(rng, K, P, per_row, s)
    col_ptr, rows = _csc_from_pairs(r, c, K, P)
    bc = rng.choice(P, n_true, replace=False)
    bv = rng.normal(0.0, sd, n_true)
    eta = _xbeta(col_ptr, rows, K, bc, bv)
    # one case per stratum, chosen with probability proportional to exp(eta)
    g = rng.gumbel(size=K) + eta
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    best = np.maximum.reduceat(g, starts)
    is_case = g == np.repeat(best, sizes)
    y = is_case.astype(np.float64)
    prob = _assemble(CONDITIONAL_LOGISTIC, K, P, col_ptr, rows, y, strata=strata)
    return prob, dict(beta_cols=bc, beta_vals=bv, sizes=sizes)


def lr_zipf(K=100_000, P=10_000, per_row=40, s=1.1, n_true=100, sd=0.5, intercept=-1.0,
            seed=DEFAULT_SEED):
    rng = np.random.default_rng(seed)
    r, c = zipf_pairs(rng, K, P, per_row, s)
    col_ptr, rows = _csc_from_pairs(r, c, K, P)
    bc = rng.choice(P, min(n_true, P), replace=False)
    bv = rng.normal(0.0, sd, len(bc))
    eta = _xbeta(col_ptr, rows, K, bc, bv) + intercept
    y = (rng.random(K) < 1.0 / (1.0 + np.exp(-eta))).astype(np.float64)
    prob = _assemble(LOGISTIC, K, P, col_ptr, rows, y, intercept=True)
    return prob, dict(beta_cols=bc, beta_vals=bv)


def to_long(p, seed=41):
    """Problem p's covariate columns (INDICATOR / SPARSE) as long-format input, as
    convertToCyclopsData takes it: distinct row ids rising with the row and
    covariate ids rising with the column (so every model's row and column order is
    p's own), outcome rows and entries in random order; an empty column is written
    as one zero-valued entry, which creates the column and stores nothing.  Returns the keyword
    arguments of api.Data.from_long; Cox / Poisson times come from p.offs / the
    exponentiated offset column."""
    from .data import DENSE, INTERCEPT, SPARSE
    rng = np.random.default_rng(seed)
    K = p.K
    row_id = 1000 + np.cumsum(rng.integers(1, 5, K))
    cols = np.flatnonzero(np.isin(p.fmt, (INDICATOR, SPARSE)))
    cov_id = 10 * np.cumsum(rng.integers(1, 7, len(cols)))
    counts = np.diff(p.col_ptr)[cols]
    take = np.concatenate([np.arange(p.col_ptr[j], p.col_ptr[j + 1]) for j in cols]) \
        if len(cols) else np.zeros(0, np.int64)
    cov = np.repeat(cov_id, counts)
    erow = row_id[p.rows[take]]
    val = np.where(np.repeat(p.fmt[cols] == SPARSE, counts), p.vals[take], 1.0)
    empty = cov_id[counts == 0]
    cov = np.concatenate([cov, empty])
    erow = np.concatenate([erow, np.full(len(empty), row_id[0])])
    val = np.concatenate([val, np.zeros(len(empty))])
    e = rng.permutation(len(cov))
    o = rng.permutation(K)
    out = dict(model=p.model, row_id=row_id[o], y=p.y[o], covariate_id=cov[e],
               entry_row_id=erow[e], value=val[e],
               add_intercept=bool(np.any(p.fmt == INTERCEPT)))
    if p.model in (CONDITIONAL_LOGISTIC, COX):
        sid = 5 + 3 * np.arange(int(p.strata.max()) + 1)
        out["stratum_id"] = sid[p.strata][o]
    if p.model == COX:
        out["time"] = p.offs[o]
    if p.model == POISSON and p.has_offset and p.fmt[0] == DENSE:
        out["time"] = np.exp(p.vals[p.col_ptr[0]:p.col_ptr[1]])[o]
    return out


def request_stream(n_requests=200, rate=50.0, K_range=(10_000, 200_000),
                   P_range=(1_000, 20_000), per_row_range=(20, 60), cox_K_max=None,
                   pool_per_model=4, models=(LOGISTIC, POISSON, COX), seed=DEFAULT_SEED,
                   progress=None):
    """Served fit requests: a pool of designs per model (rows and columns
    log-uniform in the ranges, nonzeros per row uniform, Zipf(1.1) columns) and
    `n_requests` arrivals at Poisson `rate` per second, each naming a pool design
    and its own prior variance (an outcome model / hyperparameter of its own).
    Returns (pool, requests) with requests = [(arrival_s, pool_index, variance)].
    `progress(design)` is called after each design is generated."""
    rng = np.random.default_rng(seed)

    def logu(lo, hi):
        return int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))

    pool = []
    for m in models:
        for i in range(pool_per_model):
            k_hi = K_range[1] if (m != COX or cox_K_max is None) else min(K_range[1], cox_K_max)
            K = logu(K_range[0], k_hi)
            P = logu(*P_range)
            per_row = int(rng.integers(per_row_range[0], per_row_range[1] + 1))
            sd = int(rng.integers(1 << 30))
            if m == LOGISTIC:
                prob = lr_zipf(K=K, P=P, per_row=per_row, n_true=min(100, P // 10), seed=sd)[0]
            elif m == POISSON:
                prob = poisson_zipf(K=K, P=P, per_row=per_row, n_true=min(100, P // 10),
                                    base=-5.0, seed=sd)[0]
            else:
                prob = cox_zipf(K=K, P=P, per_row=per_row, n_true=min(100, P // 10), seed=sd)[0]
            pool.append(prob)
            if progress is not None:
                progress(prob)
    t = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    idx = rng.integers(0, len(pool), n_requests)
    var = np.exp(rng.uniform(np.log(0.01), np.log(1.0), n_requests))
    return pool, [(float(t[i]), int(idx[i]), float(var[i])) for i in range(n_requests)]


def separable(model=LOGISTIC, nrows=10_000, ncovars=5, effect_sd=1.0, zero_prop=0.9,
              seed=666):
    """A design with one separable covariate, shaped like the reference's
    separability tests (tests/testthat/test-largeBernoulli.R:25-101:
    simulateCyclopsData(nstrata = 1, ncovars = 5), R/Simulation.R:37-97): rows draw
    Poisson(ncovars/100) distinct covariates, effects are log-normal for a tenth of
    the covariates; logistic outcomes ~ Bernoulli(U(0.1, 0.3) * rr), survival
    outcomes with exponential event times (rate U(0.01, 0.03) * rr, 1 + rounded) and
    uniform censoring (1 + round U(0, 499)).  Covariate ncovars + 1 repeats
    covariate 3 on the rows with an event, so its likelihood keeps rising as its
    coefficient grows (the maximum does not exist).  Logistic fits get an intercept;
    Cox rows are sorted by (-time, y, row)."""
    rng = np.random.default_rng(seed)
    sd = effect_sd * (rng.random(ncovars) < 1 - zero_prop)
    rr = np.exp(rng.normal(0.0, 1.0, ncovars) * sd)
    per_row = np.minimum(rng.poisson(ncovars / 100.0, nrows), ncovars)
    rr_row = np.ones(nrows)
    pairs = []
    for i in np.flatnonzero(per_row):
        cols = rng.choice(ncovars, per_row[i], replace=False)
        rr_row[i] = np.prod(rr[cols])
        pairs.extend((i, c) for c in cols)
    r = np.array([a for a, _ in pairs], dtype=np.int64)
    c = np.array([b for _, b in pairs], dtype=np.int64)
    if model == LOGISTIC:
        prob = rng.uniform(0.1, 0.3) * rr_row
        y = (rng.random(nrows) < prob).astype(np.float64)
        time = None
    else:
        rate = rng.uniform(0.01, 0.03) * rr_row
        t_out = 1 + np.round(rng.exponential(1.0 / rate))
        t_cens = 1 + np.round(rng.uniform(0, 499, nrows))
        y = (t_cens > t_out).astype(np.float64)
        time = np.where(t_cens < t_out, t_cens, t_out)
    sep = (c == 2) & (y[r] == 1)
    r = np.concatenate([r, r[sep]])
    c = np.concatenate([c, np.full(int(sep.sum()), ncovars)])
    P = ncovars + 1
    if model == LOGISTIC:
        col_ptr, rows = _csc_from_pairs(r, c, nrows, P)
        return _assemble(LOGISTIC, nrows, P, col_ptr, rows, y, intercept=True)
    order = np.lexsort((np.arange(nrows), y, -time))
    inv = np.empty(nrows, dtype=np.int64)
    inv[order] = np.arange(nrows)
    col_ptr, rows = _csc_from_pairs(inv[r], c, nrows, P)
    return _assemble(COX, nrows, P, col_ptr, rows, y[order], offs=time[order])


def lr_separated(seed=666, n_sep=60):
    """Logistic design of separable(LOGISTIC, seed) whose last covariate is replaced
    by an indicator on `n_sep` rows with an event and no other covariate: the
    likelihood rises without bound in that coefficient alone, so every fit from
    zero marches it up until exp(x beta) / (1 + exp(x beta)) rounds to 1, where
    the gradient and Hessian are both 0 and the step is 0 / 0."""
    from .data import build_problem
    p = separable(LOGISTIC, seed=seed)
    covered = np.zeros(p.K, dtype=bool)
    for j in range(1, p.J):
        covered[p.column(j)[0]] = True
    rng = np.random.default_rng(seed)
    rows = np.sort(rng.choice(np.flatnonzero((p.y == 1) & ~covered), n_sep, replace=False))
    cols = [(INDICATOR, p.column(j)[0], None) for j in range(1, p.J - 1)]
    cols.append((INDICATOR, rows, None))
    return build_problem(LOGISTIC, cols, p.y, intercept=True)


def poisson_overflow(seed=5, K=5000, slope=0.01):
    """Poisson design whose first covariate is SPARSE with values 500..1000 on 1% of
    the rows and counts ~ Poisson(exp(slope * x + 0.3 per indicator)) there, three
    indicator covariates, no intercept.  From zero the first Newton step is clipped
    to the initial bound 2, x beta reaches 1000..2000 and exp overflows; with the
    bound divided by 10 the fit descends from x beta ~ 200 and converges."""
    from .data import SPARSE, build_problem
    rng = np.random.default_rng(seed)
    big = np.flatnonzero(rng.random(K) < 0.01)
    xv = rng.uniform(500.0, 1000.0, len(big))
    cols = [(SPARSE, big, xv)]
    eta = np.zeros(K)
    for _ in range(3):
        r = np.flatnonzero(rng.random(K) < 0.05)
        cols.append((INDICATOR, r, None))
        eta[r] += 0.3
    eta[big] += slope * xv
    y = rng.poisson(np.exp(eta)).astype(np.float64)
    return build_problem(POISSON, cols, y)


def _request_design(args):
    m, K, P, per_row, seed = args
    n_true = min(100, max(1, P // 10))
    if m == LOGISTIC:
        return lr_zipf(K=K, P=P, per_row=per_row, n_true=n_true, seed=seed)[0]
    if m == POISSON:
        return poisson_zipf(K=K, P=P, per_row=per_row, n_true=n_true, base=-5.0, seed=seed)[0]
    return cox_zipf(K=K, P=P, per_row=per_row, n_true=n_true, seed=seed)[0]


def request_designs(n_requests=100, rate=50.0, K_range=(10_000, 2_000_000),
                    P_range=(1_000, 100_000), per_row_range=(20, 60),
                    models=(LOGISTIC, POISSON, COX), seed=DEFAULT_SEED, workers=1,
                    progress=None):
    """A served request stream with a design of its own per request: Poisson
    arrivals at `rate` per second, model uniform over `models`, rows and columns
    log-uniform in the ranges, nonzeros per row uniform, Zipf(1.1) columns, a prior
    variance log-uniform in [0.01, 1].  `workers` > 1 generates the designs in a
    pool of forked processes (call it before the process has touched the GPU).
    Returns [(arrival_s, problem, variance)]."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    jobs, var = [], []
    for _ in range(n_requests):
        m = models[int(rng.integers(len(models)))]
        K = int(round(np.exp(rng.uniform(np.log(K_range[0]), np.log(K_range[1])))))
        P = int(round(np.exp(rng.uniform(np.log(P_range[0]), np.log(P_range[1])))))
        per_row = int(rng.integers(per_row_range[0], per_row_range[1] + 1))
        var.append(float(np.exp(rng.uniform(np.log(0.01), np.log(1.0)))))
        jobs.append((m, K, P, per_row, int(rng.integers(1 << 30))))
    if workers > 1:
        import multiprocessing as mp
        with mp.get_context("fork").Pool(workers) as pool:
            designs = []
            for prob in pool.imap(_request_design, jobs):
                designs.append(prob)
                if progress is not None:
                    progress(prob)
    else:
        designs = []
        for j in jobs:
            designs.append(_request_design(j))
            if progress is not None:
                progress(designs[-1])
    return [(float(t[i]), designs[i], var[i]) for i in range(n_requests)]
