"""Write the golden fixtures under tests/golden/ (run in the build container only).

Each fixture is a small .npz holding a problem (the CSC arrays of
ccd_amd.data.CcdProblem) and the maximum-likelihood estimate computed here by
an independent full-vector Newton-Raphson solver written directly from the
likelihoods (dense X, closed-form score and information matrix, FP64,
iterated until the step is below 1e-13).  None of this code shares a line with
the coordinate-wise oracle, so a dropped term or a wrong index in either shows
up as a mismatch.

Datasets are the ones the reference's own tests hold:
  * binomial "bid" table ............ tests/testthat/test-smallBernoulli.R:10-15
  * Dobson counts ................... tests/testthat/test-smallPoisson.R:8-12
  * four 7-row Cox tables ........... tests/testthat/test-smallCox.R:7-16,50-59,73-81,94-103
  * weighted, duplicated Cox table .. tests/testthat/test-weighting.R:107-142
  * infert (CLR) .................... inst/extdata/infert_ccd.txt (gold: clogit, test-smallCLR.R)
  * 17-variable logistic file ....... inst/extdata/CCD_LOGISTIC_TEST_17var.txt
                                      (SAS: 2 logL = -1578.046, test-largeBernoulli.R:20)

Usage:  python tests/golden/make_golden.py   (reads /root/reference for the two text files)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))
from ccd_amd.data import (COX, CONDITIONAL_LOGISTIC, DENSE, INDICATOR, LOGISTIC,  # noqa: E402
                          POISSON, SPARSE, build_problem, sort_cox)

REF = "/root/reference"


# ---------------------------------------------------------------- likelihoods
def lr_terms(X, y, w, beta):
    eta = X @ beta
    p = 1.0 / (1.0 + np.exp(-eta))
    ll = np.sum(w * (y * eta - np.logaddexp(0.0, eta)))
    g = X.T @ (w * (y - p))
    H = -(X * (w * p * (1 - p))[:, None]).T @ X
    return ll, g, H


def poisson_terms(X, y, w, beta, off):
    eta = X @ beta + off
    mu = np.exp(eta)
    from scipy.special import gammaln
    ll = np.sum(w * (y * eta - mu - gammaln(y + 1)))
    g = X.T @ (w * (y - mu))
    H = -(X * (w * mu)[:, None]).T @ X
    return ll, g, H


def clr_terms(X, y, strata, beta):
    eta = X @ beta
    ll, g, H = 0.0, np.zeros(X.shape[1]), np.zeros((X.shape[1],) * 2)
    for s in np.unique(strata):
        m = strata == s
        e = np.exp(eta[m])
        n = y[m].sum()
        p = e / e.sum()
        xs = X[m]
        xbar = p @ xs
        ll += y[m] @ eta[m] - n * np.log(e.sum())
        g += y[m] @ xs - n * xbar
        H -= n * ((xs * p[:, None]).T @ xs - np.outer(xbar, xbar))
    return ll, g, H


def cox_breslow_terms(X, time, event, strata, w, beta):
    """Breslow partial likelihood by explicit risk sets (O(K^2))."""
    eta = X @ beta
    r = w * np.exp(eta)
    ll, g, H = 0.0, np.zeros(X.shape[1]), np.zeros((X.shape[1],) * 2)
    for i in np.nonzero(event == 1)[0]:
        if w[i] == 0:
            continue
        risk = (time >= time[i]) & (strata == strata[i])
        S = r[risk].sum()
        xr = X[risk]
        xbar = r[risk] @ xr / S
        ll += w[i] * (eta[i] - np.log(S))
        g += w[i] * (X[i] - xbar)
        H -= w[i] * ((xr * r[risk][:, None]).T @ xr / S - np.outer(xbar, xbar))
    return ll, g, H


def newton(fun, p, tol=1e-13, maxit=200):
    beta = np.zeros(p)
    for _ in range(maxit):
        ll, g, H = fun(beta)
        step = np.linalg.solve(H, -g)
        # damped until the likelihood does not decrease
        t = 1.0
        while fun(beta + t * step)[0] < ll - 1e-12 and t > 1e-8:
            t *= 0.5
        beta = beta + t * step
        if np.max(np.abs(t * step)) < tol:
            break
    ll, g, H = fun(beta)
    return beta, ll


# ---------------------------------------------------------------- datasets
def ds_binomial_bid():
    bid = np.array([1, 5, 10, 20, 30, 40, 50, 75, 100, 150, 200.0])
    n = np.array([31, 29, 27, 25, 23, 21, 19, 17, 15, 15, 15])
    yy = np.array([0, 3, 6, 7, 9, 13, 17, 12, 11, 14, 13])
    log_bid = np.log(np.concatenate([np.repeat(bid, n - yy), np.repeat(bid, yy)]))
    y = np.concatenate([np.zeros((n - yy).sum()), np.ones(yy.sum())])
    p = build_problem(LOGISTIC, [(DENSE, None, log_bid)], y, intercept=True)
    X = np.column_stack([np.ones(len(y)), log_bid])
    beta, ll = newton(lambda b: lr_terms(X, y, np.ones(len(y)), b), 2)
    return p, beta, ll


def ds_dobson():
    counts = np.array([18, 17, 15, 20, 10, 20, 25, 13, 12.0])
    outcome = np.tile([1, 2, 3], 3)
    treatment = np.repeat([1, 2, 3], 3)
    cols, Xc = [], [np.ones(9)]
    for lev in (2, 3):
        r = np.nonzero(outcome == lev)[0]
        cols.append((INDICATOR, r, None))
        Xc.append((outcome == lev).astype(float))
    for lev in (2, 3):
        r = np.nonzero(treatment == lev)[0]
        cols.append((INDICATOR, r, None))
        Xc.append((treatment == lev).astype(float))
    p = build_problem(POISSON, cols, counts, intercept=True)
    X = np.column_stack(Xc)
    beta, ll = newton(lambda b: poisson_terms(X, counts, np.ones(9), b, np.zeros(9)), 5)
    return p, beta, ll


def ds_poisson_offset():
    """Dobson counts with a log-exposure offset (exercises the fixed offset column)."""
    counts = np.array([18, 17, 15, 20, 10, 20, 25, 13, 12.0])
    exposure = np.array([10, 12, 9, 14, 7, 15, 20, 11, 9.0])
    off = np.log(exposure)
    outcome = np.tile([1, 2, 3], 3)
    x = (outcome == 2).astype(float)
    z = np.linspace(-1, 1, 9)
    cols = [(INDICATOR, np.nonzero(x)[0], None), (DENSE, None, z)]
    p = build_problem(POISSON, cols, counts, offset=off, intercept=True)
    X = np.column_stack([np.ones(9), x, z])
    beta, ll = newton(lambda b: poisson_terms(X, counts, np.ones(9), b, off), 3)
    return p, np.concatenate([[1.0], beta]), ll


COX_TABLES = {
    "cox_noties": ([4, 3.5, 3, 2.5, 2, 1.5, 1], [1, 1, 0, 1, 1, 0, 1],
                   [0, 2, 0, 0, 1, 1, 1], [0, 0, 1, 1, 1, 0, 0]),
    "cox_timeties": ([4, 3, 3, 2, 2, 1, 1], [1, 1, 0, 1, 0, 0, 1],
                     [0, 2, 0, 0, 1, 1, 1], [0, 0, 1, 1, 1, 0, 0]),
    "cox_failties": ([4, 3, 3, 2, 2, 2], [1, 1, 0, 1, 1, 1],
                     [0, 2, 0, 0, 1, 0], [0, 0, 1, 1, 1, 0]),
    "cox_failties_after": ([4, 3, 3, 2, 2, 1, 1], [1, 1, 0, 1, 1, 0, 1],
                           [0, 2, 0, 0, 1, 1, 1], [0, 0, 1, 1, 1, 0, 0]),
}


def ds_cox(name, strata_from_x2=False, sparse=False, dup_weights=None):
    time, event, x1, x2 = (np.array(a, dtype=float) for a in COX_TABLES[name])
    w = np.ones(len(time))
    if dup_weights is not None:
        idx = np.concatenate([np.arange(len(time)), dup_weights[0]])
        time, event, x1, x2 = time[idx], event[idx], x1[idx], x2[idx]
        w = np.asarray(dup_weights[1], dtype=float)
    strata = x2.astype(np.int32) if strata_from_x2 else np.zeros(len(time), np.int32)
    order = sort_cox(time, event, strata)
    time, event, x1, x2, w, strata = (a[order] for a in (time, event, x1, x2, w, strata))
    fmt = SPARSE if sparse else DENSE
    cols = []
    for xx in ([x1] if strata_from_x2 else [x1, x2]):
        if fmt == SPARSE:
            r = np.nonzero(xx)[0]
            cols.append((SPARSE, r, xx[r]))
        else:
            cols.append((DENSE, None, xx))
    p = build_problem(COX, cols, event, offs=time, strata=strata)
    X = np.column_stack([x1] if strata_from_x2 else [x1, x2])
    beta, ll = newton(lambda b: cox_breslow_terms(X, time, event, strata, w, b), X.shape[1])
    return p, beta, ll, (w if dup_weights is not None else None)


def ds_infert():
    rows = [ln.split() for ln in open(os.path.join(REF, "inst/extdata/infert_ccd.txt"))
            if not ln.startswith("#") and ln.strip()]
    strata = np.array([int(r[0]) for r in rows])
    y = np.array([float(r[1]) for r in rows])
    K = len(rows)
    X = np.zeros((K, 2))
    for k, r in enumerate(rows):
        for tok in r[2:]:
            c, v = tok.split(":")
            X[k, int(c) - 1] = float(v)
    strata = strata - strata.min()
    order = np.argsort(strata, kind="stable")
    strata, y, X = strata[order], y[order], X[order]
    cols = []
    for c in range(2):
        r = np.nonzero(X[:, c])[0]
        cols.append((SPARSE, r, X[r, c]))
    p = build_problem(CONDITIONAL_LOGISTIC, cols, y, strata=strata.astype(np.int32))
    beta, ll = newton(lambda b: clr_terms(X, y, strata, b), 2)
    return p, beta, ll


def ds_sas17():
    path = os.path.join(REF, "inst/extdata/CCD_LOGISTIC_TEST_17var.txt")
    y, entries = [], []
    for ln in open(path):
        if ln.startswith("#") or not ln.strip():
            continue
        t = ln.split()
        y.append(float(t[1]))
        entries.append([(int(a), float(b)) for a, b in (s.split(":") for s in t[2:])])
    y = np.array(y)
    K = len(y)
    names = sorted({c for e in entries for c, _ in e})
    cidx = {c: i for i, c in enumerate(names)}
    X = np.zeros((K, len(names)))
    for k, e in enumerate(entries):
        for c, v in e:
            X[k, cidx[c]] = v
    cols = []
    for i in range(len(names)):
        r = np.nonzero(X[:, i])[0]
        if np.all(X[r, i] == 1.0):
            cols.append((INDICATOR, r, None))
        else:
            cols.append((SPARSE, r, X[r, i]))
    p = build_problem(LOGISTIC, cols, y, intercept=True)
    Xf = np.column_stack([np.ones(K), X])
    beta, ll = newton(lambda b: lr_terms(Xf, y, np.ones(K), b), Xf.shape[1])
    return p, beta, ll


def save(name, p, beta, ll, weights=None, extra=None):
    d = dict(model=p.model, K=p.K, J=p.J, col_ptr=p.col_ptr, rows=p.rows, vals=p.vals,
             fmt=p.fmt, y=p.y, offs=p.offs, strata=p.strata, has_offset=int(p.has_offset),
             beta=np.asarray(beta, float), loglik=float(ll))
    if weights is not None:
        d["weights"] = np.asarray(weights, float)
    if extra:
        d.update(extra)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)
    print(f"{name:28s} K={p.K:6d} J={p.J:3d} ll={ll:.10f} beta={np.round(beta, 6)}")


def main():
    save("lr_binomial_bid", *ds_binomial_bid())
    save("poisson_dobson", *ds_dobson())
    save("poisson_offset", *ds_poisson_offset())
    for name in COX_TABLES:
        p, b, ll, _ = ds_cox(name)
        save(name, p, b, ll)
    p, b, ll, _ = ds_cox("cox_noties", strata_from_x2=True)
    save("cox_noties_strata", p, b, ll)
    p, b, ll, _ = ds_cox("cox_failties_after", strata_from_x2=True, sparse=True)
    save("cox_failties_strata_sparse", p, b, ll)
    wdup = ([0, 2, 3], [0.5, 1, 0.5, 0.5, 1, 1, 1, 0.5, 0.5, 0.5])
    p, b, ll, w = ds_cox("cox_noties", dup_weights=wdup)
    save("cox_weighted_dup", p, b, ll, weights=w)
    save("clr_infert", *ds_infert())
    p, b, ll = ds_sas17()
    save("lr_sas17", p, b, ll, extra=dict(sas_minus2ll=-1578.046))


if __name__ == "__main__":
    main()
