"""ccdh_cox_groups_noalloc (the allocation-free variant of ccdh_cox_groups: its
staging lives in a file-scope device array) gives exactly what ccdh_cox_groups
gives, on two calls in a row with the second over a smaller data set, under
weights and without, on FP64 and mixed data sets; a data set above the array's
row limit is refused."""
import numpy as np
import pytest

from ccd_amd import api, synth
from ccd_amd.data import COX, build_problem

pytestmark = pytest.mark.gpu

MAX_ROWS = 1 << 20


def _same(a, b):
    for x, y in zip(a[:4], b[:4]):
        assert np.array_equal(x, y)
    assert a[4] == b[4]


@pytest.mark.parametrize("precision", [api.FP64, api.MIXED])
def test_noalloc_cox_groups_equal_the_allocating_call(precision):
    big, _ = synth.cox_zipf(K=200_000, P=400, per_row=5, n_true=40, seed=31)
    small, _ = synth.cox_zipf(K=3_000, P=50, per_row=4, n_true=5, seed=32)
    rng = np.random.default_rng(4)
    for p in (big, small):  # the second call is smaller: stale contents must not leak
        print("noalloc cox groups K", p.K, "precision", precision, flush=True)
        d = api.Data(p, precision=precision)
        w = (rng.random(p.K) < 0.8).astype(float)
        for weights in (None, w):
            want = api.cox_groups(d, weights)
            got = api.cox_groups(d, weights, noalloc=True)
            _same(got, want)
            assert got[4] > 0 and (weights is None or np.all(got[0][w == 0] == got[4]))
        d.close()


def test_noalloc_cox_groups_refuse_more_rows_than_the_array_holds():
    K = MAX_ROWS + 1
    t = np.arange(K, 0, -1, dtype=float)
    p = build_problem(COX, [(0, None, np.zeros(K))], (np.arange(K) % 3 == 0).astype(float),
                      offs=t)
    d = api.Data(p)
    with pytest.raises(api.CcdhError, match="at most 1048576"):
        api.cox_groups(d, noalloc=True)
    d.close()
