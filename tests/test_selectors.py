"""Cross-validation folds and bootstrap draws (CPU): the library's
ccdh_cv_fold_weights / ccdh_bootstrap_weights and the oracle's restatement both
equal a Python restatement of libstdc++'s std::mt19937, uniform_int_distribution
and std::shuffle (tests/golden/libstdcxx_rng.py) driven as the reference's
selectors drive them; the engine itself is pinned to the value the C++ standard
fixes for it."""
import numpy as np
import pytest

from ccd_amd import api, synth
from oracle import oracle as orc
from tests.golden import libstdcxx_rng as R

pytestmark = pytest.mark.skipif(not __import__("os").path.exists(api.LIB_PATH),
                                reason="libccdhip.so not built")


def test_mt19937_is_the_standard_engine():
    # [rand.predef]: the 10000th consecutive invocation of a default-constructed
    # mt19937 produces 4123659995
    g = R.MT19937()
    v = [g() for _ in range(10000)]
    assert v[-1] == 4123659995
    # seed(5489) is the default state; another seed is another stream
    assert R.MT19937(5489)() == v[0] and R.MT19937(123)() != v[0]


def test_uniform_int_is_unbiased_and_in_range():
    g = R.MT19937(7)
    draws = np.array([R.uniform_int(g, 0, 4) for _ in range(50000)])
    assert draws.min() == 0 and draws.max() == 4
    assert np.abs(np.bincount(draws) / 50000 - 0.2).max() < 0.01
    g = R.MT19937(7)
    assert {R.uniform_int(g, 3, 3) for _ in range(10)} == {3}


@pytest.mark.parametrize("n", [1, 2, 7, 64, 1001, 65535, 65536, 70001])
def test_shuffle_is_a_permutation(n):
    # both branches: pairs of positions per draw (n <= 65535) and one per position
    perm = list(range(n))
    R.shuffle(perm, R.MT19937(11))
    assert sorted(perm) == list(range(n))
    if n > 7:
        assert perm != list(range(n))


def _unit_ids(n_units, seed=3):
    sizes = np.random.default_rng(seed).integers(1, 6, n_units)
    return np.repeat(np.arange(n_units), sizes).astype(np.int32)


@pytest.mark.parametrize("n_units", [9, 10, 401, 70001])
@pytest.mark.parametrize("by_pid", [True, False])
def test_cv_fold_weights_equal_the_restated_selector(n_units, by_pid):
    ids = _unit_ids(n_units)
    n_fold = 10 if n_units >= 10 else 3
    tasks = range(2 * n_fold) if n_units < 1000 else (0, n_fold - 1, n_fold, 2 * n_fold - 1)
    for seed in (123, 2024, -1, 7):
        for task in tasks:
            want = R.cv_fold_weights(ids, n_fold, seed, task, by_pid=by_pid)
            got = api.cv_fold_weights(ids, n_fold, seed, task, by_pid=by_pid)
            assert np.array_equal(got, want), (seed, task)
            assert np.array_equal(orc.cv_weights(ids, n_fold, seed, task, by_pid=by_pid), want)


def test_cv_folds_partition_the_units():
    ids = _unit_ids(1003)
    for seed in (123, -1):
        held = np.zeros(len(ids))
        for fold in range(10):
            w = api.cv_fold_weights(ids, 10, seed, fold)
            held += 1.0 - w
            # whole strata are held out, 100 or 101 of them
            per = np.bincount(ids, weights=1.0 - w) / np.bincount(ids)
            assert set(np.unique(per)) <= {0.0, 1.0} and per.sum() in (100, 101)
        assert np.array_equal(held, np.ones(len(ids)))
    # seed -1: no shuffle, fold f holds out the f-th interval of units
    w = api.cv_fold_weights(ids, 10, -1, 0)
    assert np.array_equal(w == 0, ids < 101)


def test_cv_fold_weights_of_the_batched_cv_workload():
    """The 160 fits of the batched-CV benchmark use folds over 200,000 strata
    (one shuffle per repetition over 200,000 units): first and last fold."""
    # the stratum sizes synth.clr_strata draws first (1 case + U{1..10} controls)
    rng = np.random.default_rng(synth.DEFAULT_SEED)
    sizes = 1 + rng.integers(1, 11, 200_000)
    strata = np.repeat(np.arange(200_000), sizes).astype(np.int32)
    for task in (0, 9):
        want = R.cv_fold_weights(strata, 10, 123, task)
        assert np.array_equal(api.cv_fold_weights(strata, 10, 123, task), want)


@pytest.mark.parametrize("seed", [1, 123, -1])
def test_bootstrap_weights_equal_the_restated_selector(seed):
    ids = _unit_ids(300)
    exclude = (np.random.default_rng(5).random(300) < 0.2).astype(float)
    for r in range(3):
        for ex in (None, exclude):
            want = R.bootstrap_weights(ids, seed, r, exclude=ex)
            assert np.array_equal(api.bootstrap_weights(ids, seed, r, exclude=ex), want)
            assert np.array_equal(orc.bootstrap_weights(ids, seed, r, exclude=ex), want)


def test_clock_seed_is_refused():
    ids = _unit_ids(20)
    with pytest.raises(api.CcdhError, match="-99"):
        api.cv_fold_weights(ids, 10, -99, 0)
    with pytest.raises(api.CcdhError, match="-99"):
        api.bootstrap_weights(ids, -99, 0)
