"""Restatement, in Python, of the libstdc++ random routines that the reference's
cross-validation and bootstrap selectors call, written from the installed GCC 11
headers' source (the same headers g++ and hipcc compile the oracle and the
library against):

* std::mt19937 ([rand.predef]; seed(): bits/random.tcc mersenne_twister_engine::seed)
  -- numpy's MT19937 with its legacy init_genrand seeding, which is the same
  recurrence x[i] = 1812433253 * (x[i-1] ^ (x[i-1] >> 30)) + i mod 2^32 and the
  same tempering;
* std::uniform_int_distribution<T>::operator() (bits/uniform_int_dist.h:273-330):
  a 32-bit engine downscaled to [0, range) by Lemire's nearly divisionless method
  (_S_nd, :243-268) with a 64-bit product;
* std::shuffle (bits/stl_algo.h:3729-3792): when (2^32 - 1) / n >= n it takes
  two swap positions from one draw over [0, b0 * b1) (__gen_two_uniform_ints,
  :3704-3712) after one up-front draw in {0, 1} for an even n; otherwise one draw
  over [0, i] per position i.

and, over them, the selectors as the reference drives them:
CrossValidationSelector (drivers/CrossValidationSelector.cpp:21-116, 138-143;
AbstractSelector.cpp:32-48 for seed -1) with the permute-per-repetition loop of
AbstractCrossValidationDriver.cpp:157-168, and BootstrapSelector
(drivers/BootstrapSelector.cpp:17-95) with BootstrapDriver::drive
(drivers/BootstrapDriver.cpp:47-71).  Pure Python loops: for the small and
medium sizes the tests use.
"""
import numpy as np

M32 = 0xFFFFFFFF
DEFAULT_SEED = 5489  # mersenne_twister_engine::default_seed


class MT19937:
    def __init__(self, seed=DEFAULT_SEED):
        self._bg = np.random.MT19937()
        self._bg._legacy_seeding(int(seed) & M32)
        self._buf = []
        self._pos = 0

    def __call__(self):
        if self._pos == len(self._buf):
            self._buf = [int(v) for v in self._bg.random_raw(4096)]
            self._pos = 0
        v = self._buf[self._pos]
        self._pos += 1
        return v


def uniform_int(g, a, b):
    """uniform_int_distribution{a, b}(g) for a 32-bit engine and b - a < 2^32 - 1."""
    rng = (b - a) + 1
    product = g() * rng
    low = product & M32
    if low < rng:
        threshold = ((1 << 32) - rng) % rng
        while low < threshold:
            product = g() * rng
            low = product & M32
    return a + (product >> 32)


def shuffle(seq, g):
    n = len(seq)
    if n == 0:
        return
    if M32 // n >= n:
        i = 1
        if n % 2 == 0:
            j = uniform_int(g, 0, 1)
            seq[i], seq[j] = seq[j], seq[i]
            i += 1
        while i != n:
            b0 = i + 1
            b1 = b0 + 1
            x = uniform_int(g, 0, b0 * b1 - 1)
            p1, p2 = x // b1, x % b1
            seq[i], seq[p1] = seq[p1], seq[i]
            i += 1
            seq[i], seq[p2] = seq[p2], seq[i]
            i += 1
        return
    for i in range(1, n):
        j = uniform_int(g, 0, i)
        seq[i], seq[j] = seq[j], seq[i]


def cv_fold_weights(ids, n_fold, seed, task, by_pid=True):
    ids = np.asarray(ids)
    K = len(ids)
    N = int(ids.max()) + 1 if by_pid else K
    start, index = [], 0
    fraction, extra = N // n_fold, N % n_fold
    for i in range(n_fold):
        start.append(index)
        index += fraction + (1 if i < extra else 0)
    start.append(N)
    g = MT19937(seed)  # reseed() seeds the engine whatever the seed
    perm = list(range(N))
    for i in range(task + 1):
        if i % n_fold == 0 and seed != -1:
            shuffle(perm, g)
    fold = task % n_fold
    held = perm[start[fold]:start[fold + 1]]
    w = np.ones(K)
    if by_pid:
        w[np.isin(ids, held)] = 0.0
    else:
        w[held] = 0.0
    return w


def bootstrap_weights(ids, seed, replicate, exclude=None):
    ids = np.asarray(ids)
    N = int(ids.max()) + 1
    if exclude is None:
        included = list(range(N))
    else:
        included = [i for i, e in enumerate(exclude) if e == 0]
    g = MT19937(DEFAULT_SEED if seed == -1 else seed)
    counts = None
    for _ in range(replicate + 2):  # the constructor's sample, then one per replicate
        counts = np.zeros(max(N, 0 if exclude is None else len(exclude)), np.int64)
        for _ in range(len(included)):
            counts[included[uniform_int(g, 0, len(included) - 1)]] += 1
    return counts[ids].astype(np.float64)
