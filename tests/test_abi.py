"""The C-ABI library loads and exports every symbol include/ccdhip.h declares (CPU only:
no compute call is made without a GPU)."""
import ctypes
import os
import re

import pytest

from ccd_amd import api

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "ccdhip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ccdh_\w+)\s*\(", text)))


def test_header_declares_the_python_export_list():
    assert header_symbols() == sorted(api.EXPORTS)


@pytest.mark.skipif(not os.path.exists(api.LIB_PATH), reason="libccdhip.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(api.LIB_PATH)
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(api.LIB_PATH), reason="libccdhip.so not built")
def test_invalid_arguments_are_rejected_without_a_device():
    lib = api.load_library()
    assert lib.ccdh_data_create(None, None) == -1
    assert b"null" in lib.ccdh_last_error()
    assert lib.ccdh_fit_plain(None, None, None, None) == -1


@pytest.mark.skipif(not os.path.exists(api.LIB_PATH), reason="libccdhip.so not built")
def test_server_arguments_are_rejected_without_a_device():
    lib = api.load_library()
    assert lib.ccdh_server_create(None, None) == -1
    opt = api.ServerOptions(7, 0, 0, 0, 0.0, 0)  # unknown precision
    h = ctypes.c_void_p()
    assert lib.ccdh_server_create(ctypes.byref(opt), ctypes.byref(h)) == -1
    assert b"precision" in lib.ccdh_last_error()
    assert lib.ccdh_server_submit(None, None, None, None) == -1
    assert lib.ccdh_server_wait(None, 1, -1.0, None, None, None) == -1
    assert lib.ccdh_server_destroy(None) == -1


def test_serving_arena_allocator(tmp_path):
    """First-fit arena of the server (header-only C++), built with g++ and run."""
    import subprocess
    exe = tmp_path / "arena_check"
    src = os.path.join(ROOT, "tests", "native", "arena_check.cpp")
    subprocess.run(["g++", "-O1", "-std=c++17", "-o", str(exe), src], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout


def test_bootstrap_weights_match_the_selector_restatement():
    """Host logic of the C ABI (no device): bootstrap weights equal the oracle's
    multiset restatement of BootstrapSelector for several seeds and replicates,
    with and without excluded pids, and are multinomial resampling counts."""
    import numpy as np
    from oracle import oracle as orc
    rng = np.random.default_rng(3)
    sizes = rng.integers(1, 6, 400)  # uneven strata
    ids = np.repeat(np.arange(len(sizes)), sizes).astype(np.int32)
    N = len(sizes)
    exclude = (rng.random(N) < 0.2).astype(float)
    for seed in (1, 123, 99991):
        prev = None
        for r in range(4):
            for ex in (None, exclude):
                w = api.bootstrap_weights(ids, seed, r, exclude=ex)
                assert np.array_equal(w, orc.bootstrap_weights(ids, seed, r, exclude=ex))
                per_pid = np.bincount(ids, weights=w, minlength=N) / sizes
                assert np.array_equal(per_pid, np.round(per_pid)) and per_pid.min() >= 0
                drawn = N if ex is None else int((ex == 0).sum())
                assert per_pid.sum() == drawn  # one draw per included pid
                if ex is not None:
                    assert np.all(per_pid[ex != 0] == 0)
                assert per_pid.max() > 1  # drawn with replacement
            if prev is not None:
                assert not np.array_equal(w, prev)
            prev = w
    with pytest.raises(api.CcdhError):
        api.bootstrap_weights(np.array([-1], np.int32), 1, 0)


# UniModalSearch's own unit example (utils/HParSearch.cpp, UNITTEST_ block): ten
# (variance, log likelihood) points with R's quadratic fit in log variance and
# the search's next step.
_SEARCH_KAT = [(2.000000e+04, -408.199), (2.002884e+03, -328.581), (2.000000e+02, -281.788),
               (2.002884e+01, -229.966), (2.000000e+00, -184.908), (2.002884e-01, -150.792),
               (2.000000e-02, -132.081), (2.002884e-03, -127.708), (2.000000e-04, -146.397),
               (2.002884e-05, -183.208)]


def test_search_matches_the_reference_example_and_the_oracle():
    """Host logic of the C ABI: the unimodal search reproduces the reference's
    unit example (R: c0 -171.030466, log coefficient -12.662184, log^2 coefficient
    -1.186039, next step 0.004805411 with the default step 100), and agrees with
    the oracle's restatement (Cramer's rule instead of pivoted elimination) on
    every branch."""
    import numpy as np
    from oracle import oracle as orc
    tried = [(x, y, 0.0) for x, y in _SEARCH_KAT]
    c = api.search_quadratic(tried)
    np.testing.assert_allclose(c, [-171.030466, -12.662184, -1.186039], rtol=2e-6)
    cont, nxt, _ = api.search_step(tried, stdstep=100.0)
    np.testing.assert_allclose(nxt, 0.004805411, rtol=2e-6)
    assert not cont  # the fitted maximum (-137.2) is below the best point tried (-127.7)
    rng = np.random.default_rng(8)
    cases = [[], [(0.5, -10.0, 1.0)], [(2.0, -10.0, 1.0)], [(1.0, -9.0, 0.1), (10.0, -8.0, 0.2)],
             [(1.0, -8.0, 0.1), (10.0, -9.0, 0.2)]]
    for _ in range(40):
        n = int(rng.integers(3, 9))
        xs = np.exp(rng.uniform(-8, 2, n))
        ys = -(np.log(xs) + 3.0) ** 2 + rng.normal(0, 0.3, n) - 50
        sds = np.where(rng.random(n) < 0.2, 0.0, rng.uniform(0.1, 2.0, n))
        cases.append(list(zip(xs, ys, sds)))
    cases.append([(1.0, -5.0, 0.1), (0.1, -6.0, 0.1), (0.01, -7.0, 0.1)])  # best at the right end
    cases.append([(1.0, -7.0, 0.1), (0.1, -6.0, 0.1), (0.01, -5.0, 0.1)])  # best at the left end
    cases.append([(1.0, -5.0, 0.1), (1.0, -7.0, 0.1), (0.1, -6.0, 0.0), (0.01, -8.0, 0.0)])
    for t in cases:
        a, o = api.search_step(t), orc.search_step(t)
        assert a[0] == o[0]
        # (normal equations in log x reach (log x)^4: the two solvers agree to ~1e-9)
        np.testing.assert_allclose(a[1:], o[1:], rtol=1e-7, atol=1e-12)
        if len({x for x, _, _ in t}) >= 3:
            np.testing.assert_allclose(api.search_quadratic(t), orc.search_quadratic(t),
                                       rtol=1e-7, atol=1e-9)
