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
