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
