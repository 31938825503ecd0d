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
