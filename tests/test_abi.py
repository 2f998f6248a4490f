"""The C-ABI library loads on a host without a GPU and exports every entry
point that include/pir.h declares (no compute calls are made here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pir.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    names = re.findall(r"\b(pir_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


@pytest.fixture(scope="module")
def libpir():
    from piranha_amd import LIB_PATH
    from piranha_amd import build as B

    if not os.path.exists(LIB_PATH):
        B.build()
    return ctypes.CDLL(LIB_PATH)


def test_header_declares_entry_points():
    names = declared_functions()
    assert len(names) > 40
    for must in ("pir_gemm", "pir_conv_fprop", "pir_rss_matmul", "pir_rss_relu", "pir_rss_maxpool", "pir_net_forward"):
        assert must in names


def test_every_declared_symbol_is_exported(libpir):
    missing = [n for n in declared_functions() if not hasattr(libpir, n)]
    assert not missing, f"declared in include/pir.h but not exported: {missing}"


def test_python_signatures_match_exports(libpir):
    import piranha_amd as P

    declared = set(declared_functions())
    assert set(P._SIGS) <= declared, sorted(set(P._SIGS) - declared)


def test_version_and_error_string(libpir):
    libpir.pir_version.restype = ctypes.c_int
    libpir.pir_last_error.restype = ctypes.c_char_p
    assert libpir.pir_version() >= 1
    assert isinstance(libpir.pir_last_error(), bytes)
