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


def test_autotune_table_save_load_roundtrip(libpir, tmp_path):
    """The autotune table is host-side: load, count, save and reload a plan file
    (no GPU needed); malformed entries are rejected."""
    for name in ("pir_autotune_entries", "pir_autotune_clear", "pir_autotune_save", "pir_autotune_load"):
        getattr(libpir, name).restype = ctypes.c_int
    libpir.pir_autotune_save.argtypes = [ctypes.c_char_p]
    libpir.pir_autotune_load.argtypes = [ctypes.c_char_p]
    assert libpir.pir_autotune_clear() == 0
    good = tmp_path / "plans.txt"
    good.write_text("4096 256 1152 1152 6 0 1 1 8 8 128 3 3 1 1 1 8 8 : 1 2\n"
                    "65536 64 576 576 6 0 1 1 32 32 64 3 3 1 1 1 32 32 : 0 1\n")
    assert libpir.pir_autotune_load(str(good).encode()) == 0
    assert libpir.pir_autotune_entries() == 2
    out = tmp_path / "saved.txt"
    assert libpir.pir_autotune_save(str(out).encode()) == 0
    assert sorted(out.read_text().splitlines()) == sorted(good.read_text().splitlines())
    bad = tmp_path / "bad.txt"
    bad.write_text("65536 64 576 576 6 0 1 1 32 32 64 3 3 1 1 1 32 32 : 1 1\n")  # 32x128 tile with N = 64
    assert libpir.pir_autotune_load(str(bad).encode()) != 0
    assert libpir.pir_autotune_entries() == 2
    assert libpir.pir_autotune_clear() == 0 and libpir.pir_autotune_entries() == 0
