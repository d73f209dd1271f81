"""CPU checks of the C-ABI boundary: the HIP library loads and exports every declared symbol."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "peclet_dem.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for must in ("pd_create", "pd_step", "pd_set_positions", "pd_get_positions", "pd_destroy"):
        assert must in syms


def test_library_exports_every_declared_symbol():
    from peclet.dem import library_path

    path = library_path()
    if not os.path.exists(path):
        pytest.skip("HIP library not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, f"missing exports: {missing}"


def test_binding_table_matches_header():
    from peclet.dem._binding import PRODUCT_SIGNATURES, SIGNATURES

    syms = set(declared_symbols())
    for name in list(SIGNATURES) + list(PRODUCT_SIGNATURES):
        assert "pd_" + name in syms, name


def test_oracle_mirrors_product_api():
    from oracle_binding import ORACLE_LIB
    from peclet.dem._binding import SIGNATURES

    if not os.path.exists(ORACLE_LIB):
        pytest.skip("oracle not built")
    lib = ctypes.CDLL(ORACLE_LIB)
    missing = [n for n in SIGNATURES if not hasattr(lib, "pdref_" + n)]
    assert not missing, missing
