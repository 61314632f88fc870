"""The C-ABI library loads and exports every symbol include/phrike_hip.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "phrike_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ph_[a-z0-9_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    from phrike_amd import _lib
    assert header_symbols() == sorted(_lib.symbols())


def test_library_exports_all_symbols():
    from phrike_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libphrike_hip.so not built (run python -m phrike_amd.build)")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for s in header_symbols():
        assert hasattr(lib, s), s
    lib.ph_abi_version.restype = ctypes.c_int
    assert lib.ph_abi_version() == 1


def test_product_does_not_import_oracle():
    pkg = os.path.join(ROOT, "phrike_amd")
    for dp, _, fs in os.walk(pkg):
        for f in fs:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                src = open(os.path.join(dp, f)).read()
                assert not re.search(r"^\s*(from|import)\s+oracle", src, re.M), f
