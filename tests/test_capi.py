"""The C-ABI library loads and exports every function include/ucy.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ucy.h")
LIB = os.path.join(ROOT, "unicycle_amd", "libucy.so")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(ucy_[a-z0-9_]+)\s*\(", src)
    return sorted(set(names))


def test_header_declares_functions():
    names = declared_functions()
    assert "ucy_odefun" in names and "ucy_integrate" in names and len(names) > 30


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libucy.so is not built (make -C unicycle_amd/csrc)")
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_python_binding_covers_header():
    from unicycle_amd import _lib
    assert set(declared_functions()) == set(_lib.EXPORTED)
