"""The C-ABI library loads and exports every entry point include/lekf.h declares (CPU)."""
import os
import re

import pytest

from letkf_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(lekf_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_loads():
    L = _lib.lib()
    assert L.lekf_version().decode().startswith("lekf-mi355x")


def test_every_declared_symbol_is_exported():
    L = _lib.lib()
    names = _declared()
    assert "lekf_core_f64" in names and "lekf_das_run" in names
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_binding_list_matches_header():
    assert set(_lib.EXPORTED) >= _declared() - {"lekf_last_error"}


def test_fortran_shim_built():
    path = os.path.join(ROOT, "letkf_amd", "fortran", "liblekf_fortran.so")
    if not os.path.exists(path):
        pytest.skip("amdflang not available")
    import ctypes
    ctypes.CDLL(_lib.LIB_PATH)
    ctypes.CDLL(path)
