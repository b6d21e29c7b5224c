"""The C ABI: every symbol declared in include/ibx.h is exported by libibx.so."""
import ctypes
import os
import re

import pytest

from ibisco_amd import engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "ibx.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*\*?(ibx_\w+)\s*\(", txt, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "ibx_create" in names and "ibx_step" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    if not os.path.exists(engine.LIB_PATH):
        pytest.skip("libibx.so not built")
    lib = ctypes.CDLL(engine.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(engine.EXPORTS) == declared()


def test_engine_binding_loads_without_device():
    if not os.path.exists(engine.LIB_PATH):
        pytest.skip("libibx.so not built")
    L = engine.load_library()
    assert L.ibx_last_error(None) is not None
