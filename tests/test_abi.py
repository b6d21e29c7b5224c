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


def test_server_without_device_fails_cleanly():
    if not os.path.exists(engine.LIB_PATH):
        pytest.skip("libibx.so not built")
    try:
        srv = engine.Server()
    except engine.IbxError as e:
        assert e.code == -4 and "device" in str(e)
        return
    srv.close()
    pytest.skip("a HIP device is present: the serving loop is covered by the GPU tests")


def test_footprint_estimate_without_device():
    """ibx_estimate_bytes needs no device: the list dominates (cap ~ 1.5 x the neighbours within rlist)."""
    if not os.path.exists(engine.LIB_PATH):
        pytest.skip("libibx.so not built")
    from ibisco_amd import system
    s = system.cg_melt(nchains=20, nbeads=50, seed=4, avoid=0.8, dt=0.002)
    b0, b1 = engine.estimate_bytes(s), engine.estimate_bytes(s, 150)
    assert b1 - b0 == 150 * 8 + 4 + 8   # the histogram, its sample count and its sum of 1/volume
    import math
    nbr = 4.0 / 3.0 * math.pi * 1.8 ** 3 * 8.0
    cap = (int(1.5 * nbr) + 64 + 7) // 8 * 8
    assert 1024 * cap * 4 < b0 < 1024 * (cap * 4 + 600)
