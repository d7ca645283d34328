"""The C-ABI library loads and exports every symbol include/profit_amd.h declares."""
import ctypes
import os
import re

import pytest

import profit_amd as P

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "profit_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(profit_amd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "profit_amd_make_model" in names
    assert "profit_amd_engine_like_grad" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    if not os.path.exists(P.LIB_PATH):
        pytest.skip("libprofit_amd.so not built")
    lib = ctypes.CDLL(P.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    bound = {s[0] for s in P.SYMBOLS}
    assert set(declared_functions()) == bound


def test_struct_layouts():
    # profit_amd_profile mirrors the engine's ProfileParams (static_assert in capi.cpp)
    assert ctypes.sizeof(P.Profile) == 4 * 6 + 8 * 9 + 4 * 2 + 8 * 5


def test_no_gpu_calls_without_device():
    # library metadata is available without touching the GPU
    if not os.path.exists(P.LIB_PATH):
        pytest.skip("libprofit_amd.so not built")
    assert b"brute" in P.lib().profit_amd_convolvers()
    p = P.make_profile("sersic", re=3.0)
    assert p.shape[0] == 3.0 and p.resolution == 9 and p.max_recursions == 2 and p.adjust == 1
