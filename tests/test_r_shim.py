"""The R binding (r/r_profit_amd.cpp): a syntax-only compile against
include/profit_amd.h (with declarations of the R API it uses, since R is not
installed), so every call into the C-ABI matches its declaration; and its
routine table against the twelve routines ProFit's R code calls
(src/r_profit.cpp:832-850: names and arities), plus the engine's."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "r", "r_profit_amd.cpp")

# src/r_profit.cpp:832-850
REFERENCE_ROUTINES = {
    "R_profit_make_model": 2, "R_profit_convolvers": 0, "R_profit_make_convolver": 7, "R_profit_convolve": 4,
    "R_profit_has_openmp": 0, "R_profit_has_fftw": 0, "R_profit_openclenv_info": 0, "R_profit_openclenv": 3,
    "R_profit_downsample": 2, "R_profit_upsample": 2, "R_profit_adjust_mask": 4, "R_profit_clear_cache": 0,
}
ENGINE_ROUTINES = {"R_profit_amd_engine": 1, "R_profit_amd_add_galaxy": 2, "R_profit_amd_set_batch": 3,
                   "R_profit_amd_like": 2, "R_profit_amd_like_grad": 2}


def test_shim_compiles_against_the_c_abi():
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not available")
    cmd = [gxx, "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
           # R's routine table casts every routine to DL_FUNC
           "-Wno-cast-function-type",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "tests", "native", "r_api"), SHIM]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def _table():
    text = open(SHIM).read()
    return {m.group(1): int(m.group(2))
            for m in re.finditer(r'\{"(R_[a-z_]+)",\s*\(DL_FUNC\)&\1,\s*(\d+)\}', text)}


def test_routine_table_keeps_the_reference_routines():
    table = _table()
    for name, arity in REFERENCE_ROUTINES.items():
        assert table.get(name) == arity, name
    for name, arity in ENGINE_ROUTINES.items():
        assert table.get(name) == arity, name
    assert set(table) == set(REFERENCE_ROUTINES) | set(ENGINE_ROUTINES)


def test_routines_have_the_registered_arity():
    text = open(SHIM).read()
    for name, arity in _table().items():
        m = re.search(r"SEXP\s+" + name + r"\(([^)]*)\)\s*\{", text)
        assert m, name
        args = [a for a in m.group(1).split(",") if a.strip()]
        assert len(args) == arity, name
