"""CPU checks of the C ABI: the library loads and exports every symbol that
include/spx.h declares; host-only argument normalisation matches the
reference facade.  No kernel is launched here."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import spx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(spx.LIB_PATH):
        pytest.fail("libspx.so is not built (run spams-devel_amd/build.py)")
    return spx.lib()


def test_every_declared_symbol_is_exported(lib):
    names = spx.exported_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_string(lib):
    assert b"gfx950" in lib.spx_version()


def test_lasso_L_normalisation_matches_facade():
    # cpp_library/spams.h:120-131
    assert spx.lasso_L(64, 256) == 64
    assert spx.lasso_L(256, 64) == 64
    assert spx.lasso_L(64, 256, L=10) == 10
    assert spx.lasso_L(64, 256, L=100) == 64
    # PENALTY with lambda1 == 0, no pos and lambda2 > 0 keeps L up to p
    assert spx.lasso_L(64, 256, L=100, lambda1=0.0, lambda2=0.1) == 100
    assert spx.lasso_L(64, 256, L=100, lambda1=0.0, lambda2=0.1, pos=True) == 64
    assert spx.lasso_L(64, 256, L=0) == 0


def test_omp_L_normalisation_matches_facade():
    assert spx.omp_L(64, 256) == 64
    assert spx.omp_L(64, 256, 10) == 10
    assert spx.omp_L(16, 256, 32) == 16
    assert spx.omp_L(64, 20, 32) == 20
    assert spx.omp_L(64, 256, np.array([3, 9, 4]), n=3) == 9
    assert spx.omp_L(64, 256, 0) == 0


def test_errors_without_gpu_are_reported_not_crashed(lib):
    """Creating a context on a device that does not exist fails cleanly."""
    h = ctypes.c_void_p()
    rc = lib.spx_ctx_create(ctypes.byref(h), 4096)
    assert rc != 0
    assert len(lib.spx_last_error()) > 0


def test_cpp_dropin_header_compiles():
    """include/spams_amd.h (the spams.h-compatible templates) compiles as
    C++14 against the C ABI header, as a user translation unit would."""
    src = os.path.join(ROOT, "tests", "cpp", "dropin_compile.cpp")
    r = subprocess.run(["g++", "-std=c++14", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
