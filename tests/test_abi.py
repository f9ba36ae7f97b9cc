"""The C-ABI library loads and exports every entry point include/lpm_gpu.h
declares; host-only entry points (unit cell, argument checks) behave like the
reference.  Nothing here launches a GPU kernel."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lpm_gpu.h")


@pytest.fixture(scope="module")
def lib():
    import lpmg

    if not os.path.exists(lpmg.LIB_PATH):
        subprocess.check_call(["make", "-j8", "-C", lpmg.CSRC])
    return lpmg.lib()


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lpmg_\w+)\s*\(", text)))


def test_header_declares_the_binding_list():
    import lpmg

    assert declared_functions() == sorted(lpmg.EXPORTS)


def test_library_exports_every_declared_symbol(lib):
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.check_output(["nm", "-D", "--defined-only", os.path.join(ROOT, "lpm-cpp_amd", "liblpmg.so")],
                                  text=True)
    exported = set(re.findall(r"\bT (lpmg_\w+)", out))
    assert set(declared_functions()) <= exported


def test_header_is_plain_c():
    subprocess.check_call(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-x", "c", HEADER])


def test_unit_cell_matches_oracle(lib):
    import lpmg
    from oracle import unit_cell

    for L in range(5):
        np.testing.assert_array_equal(lpmg.unit_cell(L, 0.13), unit_cell(L, 0.13))


def test_bad_arguments_are_errors_not_exits(lib):
    h = ctypes.c_void_p()
    xyz = np.zeros(3)
    dp = xyz.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert lib.lpmg_create(99, 0.1, 0, 1, dp, ctypes.byref(h)) == -1
    assert b"lattice" in lib.lpmg_last_error()
    assert lib.lpmg_create(2, -1.0, 0, 1, dp, ctypes.byref(h)) == -1
    assert lib.lpmg_create(2, 0.1, 7, 1, dp, ctypes.byref(h)) == -1
    assert lib.lpmg_forward(None) == -1
    assert b"null handle" in lib.lpmg_last_error()
    out = np.zeros(2)
    assert lib.lpmg_update_rr(None, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == -1
