"""The C-ABI library loads and exports every symbol include/powergrid_amd.h
declares (no compute: this container has no GPU)."""
import ctypes
import os
import subprocess

import pytest

from powergrid_amd import _lib


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _lib.header_symbols()
    assert len(declared) >= 30
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing


def test_status_strings_and_version():
    L = _lib.lib()
    assert L.pg_version().decode().count(".") == 2
    for code in range(7):
        assert L.pg_status_string(code)
    assert L.pg_status_string(0) == b"ok"


def test_argument_errors_are_reported_without_a_device():
    L = _lib.lib()
    h = ctypes.c_void_p()
    # empty trajectory -> invalid argument, no crash, message recorded
    st = L.pg_gnufft_create(ctypes.byref(h), 0, 2.0, 16, 16, 1, None, None, None, None, None,
                            None, 4.0)
    assert st == 1
    assert b"empty" in L.pg_last_error()
    assert L.pg_op_destroy(None) == 0
    assert L.pg_op_set_mode(ctypes.c_void_p(0), 0) == 1


def test_header_compiles_as_c_and_cxx(tmp_path):
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    src = tmp_path / "t.c"
    src.write_text('#include "powergrid_amd.h"\nint main(void){return PG_OK;}\n')
    subprocess.check_call(["gcc", "-std=c99", "-Wall", "-Werror", "-I", inc, "-c", str(src),
                           "-o", str(tmp_path / "t.o")])
    subprocess.check_call(["g++", "-std=c++17", "-Wall", "-Werror", "-I", inc, "-x", "c++",
                           "-c", str(src), "-o", str(tmp_path / "t2.o")])


def test_cpp_api_header_compiles(tmp_path):
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    src = tmp_path / "t.cpp"
    src.write_text("""
#include "powergrid/powergrid.hpp"
using namespace powergrid;
int main() {
    Col<float> kx(4), ky(4), kz(4), ix, iy, iz;
    initImageSpaceCoords(ix, iy, iz, 2, 2, 1);
    return (int)ix.n_rows - 4;
}
""")
    subprocess.check_call(["g++", "-std=c++17", "-Wall", "-Werror", "-I", inc, "-fsyntax-only",
                           str(src)])
