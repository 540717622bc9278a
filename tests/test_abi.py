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


def test_unsupported_and_invalid_shapes_are_rejected_without_a_device():
    # checks that fail before any allocation: status code and message
    L = _lib.lib()
    h = ctypes.c_void_p()
    k = (ctypes.c_float * 8)(*([0.0] * 8))

    def gnufft(M, os_, Nx, Ny, Nz, W):
        return L.pg_gnufft_create(ctypes.byref(h), M, os_, Nx, Ny, Nz, k, k, k, None, None,
                                  None, W)

    # Nx != Ny: the reference's x-fastest padding only works for square images
    assert gnufft(8, 2.0, 16, 32, 1, 4.0) == 2 and b"Nx != Ny" in L.pg_last_error()
    # taps = kernelWidth * gridos: 6, 8 and 12 are supported
    assert gnufft(8, 2.0, 16, 16, 1, 5.0) == 2 and b"taps" in L.pg_last_error()
    # gridos * N must be an even integer
    assert gnufft(8, 1.5, 15, 15, 1, 4.0) == 2
    # null trajectory
    assert L.pg_gnufft_create(ctypes.byref(h), 8, 2.0, 16, 16, 1, None, None, None, None, None,
                              None, 4.0) == 1
    # one MPI rank's share: rank below nranks, tasks for the rank, whole sub-shots
    f = (ctypes.c_float * 64)(*([0.0] * 64))

    def rank(ns, nd, nc, r, nr):
        return L.pg_pcsense_rank_create(ctypes.byref(h), f, f, None, 2, 2, 1, nc, f, 1, 1, f,
                                        f, f, ns, nd, 4.0, r, nr)

    assert rank(2, 4, 2, 3, 3) == 1 and b"rank" in L.pg_last_error()
    assert rank(2, 4, 2, 5, 8) == 1 and b"no tasks" in L.pg_last_error()
    assert rank(3, 4, 2, 0, 2) == 1 and b"multiple" in L.pg_last_error()
    # the rest of the surface
    assert L.pg_mem_info(None, None) == 1
    assert L.pg_set_tuning(b"no_such_key", 1) == 1 and b"unknown" in L.pg_last_error()
    assert L.pg_pcsense_tasks(None, None, None, None, None) == 1
