"""ctypes binding of libpowergrid_amd.so (the C ABI in include/powergrid_amd.h).

The library is the product: if it is missing this module raises instead of
falling back to anything else.  Build it with `python -m powergrid_amd.build`.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpowergrid_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "powergrid_amd.h")

_lib = None

vp = C.c_void_p
sz = C.c_size_t
f32 = C.c_float
i32 = C.c_int
dp = C.POINTER(C.c_double)


class CgInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("exit_reason", C.c_int), ("final_resid", C.c_double)]


_SIGS = {
    "pg_status_string": (C.c_char_p, [i32]),
    "pg_last_error": (C.c_char_p, []),
    "pg_version": (C.c_char_p, []),
    "pg_set_device": (i32, [i32]),
    "pg_malloc": (i32, [C.POINTER(vp), sz]),
    "pg_free": (i32, [vp]),
    "pg_mem_info": (i32, [C.POINTER(sz), C.POINTER(sz)]),
    "pg_op_tuning": (i32, [vp, i32, C.POINTER(i32), C.POINTER(i32)]),
    "pg_gnufft_zclasses": (i32, [vp, C.POINTER(sz)]),
    "pg_memcpy_htod": (i32, [vp, vp, sz, vp]),
    "pg_memcpy_dtoh": (i32, [vp, vp, sz, vp]),
    "pg_memcpy_dtod": (i32, [vp, vp, sz, vp]),
    "pg_stream_create": (i32, [C.POINTER(vp)]),
    "pg_stream_destroy": (i32, [vp]),
    "pg_stream_synchronize": (i32, [vp]),
    "pg_gnufft_create": (i32, [C.POINTER(vp), sz, f32, sz, sz, sz, vp, vp, vp, vp, vp, vp, f32]),
    "pg_gdft_create": (i32, [C.POINTER(vp), sz, sz, vp, vp, vp, vp, vp, vp, vp, vp, i32]),
    "pg_gdftr2_create": (i32, [C.POINTER(vp), sz, sz, vp, vp, vp, vp, vp, vp, vp, vp, sz, sz,
                               sz]),
    "pg_ts_create": (i32, [C.POINTER(vp), vp, vp, vp, sz, sz, sz, sz, sz]),
    "pg_sense_create": (i32, [C.POINTER(vp), vp, vp, sz, sz, sz]),
    "pg_pcsense_create": (i32, [C.POINTER(vp), vp, vp, vp, sz, sz, sz, sz, vp, sz, sz, vp, vp, vp,
                                sz, sz, C.c_float]),
    "pg_pcsense_gdft_create": (i32, [C.POINTER(vp), vp, vp, vp, sz, sz, sz, sz, vp, vp, vp, vp,
                                     sz, sz]),
    "pg_pcsense_rank_create": (i32, [C.POINTER(vp), vp, vp, vp, sz, sz, sz, sz, vp, sz, sz, vp, vp,
                                     vp, sz, sz, C.c_float, sz, sz]),
    "pg_pcsense_tasks": (i32, [vp, C.POINTER(sz), vp, vp, vp]),
    "pg_op_destroy": (i32, [vp]),
    "pg_op_shape": (i32, [vp, C.POINTER(sz), C.POINTER(sz)]),
    "pg_op_set_mode": (i32, [vp, i32]),
    "pg_op_forward": (i32, [vp, vp, vp, vp]),
    "pg_op_adjoint": (i32, [vp, vp, vp, vp]),
    "pg_op_forward_host": (i32, [vp, vp, vp]),
    "pg_op_adjoint_host": (i32, [vp, vp, vp]),
    "pg_dft_apply_static": (i32, [vp, C.c_int, vp, vp, vp]),
    "pg_trace_ranges": (i32, [C.POINTER(C.c_int), C.POINTER(C.c_ulonglong)]),
    "pg_gnufft_grid_size": (i32, [vp, C.POINTER(sz), C.POINTER(sz), C.POINTER(sz)]),
    "pg_gnufft_forward_spatial": (i32, [vp, vp, vp, vp]),
    "pg_gnufft_adjoint_spatial": (i32, [vp, vp, vp, vp]),
    "pg_density_compensation": (i32, [vp, vp, vp, sz, sz, sz, vp, vp]),
    "pg_grid_coil_images": (i32, [vp, vp, vp, sz, vp, vp]),
    "pg_sum_of_squares": (i32, [vp, sz, sz, vp, vp]),
    "pg_write_nifti_real": (i32, [C.c_char_p, vp, sz, sz, sz, sz, sz]),
    "pg_write_nifti_mag_phase": (i32, [C.c_char_p, vp, sz, sz, sz, sz, sz, vp]),
    "pg_penalty_create": (i32, [C.POINTER(vp), i32, sz, sz, sz, f32, f32, sz]),
    "pg_penalty_destroy": (i32, [vp]),
    "pg_penalty_gradient": (i32, [vp, vp, vp, vp]),
    "pg_penalty_denom": (i32, [vp, vp, vp, dp, vp]),
    "pg_penalty_value": (i32, [vp, vp, dp, vp]),
    "pg_solve_pwls_pcg": (i32, [vp, vp, vp, vp, vp, sz, i32, vp, dp, C.POINTER(CgInfo), vp]),
    "pg_recon_solve": (i32, [vp, vp, vp, sz, i32, vp, dp, C.POINTER(CgInfo), vp]),
    "pg_solve_pwls_pcg_tol": (i32, [vp, vp, vp, vp, vp, sz, C.c_double, i32, vp, dp,
                                    C.POINTER(CgInfo), vp]),
    "pg_recon_solve_tol": (i32, [vp, vp, vp, sz, C.c_double, i32, vp, dp, C.POINTER(CgInfo), vp]),
    "pg_recon_solve_batch": (i32, [sz, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), sz,
                                   C.c_double, C.POINTER(vp), C.POINTER(CgInfo), vp]),
    "pg_set_tuning": (i32, [C.c_char_p, i32]),
    "pg_profile_enable": (i32, [i32]),
    "pg_profile_reset": (i32, []),
    "pg_profile_query": (i32, [C.c_char_p, C.POINTER(i32), dp]),
}


def header_symbols():
    """Names of the functions declared in include/powergrid_amd.h."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pg_[a-z0-9_]+)\s*\(", txt)))


class PgError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{msg} (status {status})")
        self.status = status


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python -m powergrid_amd.build` "
                "(there is no fallback implementation)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != 0:
        L = lib()
        raise PgError(status, (L.pg_last_error() or b"").decode() or
                      L.pg_status_string(status).decode())
    return status
