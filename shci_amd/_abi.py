"""ctypes declarations for every export of include/shci_hip.h."""
from __future__ import annotations

import ctypes as C

_u64p = C.POINTER(C.c_uint64)
_i64p = C.POINTER(C.c_int64)
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p

SHCI_OK = 0
SHCI_STORE_FP64 = 1 << 0
SHCI_STORE_FP32 = 1 << 1
SHCI_KEEP_ZEROS = 1 << 2
SHCI_DEVICE_PTRS = 1 << 3
SHCI_FP32 = 1 << 4
SHCI_LAYOUT_SEGMENTS = 1 << 5
SHCI_F64_VECTORS = 1 << 6
SHCI_TIME_SYM = 1 << 7
SHCI_DAVIDSON_FP64 = 0
SHCI_DAVIDSON_MIXED = 1

# name -> (restype, argtypes)
EXPORTS = {
    "shci_ctx_create": (C.c_int, [C.POINTER(_vp), C.c_int]),
    "shci_ctx_destroy": (C.c_int, [_vp]),
    "shci_last_error": (C.c_char_p, [_vp]),
    "shci_set_integrals": (C.c_int, [_vp, C.c_uint32, _dp, _dp, C.c_double]),
    "shci_build_hamiltonian": (C.c_int, [_vp, _vp, _vp, C.c_uint64, C.c_uint64, C.c_uint32]),
    "shci_rebuild_hamiltonian": (C.c_int, [_vp, C.c_uint32]),
    "shci_update_existing_elems": (C.c_int, [_vp]),
    "shci_clear": (C.c_int, [_vp]),
    "shci_get_counts": (C.c_int, [_vp, _u64p, _u64p, _u64p]),
    "shci_get_diag": (C.c_int, [_vp, _dp]),
    "shci_spmv": (C.c_int, [_vp, _vp, _vp, C.c_uint32]),
    "shci_spmv_noalloc": (C.c_int, [_vp, _vp, _vp, C.c_uint32]),
    "shci_spmm": (C.c_int, [_vp, _vp, _vp, C.c_uint32, C.c_uint32]),
    "shci_davidson": (C.c_int, [_vp, C.c_uint32, _dp, C.c_double, C.c_uint32, C.c_uint32, _dp, _dp, _ip, _dp, _dp,
                                C.c_int, _ip]),
    "shci_rdm": (C.c_int, [_vp, C.c_uint32, _dp, _dp, _dp]),
    "shci_export_upper": (C.c_int, [_vp, _i64p, _i64p, _dp]),
    "shci_select": (C.c_int, [_vp, _u64p, _u64p, C.c_uint64, _dp, C.c_uint32, C.c_double, _dp, _u64p, C.c_uint32]),
    "shci_select_result": (C.c_int, [_vp, _u64p, _u64p, C.c_uint64]),
    "shci_select_timings": (C.c_int, [_vp, _fp]),
    "shci_pt_dtm": (C.c_int, [_vp, _u64p, _u64p, C.c_uint64, _dp, C.c_double, C.c_double, C.c_uint32, _dp,
                             _u64p]),
    "shci_get_build_timings": (C.c_int, [_vp, _fp]),
    "shci_time_spmv": (C.c_int, [_vp, C.c_uint32, C.c_int, _fp, _dp]),
    "shci_time_spmm": (C.c_int, [_vp, C.c_uint32, C.c_int, _fp, _dp]),
    "shci_capture_ritz": (C.c_int, [_vp, C.c_uint32]),
    "shci_ritz_energies": (C.c_int, [_vp, _dp, C.c_uint32, C.POINTER(C.c_uint32)]),
    "shci_memory": (C.c_int, [_vp, _u64p]),
    "shci_layout_stats": (C.c_int, [_vp, _u64p]),
    "shci_segment_ranges": (C.c_int, [_vp, _u64p]),
    "shci_segment_kept": (C.c_int, [_vp, _u64p]),
}


def declare(lib):
    for name, (res, args) in EXPORTS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
