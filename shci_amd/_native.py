"""Loaders for the in-tree native libraries.  Missing libraries are a hard
error: the GPU path has no CPU fallback."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

_PKG = Path(__file__).resolve().parent
_HIP = _PKG / "libshci_hip.so"
_HOST = _PKG / "libshci_host.so"
_hip = None
_host = None

_u64p = C.POINTER(C.c_uint64)
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_vp = C.c_void_p


def hip_lib():
    """libshci_hip.so with argtypes for every export of include/shci_hip.h."""
    global _hip
    if _hip is not None:
        return _hip
    if not _HIP.exists():
        raise RuntimeError(f"HIP library missing: {_HIP}; run __graft_entry__.build()")
    lib = C.CDLL(str(_HIP), mode=C.RTLD_GLOBAL)
    from . import _abi

    _abi.declare(lib)
    _hip = lib
    return lib


def host_lib():
    global _host
    if _host is not None:
        return _host
    if not _HOST.exists():
        raise RuntimeError(f"host library missing: {_HOST}; run __graft_entry__.build()")
    hip_lib()  # resolves libshci_hip.so for the host library's link
    lib = C.CDLL(str(_HOST))
    lib.shci_host_last_error.restype = C.c_char_p
    lib.shci_host_system_from_fcidump.restype = _vp
    lib.shci_host_system_from_fcidump.argtypes = [C.c_char_p, C.c_uint, C.c_uint, C.POINTER(C.c_uint), C.c_uint,
                                                  C.POINTER(C.c_uint), C.c_uint, C.c_int]
    lib.shci_host_system_from_arrays.restype = _vp
    lib.shci_host_system_from_arrays.argtypes = [C.c_uint, C.c_uint, C.c_uint, _dp, _dp, C.c_double, C.c_uint64,
                                                 C.c_uint64]
    lib.shci_host_system_free.argtypes = [_vp]
    lib.shci_host_system_info.argtypes = [_vp, C.POINTER(C.c_uint), _dp, _u64p, _u64p, C.POINTER(C.c_size_t), _dp]
    lib.shci_host_system_integrals.argtypes = [_vp, _dp, _dp, C.POINTER(C.c_uint)]
    lib.shci_host_reset_dets.argtypes = [_vp]
    lib.shci_host_n_dets.restype = C.c_size_t
    lib.shci_host_n_dets.argtypes = [_vp]
    lib.shci_host_get_dets.argtypes = [_vp, _u64p, _u64p]
    lib.shci_host_set_dets.argtypes = [_vp, _u64p, _u64p, C.c_size_t]
    lib.shci_host_select.restype = C.c_longlong
    lib.shci_host_select.argtypes = [_vp, _dp, C.c_uint, C.c_double]
    lib.shci_host_dump_1rdm.argtypes = [C.c_char_p, C.c_uint, C.POINTER(C.c_uint), _dp]
    lib.shci_host_dump_2rdm.argtypes = [C.c_char_p, C.c_uint, C.POINTER(C.c_uint), _dp, C.c_int]
    lib.shci_host_io_last_error.restype = C.c_char_p
    lib.shci_host_save_wf.argtypes = [C.c_char_p, C.c_uint, C.c_uint, _u64p, _u64p, C.c_size_t, _dp, C.c_uint,
                                      C.c_double, _dp, C.c_int]
    lib.shci_host_wf_info.argtypes = [C.c_char_p, C.POINTER(C.c_uint), C.POINTER(C.c_uint), C.POINTER(C.c_size_t),
                                      C.POINTER(C.c_uint), C.POINTER(C.c_int)]
    lib.shci_host_load_wf.argtypes = [C.c_char_p, _u64p, _u64p, C.c_size_t, _dp, C.c_uint, _dp, _dp]
    lib.shci_host_wf_last_error.restype = C.c_char_p
    lib.shci_host_run_variation.argtypes = [_vp, _dp, C.c_uint, _dp, C.c_uint, C.c_double, C.c_uint, C.c_int,
                                            C.c_int, C.c_int, _dp]
    _host = lib
    return lib
