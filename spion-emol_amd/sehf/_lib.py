"""ctypes loader for the in-tree HIP library (spion-emol_amd/_build/libsehf.so).

There is no host fallback: if the library is missing or no device is visible,
every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_ROOT, "_build", "libsehf.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p

# (name, restype, argtypes) for every symbol include/sehf.h declares.
SIGNATURES = [
    ("sehf_last_error", ctypes.c_char_p, []),
    ("sehf_gpu_available", ctypes.c_int, []),
    ("sehf_ctx_create", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    ("sehf_ctx_destroy", None, [_vp]),
    ("sehf_ctx_info", ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t), _ip]),
    ("sehf_frp_create", ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp)]),
    ("sehf_frp_init_R_prev_inv", ctypes.c_int, [_vp, _dp]),
    ("sehf_frp_step", ctypes.c_int, [_vp, _dp, _dp, _dp]),
    ("sehf_frp_enable_symmetric_inverse", ctypes.c_int, [_vp, ctypes.c_int]),
    ("sehf_frp_step_noalloc", ctypes.c_int, [ctypes.c_int, _vp, _vp, _vp, _dp]),
    ("sehf_frp_is_symmetric_inverse_active", ctypes.c_int, [_vp]),
    ("sehf_frp_destroy", None, [_vp]),
    ("sehf_sinv_create", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    ("sehf_sinv_step", ctypes.c_int, [_vp, _dp, _dp, _dp, _dp, _dp]),
    ("sehf_sinv_step_inverse_only", ctypes.c_int, [_vp, _dp, _dp, _dp]),
    ("sehf_sinv_destroy", None, [_vp]),
    ("sehf_bp_create", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    ("sehf_bp_init_Z", ctypes.c_int, [_vp, _dp]),
    ("sehf_bp_step", ctypes.c_int, [_vp, _dp, _dp, _dp, _dp, ctypes.c_int]),
    ("sehf_bp_get_Z", ctypes.c_int, [_vp, _dp]),
    ("sehf_bp_destroy", None, [_vp]),
    ("sehf_dme_create", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    ("sehf_dme_step", ctypes.c_int, [_vp, _dp, _dp, _dp]),
    ("sehf_dme_destroy", None, [_vp]),
    ("sehf_frp_stats", ctypes.c_int, [_vp, _vp]),
    ("sehf_frp_reset_stats", ctypes.c_int, [_vp]),
    ("sehf_sinv_stats", ctypes.c_int, [_vp, _vp]),
    ("sehf_sinv_reset_stats", ctypes.c_int, [_vp]),
    ("sehf_bp_stats", ctypes.c_int, [_vp, _vp]),
    ("sehf_bp_reset_stats", ctypes.c_int, [_vp]),
    ("sehf_dme_stats", ctypes.c_int, [_vp, _vp]),
    ("sehf_dme_reset_stats", ctypes.c_int, [_vp]),
    ("sehf_scan_create", ctypes.c_int, [_vp, _vp, ctypes.POINTER(_vp)]),
    ("sehf_scan_info", ctypes.c_int, [_vp, _ip]),
    ("sehf_scan_run", ctypes.c_int, [_vp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp]),
    ("sehf_scan_run_tangent", ctypes.c_int, [_vp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp]),
    ("sehf_scan_run_segment", ctypes.c_int, [_vp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _ip]),
    ("sehf_scan_autotune", ctypes.c_int, [_vp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _ip, _ip, _dp,
                                          _ip]),
    ("sehf_set_gemm_tiles", ctypes.c_int, [ctypes.c_int]),
    ("sehf_set_gj_pivoting", ctypes.c_int, [ctypes.c_int]),
    ("sehf_scan_destroy", None, [_vp]),
    ("sehf_scan_stats", ctypes.c_int, [_vp, _dp]),
    ("sehf_scan_set_groups", ctypes.c_int, [_vp, ctypes.c_int]),
    ("sehf_scan_set_dipole_path", ctypes.c_int, [_vp, ctypes.c_int]),
    ("sehf_prof_begin", ctypes.c_int, [ctypes.c_int]),
    ("sehf_prof_end", ctypes.c_int, [_dp]),
    ("sehf_time_gemm", ctypes.c_int, [ctypes.c_int] * 9 + [_dp]),
    ("sehf_time_inverse", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    ("sehf_time_spd_inverse", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    ("sehf_copy_probe", ctypes.c_int, [ctypes.c_size_t, _dp]),
    ("sehf_host_potential", ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    ("sehf_host_exchange_B", ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    ("sehf_host_sizes", ctypes.c_int, [_vp, _ip]),
    ("sehf_host_dipole_tables", ctypes.c_int, [_vp, _dp, _dp, _ip]),
    ("sehf_host_ric_j", ctypes.c_double, [ctypes.c_int, ctypes.c_double]),
    ("sehf_host_ric_y", ctypes.c_double, [ctypes.c_int, ctypes.c_double]),
    ("sehf_host_real_gaunt", ctypes.c_double, [ctypes.c_int] * 6),
    ("sehf_host_min_eig", ctypes.c_double, [ctypes.c_int, _dp]),
]

_lib = None


class SehfError(RuntimeError):
    pass


def load():
    """Load libsehf.so (built by __graft_entry__.build()); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SehfError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = load().sehf_last_error().decode()
        raise SehfError(f"{what}: {msg}" if what else msg)


def dptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None
