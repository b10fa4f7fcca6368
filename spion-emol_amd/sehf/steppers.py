"""Python mirror of the reference's GPU stepper classes over the C ABI.

GpuContext / GpuForwardStepper / GpuSinvStepper / GpuBackStepper /
GpuDipoleEngine (src/scatt/GpuPropagate.hpp:52-279, GpuDipoleEngine.hpp:85-149):
same constructors, method names, argument meaning and error behaviour
(a failed call raises).  Arrays are numpy float64; matrices are passed in
column-major (Fortran) order like the reference's Eigen matrices.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import SehfError, check, dptr, load

_vp = ctypes.c_void_p


def _f(a, shape=None):
    a = np.asfortranarray(a, dtype=np.float64)
    if shape is not None and a.shape != shape:
        raise SehfError(f"wrong shape {a.shape}, expected {shape}")
    return a


# Field names of the steppers' per-phase timing counters (nanoseconds), in
# the order of the sehf_*_stats_t structs of include/sehf.h.
_STATS_FIELDS = {
    "frp": ("t_upload_ns", "t_u_combine_ns", "t_inverse_ns", "t_download_ns", "n_steps"),
    "sinv": ("t_upload_ns", "t_schur_ns", "t_inverse_ns", "t_download_ns", "n_steps"),
    "bp": ("t_upload_ns", "t_gemm_z_ns", "t_gemm_y_ns", "t_extract_ns", "n_steps"),
    "dme": ("t_upload_ns", "t_gemm_ns", "t_download_ns", "n_steps"),
}


class _Stats:
    """stats() / reset_stats() of the reference's steppers (GpuPropagate.hpp
    Stats structs): accumulated per-phase nanoseconds and the step count."""
    _kind = ""

    def stats(self) -> dict:
        names = _STATS_FIELDS[self._kind]
        buf = (ctypes.c_uint64 * len(names))()
        check(getattr(load(), f"sehf_{self._kind}_stats")(self._h, ctypes.cast(buf, _vp)), "stats")
        return dict(zip(names, (int(v) for v in buf)))

    def reset_stats(self):
        check(getattr(load(), f"sehf_{self._kind}_reset_stats")(self._h), "reset_stats")


class GpuContext:
    def __init__(self, prefer_gpu: bool = True):
        L = load()
        h = _vp()
        check(L.sehf_ctx_create(1 if prefer_gpu else 0, ctypes.byref(h)), "GpuContext")
        self._h = h

    @staticmethod
    def gpu_available() -> bool:
        return bool(load().sehf_gpu_available())

    def info(self) -> dict:
        name = ctypes.create_string_buffer(256)
        mem = ctypes.c_size_t()
        is_gpu = ctypes.c_int()
        check(load().sehf_ctx_info(self._h, name, 256, ctypes.byref(mem), ctypes.byref(is_gpu)))
        return {"device_name": name.value.decode(), "global_mem_bytes": mem.value, "is_gpu": bool(is_gpu.value)}

    def __del__(self):
        if getattr(self, "_h", None):
            load().sehf_ctx_destroy(self._h)
            self._h = None


class GpuForwardStepper(_Stats):
    _kind = "frp"

    def __init__(self, ctx: GpuContext, N_total: int):
        self.ctx, self.N = ctx, N_total
        h = _vp()
        check(load().sehf_frp_create(ctx._h, N_total, ctypes.byref(h)), "GpuForwardStepper")
        self._h = h

    def init_R_prev_inv(self, R):
        R = _f(R, (self.N, self.N))
        check(load().sehf_frp_init_R_prev_inv(self._h, dptr(R)), "init_R_prev_inv")

    def step(self, W_inv):
        """Returns (Rinv, max|Rinv - Rinv^T|)."""
        W = _f(W_inv, (self.N, self.N))
        out = np.zeros((self.N, self.N), order="F")
        sym = ctypes.c_double()
        check(load().sehf_frp_step(self._h, dptr(W), dptr(out), ctypes.byref(sym)), "GpuForwardStepper.step")
        return out, sym.value

    def enable_symmetric_inverse(self, enable: bool):
        check(load().sehf_frp_enable_symmetric_inverse(self._h, 1 if enable else 0))

    def is_symmetric_inverse_active(self) -> bool:
        return bool(load().sehf_frp_is_symmetric_inverse_active(self._h))

    def __del__(self):
        if getattr(self, "_h", None):
            load().sehf_frp_destroy(self._h)
            self._h = None


class GpuSinvStepper(_Stats):
    _kind = "sinv"

    def __init__(self, ctx: GpuContext, N_psi: int, N_f: int):
        self.ctx, self.Np, self.Nf = ctx, N_psi, N_f
        h = _vp()
        check(load().sehf_sinv_create(ctx._h, N_psi, N_f, ctypes.byref(h)), "GpuSinvStepper")
        self._h = h

    def step(self, A, B, Dinv):
        A = _f(A, (self.Np, self.Np))
        B = _f(B, (self.Np, self.Nf))
        D = np.ascontiguousarray(Dinv, dtype=np.float64)
        out = np.zeros((self.Np, self.Np), order="F")
        sym = ctypes.c_double()
        check(load().sehf_sinv_step(self._h, dptr(A), dptr(B), dptr(D), dptr(out), ctypes.byref(sym)),
              "GpuSinvStepper.step")
        return out, sym.value

    def step_inverse_only(self, A):
        A = _f(A, (self.Np, self.Np))
        out = np.zeros((self.Np, self.Np), order="F")
        sym = ctypes.c_double()
        check(load().sehf_sinv_step_inverse_only(self._h, dptr(A), dptr(out), ctypes.byref(sym)),
              "GpuSinvStepper.step_inverse_only")
        return out, sym.value

    def __del__(self):
        if getattr(self, "_h", None):
            load().sehf_sinv_destroy(self._h)
            self._h = None


class GpuBackStepper(_Stats):
    _kind = "bp"

    def __init__(self, ctx: GpuContext, N_total: int, N_psi: int, N_f: int):
        self.ctx, self.Nt, self.Np, self.Nf = ctx, N_total, N_psi, N_f
        h = _vp()
        check(load().sehf_bp_create(ctx._h, N_total, N_psi, N_f, ctypes.byref(h)), "GpuBackStepper")
        self._h = h

    def init_Z(self, Z):
        Z = _f(Z, (self.Nt, self.Np))
        check(load().sehf_bp_init_Z(self._h, dptr(Z)), "init_Z")

    def step(self, Rinv, W_inv, compute_f: bool = False):
        """Returns (psi, f or None)."""
        R = _f(Rinv, (self.Nt, self.Nt))
        W = _f(W_inv, (self.Nt, self.Nt))
        psi = np.zeros((self.Np, self.Np), order="F")
        f = np.zeros((self.Nf, self.Np), order="F") if compute_f else None
        check(load().sehf_bp_step(self._h, dptr(R), dptr(W), dptr(psi), dptr(f), 1 if compute_f else 0),
              "GpuBackStepper.step")
        return psi, f

    def get_Z(self):
        Z = np.zeros((self.Nt, self.Np), order="F")
        check(load().sehf_bp_get_Z(self._h, dptr(Z)), "get_Z")
        return Z

    def __del__(self):
        if getattr(self, "_h", None):
            load().sehf_bp_destroy(self._h)
            self._h = None


class GpuDipoleEngine(_Stats):
    _kind = "dme"

    def __init__(self, ctx: GpuContext, N_psi: int, n_slots: int):
        self.ctx, self.Np, self.ns = ctx, N_psi, n_slots
        h = _vp()
        check(load().sehf_dme_create(ctx._h, N_psi, n_slots, ctypes.byref(h)), "GpuDipoleEngine")
        self._h = h

    def step(self, V, psi_ir):
        V = _f(V, (self.ns, self.Np))
        P = _f(psi_ir, (self.Np, self.Np))
        out = np.zeros((self.ns, self.Np), order="F")
        check(load().sehf_dme_step(self._h, dptr(V), dptr(P), dptr(out)), "GpuDipoleEngine.step")
        return out

    def __del__(self):
        if getattr(self, "_h", None):
            load().sehf_dme_destroy(self._h)
            self._h = None
