"""ctypes binding of the CPU FP64 oracle -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, where the oracle is the checker.  The product never imports this.
"""
from __future__ import annotations

import ctypes
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libsehf_oracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class _Problem(ctypes.Structure):
    _fields_ = [("l_cont", ctypes.c_int), ("n_occ", ctypes.c_int), ("Nr", ctypes.c_int),
                ("dr", ctypes.c_double), ("n_atoms", ctypes.c_int), ("Z", _ip), ("xyz", _dp),
                ("nlm_vh", ctypes.c_int), ("V_H", _dp), ("n_orb", ctypes.c_int),
                ("nlm_orb", ctypes.c_int), ("psi_lm", _dp), ("has_init", ctypes.c_int),
                ("init_psi_lm", _dp), ("pol_kind", ctypes.c_int), ("pol_alpha", ctypes.c_double),
                ("pol_rc", ctypes.c_double), ("allow_closed", ctypes.c_int)]


class _Out(ctypes.Structure):
    _fields_ = [("K_kme", _dp), ("A", _dp), ("B", _dp), ("K_fit", _dp), ("b_overlap", _dp),
                ("d_raw", _dp), ("D_raw", _dp), ("D", _dp), ("d_corr", _dp), ("diag", _dp),
                ("n_dump_sinv", ctypes.c_int), ("dump_sinv_n", _ip), ("dump_sinv", _dp),
                ("n_dump_rinv", ctypes.c_int), ("dump_rinv_n", _ip), ("dump_rinv", _dp),
                ("n_dump_psi", ctypes.c_int), ("dump_psi_n", _ip), ("dump_psi", _dp),
                ("t_stage", _dp), ("tangent", ctypes.c_int), ("tan_E", ctypes.c_double),
                ("tan_alpha", ctypes.c_double), ("dK", _dp), ("dA", _dp), ("dB", _dp),
                ("db_overlap", _dp), ("dd_raw", _dp), ("dD", _dp), ("dump_dpsi", _dp),
                ("dump_drinv", _dp)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = ctypes.CDLL(LIB_PATH)
        for name in ("sehf_oracle_w3j", "sehf_oracle_gaunt_real"):
            getattr(L, name).restype = ctypes.c_double
            getattr(L, name).argtypes = [ctypes.c_int] * 6
        L.sehf_oracle_real_ylm.restype = ctypes.c_double
        L.sehf_oracle_real_ylm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        for name in ("sehf_oracle_riccati_j", "sehf_oracle_riccati_y"):
            getattr(L, name).restype = ctypes.c_double
            getattr(L, name).argtypes = [ctypes.c_int, ctypes.c_double]
        L.sehf_oracle_min_eig_sym.restype = ctypes.c_double
        L.sehf_oracle_simpson.restype = ctypes.c_double
        L.sehf_oracle_simpson.argtypes = [_dp, ctypes.c_int, ctypes.c_double]
        L.sehf_oracle_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


class ProblemRef:
    """Keeps the numpy buffers alive while the C struct points at them."""

    def __init__(self, p):
        self.keep = []

        def arr(x, dt=np.float64):
            y = np.ascontiguousarray(x, dtype=dt)
            self.keep.append(y)
            return y
        Z = arr(p.Z, np.int32)
        xyz = arr(p.xyz)
        vh = arr(p.V_H)
        psi = arr(p.psi_lm)
        init = arr(p.init_psi_lm) if p.init_psi_lm is not None else None
        self.s = _Problem(p.l_cont, p.n_occ, p.Nr, p.dr, len(Z), Z.ctypes.data_as(_ip), _ptr(xyz),
                          vh.shape[0], _ptr(vh), psi.shape[0], psi.shape[1], _ptr(psi),
                          1 if init is not None else 0, _ptr(init), p.pol_kind, p.pol_alpha,
                          p.pol_rc, 1 if p.allow_closed else 0)


def solve(p, k, dump_sinv=(), dump_rinv=(), dump_psi=(), tangent=None):
    """Run the oracle for one k; returns a dict of numpy outputs.

    tangent=(tE, talpha) adds the directional derivatives dK, dA, dB,
    db_overlap, dd_raw, dD along E (Hartree) and the polarisation strength."""
    L = lib()
    ref = ProblemRef(p)
    N, no = p.n_psi, p.n_occ
    o = {"K_kme": np.zeros((N, N), order="F"), "A": np.zeros((N, N), order="F"),
         "B": np.zeros((N, N), order="F"), "K_fit": np.zeros((N, N), order="F"),
         "b_overlap": np.zeros((N, no), order="F"), "d_raw": np.zeros((6, N)),
         "D_raw": np.zeros((6, N), dtype=np.complex128), "D": np.zeros((6, N), dtype=np.complex128),
         "d_corr": np.zeros((6, no)), "diag": np.zeros(16), "t_stage": np.zeros(8)}
    ds = np.asarray(dump_sinv, dtype=np.int32)
    dri = np.asarray(dump_rinv, dtype=np.int32)
    dp = np.asarray(dump_psi, dtype=np.int32)
    o["sinv"] = np.zeros((len(ds), N, N))
    o["rinv"] = np.zeros((len(dri), p.n_tot, p.n_tot))
    o["psi"] = np.zeros((len(dp), N, N))
    out = _Out(_ptr(o["K_kme"]), _ptr(o["A"]), _ptr(o["B"]), _ptr(o["K_fit"]), _ptr(o["b_overlap"]),
               _ptr(o["d_raw"]), o["D_raw"].ctypes.data_as(_dp), o["D"].ctypes.data_as(_dp),
               _ptr(o["d_corr"]), _ptr(o["diag"]),
               len(ds), ds.ctypes.data_as(_ip), _ptr(o["sinv"]),
               len(dri), dri.ctypes.data_as(_ip), _ptr(o["rinv"]),
               len(dp), dp.ctypes.data_as(_ip), _ptr(o["psi"]), _ptr(o["t_stage"]))
    if tangent is not None:
        o.update({"dK": np.zeros((N, N), order="F"), "dA": np.zeros((N, N), order="F"),
                  "dB": np.zeros((N, N), order="F"), "db_overlap": np.zeros((N, no), order="F"),
                  "dd_raw": np.zeros((6, N)), "dD": np.zeros((6, N), dtype=np.complex128)})
        out.tangent = 1
        out.tan_E, out.tan_alpha = float(tangent[0]), float(tangent[1])
        out.dK, out.dA, out.dB = _ptr(o["dK"]), _ptr(o["dA"]), _ptr(o["dB"])
        out.db_overlap, out.dd_raw = _ptr(o["db_overlap"]), _ptr(o["dd_raw"])
        out.dD = o["dD"].ctypes.data_as(_dp)
        o["dpsi"] = np.zeros((len(dp), N, N))
        out.dump_dpsi = _ptr(o["dpsi"])
        o["drinv"] = np.zeros((len(dri), p.n_tot, p.n_tot))
        out.dump_drinv = _ptr(o["drinv"])
    rc = L.sehf_oracle_solve(ctypes.byref(ref.s), ctypes.c_double(k), ctypes.byref(out))
    if rc != 0:
        raise RuntimeError("oracle: " + L.sehf_oracle_last_error().decode())
    # dumps were written column-major per matrix
    o["sinv"] = np.ascontiguousarray(o["sinv"].transpose(0, 2, 1))
    o["rinv"] = np.ascontiguousarray(o["rinv"].transpose(0, 2, 1))
    o["psi"] = np.ascontiguousarray(o["psi"].transpose(0, 2, 1))
    if "drinv" in o:
        o["drinv"] = np.ascontiguousarray(o["drinv"].transpose(0, 2, 1))
    if "dpsi" in o:
        o["dpsi"] = np.ascontiguousarray(o["dpsi"].transpose(0, 2, 1))
    return o


def solve_scan(p, ks, threads=None, **kw):
    """Oracle over a k list; energies run concurrently (the C code releases the GIL)."""
    threads = threads or min(len(ks), os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(lambda k: solve(p, float(k), **kw), ks))


def noise_floor(p, ks, **kw):
    """Oracle run with Gauss-Jordan elimination in place of LU: the spread to
    the LU run measures the reference algorithm's own rounding noise."""
    L = lib()
    L.sehf_oracle_set_inverse_mode(1)
    try:
        return solve_scan(p, ks, **kw)
    finally:
        L.sehf_oracle_set_inverse_mode(0)


def solve_scan_rounded_inverses(p, ks, **kw):
    """Oracle run with every dense inverse evaluated in 80-bit extended
    precision and rounded once to FP64 (correctly rounded inverses; all other
    operations as in the reference): the referee for outputs whose FP64
    value the reference's own inverse rounding already moves."""
    L = lib()
    L.sehf_oracle_set_inverse_mode(32)
    try:
        return solve_scan(p, ks, **kw)
    finally:
        L.sehf_oracle_set_inverse_mode(0)


def time_steps(p, k, n_steps):
    L = lib()
    ref = ProblemRef(p)
    t = np.zeros(3)
    rc = L.sehf_oracle_time_steps(ctypes.byref(ref.s), ctypes.c_double(k), n_steps, _ptr(t))
    if rc != 0:
        raise RuntimeError("oracle: " + L.sehf_oracle_last_error().decode())
    return t


def time_steps_multi(p, ks, n_steps):
    """Per-step seconds [nk, 3] for nk concurrent energies, plus (setup s, wall s)."""
    L = lib()
    ref = ProblemRef(p)
    ks = np.ascontiguousarray(ks, dtype=np.float64)
    t = np.zeros((len(ks), 3))
    wall = np.zeros(2)
    rc = L.sehf_oracle_time_steps_multi(ctypes.byref(ref.s), _ptr(ks), len(ks), n_steps, _ptr(t), _ptr(wall))
    if rc != 0:
        raise RuntimeError("oracle: " + L.sehf_oracle_last_error().decode())
    return t, wall


def matrix_fn(name, p, ir, shape):
    L = lib()
    ref = ProblemRef(p)
    M = np.zeros(shape[::-1])
    rc = getattr(L, name)(ctypes.byref(ref.s), ir, _ptr(M))
    if rc != 0:
        raise RuntimeError("oracle: " + L.sehf_oracle_last_error().decode())
    return M.T.copy()


def potential(p, ir):
    return matrix_fn("sehf_oracle_potential", p, ir, (p.n_psi, p.n_psi))


def exchange_Q(p, ir):
    return matrix_fn("sehf_oracle_exchange_Q", p, ir, (p.n_psi, p.n_f))


def inverse_general(A):
    A = np.asfortranarray(A, dtype=np.float64)
    X = np.zeros_like(A, order="F")
    lib().sehf_oracle_inverse_general(A.shape[0], _ptr(A), _ptr(X))
    return X


def min_eig_sym(A):
    A = np.asfortranarray(A, dtype=np.float64)
    return lib().sehf_oracle_min_eig_sym(A.shape[0], _ptr(A))


def frp_step(Winv, Rprev):
    W = np.asfortranarray(Winv, dtype=np.float64)
    R = np.asfortranarray(Rprev, dtype=np.float64)
    X = np.zeros_like(W, order="F")
    lib().sehf_oracle_frp_step(W.shape[0], _ptr(W), _ptr(R), _ptr(X))
    return X


def five_point_derivative(f, dr):
    f = np.ascontiguousarray(f, dtype=np.float64)
    df = np.zeros_like(f)
    lib().sehf_oracle_five_point_derivative(_ptr(f), f.size, ctypes.c_double(dr), _ptr(df))
    return df


def simpson(f, h):
    f = np.ascontiguousarray(f, dtype=np.float64)
    return lib().sehf_oracle_simpson(_ptr(f), f.size, h)


def set_threads(n: int):
    """Threads for the loops over independent columns inside one solve
    (bitwise the same results as one thread)."""
    lib().sehf_oracle_set_threads(int(n))
