"""Batched energy scan through the C ABI (sehf_scan_*).

`solve_scan(problem, ks)` runs the whole hot path -- Schur inverse, forward
ratio recursion, K matrix, back-propagation, asymptotic fit and the six
dipole slices -- for every k on the device and returns one dict per energy
with the DipoleEnergyPayload fields (src/scatt/DipoleIO.hpp:82-112).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from ._lib import check, load

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class Problem_C(ctypes.Structure):
    _fields_ = [("l_cont", ctypes.c_int), ("n_occ", ctypes.c_int), ("Nr", ctypes.c_int),
                ("dr", ctypes.c_double), ("n_atoms", ctypes.c_int), ("Z", _ip), ("xyz", _dp),
                ("nlm_vh", ctypes.c_int), ("V_H", _dp), ("n_orb", ctypes.c_int),
                ("nlm_orb", ctypes.c_int), ("psi_lm", _dp), ("has_init", ctypes.c_int),
                ("init_psi_lm", _dp), ("pol_kind", ctypes.c_int), ("pol_alpha", ctypes.c_double),
                ("pol_rc", ctypes.c_double), ("allow_closed", ctypes.c_int)]


class ScanOut_C(ctypes.Structure):
    _fields_ = [(n, _dp) for n in ("K", "A", "B", "K_fit", "b_overlap", "d_raw", "D_raw", "D",
                                   "d_corr", "diag")]


class TangentOut_C(ctypes.Structure):
    _fields_ = [(n, _dp) for n in ("dK", "dA", "dB", "db_overlap", "dd_raw", "dD")]


class PlanIn_C(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("Nr", "n_psi", "n_f", "n_transition", "n_occ", "n_exp", "n_lambda",
                                            "tangent", "dipole_path", "n_energies", "batch_cap")] + \
               [(n, ctypes.c_size_t) for n in ("device_bytes", "host_bytes", "runtime_bytes")] + \
               [("reserve_fraction", ctypes.c_double)]


class Plan_C(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("tables", "ckpt", "seg_len", "n_ckpt", "batch")] + \
               [(n, ctypes.c_size_t) for n in ("budget_bytes", "shared_bytes", "per_energy_bytes",
                                               "host_per_energy_bytes")] + \
               [("report", ctypes.c_char * 1536)]


TABLES = {"resident": 0, "streamed": 1}      # SEHF_TABLES_*
CHECKPOINTS = {"device": 0, "host": 1}       # SEHF_CKPT_*


def plan_storage(problem, n_energies: int, tangent: bool = False, dipole_path: int = 0, device_bytes: int = 0,
                 host_bytes: int = 0, runtime_bytes: int = 2 << 30, reserve_fraction: float = 0.10,
                 batch_cap: int = 0, n_transition: int | None = None) -> dict:
    """Device storage plan for scanning `n_energies` energies of `problem`
    (sehf_plan_storage): tables resident or streamed, checkpoint level and
    spacing, batch width, and the report.  n_transition defaults to Nr (the
    exchange block kept over the whole grid, the upper bound)."""
    L = _register(load())
    p = problem
    q = PlanIn_C(p.Nr, p.n_psi, p.n_f, p.Nr if n_transition is None else int(n_transition), p.n_occ,
                 (2 * p.l_cont + 1) ** 2, p.n_lambda, 1 if tangent else 0, int(dipole_path), int(n_energies),
                 int(batch_cap), int(device_bytes), int(host_bytes), int(runtime_bytes), float(reserve_fraction))
    out = Plan_C()
    check(L.sehf_plan_storage(ctypes.byref(q), ctypes.byref(out)), "sehf_plan_storage")
    inv_t = {v: k for k, v in TABLES.items()}
    inv_c = {v: k for k, v in CHECKPOINTS.items()}
    return {"tables": inv_t[out.tables], "checkpoints": inv_c[out.ckpt], "seg_len": out.seg_len,
            "n_ckpt": out.n_ckpt, "batch": out.batch, "budget_bytes": out.budget_bytes,
            "shared_bytes": out.shared_bytes, "per_energy_bytes": out.per_energy_bytes,
            "host_per_energy_bytes": out.host_per_energy_bytes, "report": out.report.decode()}


def _register(L):
    if getattr(L, "_scan_registered", False):
        return L
    L.sehf_scan_create.argtypes = [ctypes.c_void_p, ctypes.POINTER(Problem_C), ctypes.POINTER(ctypes.c_void_p)]
    L.sehf_scan_create.restype = ctypes.c_int
    L.sehf_scan_info.argtypes = [ctypes.c_void_p, _ip]
    L.sehf_scan_info.restype = ctypes.c_int
    L.sehf_scan_run.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.POINTER(ScanOut_C)]
    L.sehf_scan_run.restype = ctypes.c_int
    L.sehf_scan_run_tangent.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ScanOut_C), ctypes.POINTER(TangentOut_C)]
    L.sehf_scan_run_tangent.restype = ctypes.c_int
    L.sehf_scan_destroy.argtypes = [ctypes.c_void_p]
    L.sehf_scan_destroy.restype = None
    L.sehf_scan_create_with.argtypes = [ctypes.c_void_p, ctypes.POINTER(Problem_C), ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_void_p)]
    L.sehf_scan_create_with.restype = ctypes.c_int
    L.sehf_scan_set_checkpoints.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.sehf_scan_set_checkpoints.restype = ctypes.c_int
    L.sehf_scan_memory.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)]
    L.sehf_scan_memory.restype = ctypes.c_int
    L.sehf_plan_storage.argtypes = [ctypes.POINTER(PlanIn_C), ctypes.POINTER(Plan_C)]
    L.sehf_plan_storage.restype = ctypes.c_int
    L.sehf_scan_set_psi_store.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
    L.sehf_scan_set_psi_store.restype = ctypes.c_int
    L.sehf_scan_dipole_only.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _dp, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ScanOut_C)]
    L.sehf_scan_dipole_only.restype = ctypes.c_int
    L._scan_registered = True
    return L


def _c_problem(p, keep):
    def arr(x, dt=np.float64):
        y = np.ascontiguousarray(x, dtype=dt)
        keep.append(y)
        return y
    Z = arr(p.Z, np.int32)
    xyz = arr(p.xyz)
    vh = arr(p.V_H)
    psi = arr(p.psi_lm)
    init = arr(p.init_psi_lm) if p.init_psi_lm is not None else None
    dp = lambda a: a.ctypes.data_as(_dp) if a is not None else None  # noqa: E731
    return Problem_C(p.l_cont, p.n_occ, p.Nr, p.dr, len(Z), Z.ctypes.data_as(_ip), dp(xyz), vh.shape[0],
                     dp(vh), psi.shape[0], psi.shape[1], dp(psi), 1 if init is not None else 0, dp(init),
                     p.pol_kind, p.pol_alpha, p.pol_rc, 1 if p.allow_closed else 0)


class Scan:
    """Energy-independent device state for one molecule (sehf_scan handle)."""

    def __init__(self, problem, ctx=None, tables: str | None = None):
        """tables: None (sehf_scan_create: resident unless $SEHF_TABLES=streamed),
        "resident" or "streamed" (sehf_scan_create_with)."""
        from .steppers import GpuContext
        L = _register(load())
        self.ctx = ctx or GpuContext(True)
        self.p = problem
        keep = []
        cp = _c_problem(problem, keep)
        h = ctypes.c_void_p()
        if tables is None:
            check(L.sehf_scan_create(self.ctx._h, ctypes.byref(cp), ctypes.byref(h)), "sehf_scan_create")
        else:
            check(L.sehf_scan_create_with(self.ctx._h, ctypes.byref(cp), TABLES[tables], ctypes.byref(h)),
                  "sehf_scan_create_with")
        self._h = h
        info = (ctypes.c_int * 8)()
        check(L.sehf_scan_info(h, info))
        self.info = dict(zip(("n_psi", "n_f", "n_tot", "n_transition", "n_start", "n_pts", "Nr", "l_exch"), list(info)))

    def run(self, ks, batch: int = 0, seg_len: int = 0, groups: int = 0, tangent=None):
        """Solve every k.  `batch` energies advance through the radial grid
        together; `groups` (1..4) batches run concurrently on their own
        streams (0 keeps the handle's current setting).

        tangent: None, one direction (t_E, t_alpha) for every energy, or an
        [nk, 2] array of directions.  Each result then also holds the
        derivatives dK, dA, dB, db_overlap, dd_raw and dD along its direction
        (sehf_scan_run_tangent)."""
        L = load()
        if groups:
            check(L.sehf_scan_set_groups(self._h, int(groups)), "sehf_scan_set_groups")
        ks = np.ascontiguousarray(ks, dtype=np.float64)
        nk, N, no = len(ks), self.p.n_psi, self.p.n_occ
        o, out = self._outputs(nk)
        if tangent is None:
            check(L.sehf_scan_run(self._h, ks.ctypes.data_as(_dp), nk, batch, seg_len, ctypes.byref(out)),
                  "sehf_scan_run")
        else:
            d = np.asarray(tangent, dtype=np.float64)
            d = np.ascontiguousarray(np.broadcast_to(d, (nk, 2)) if d.ndim == 1 else d)
            if d.shape != (nk, 2):
                raise ValueError(f"tangent directions must have shape ({nk}, 2), got {d.shape}")
            o.update({"dK": np.zeros((nk, N, N)), "dA": np.zeros((nk, N, N)), "dB": np.zeros((nk, N, N)),
                      "db_overlap": np.zeros((nk, no, N)), "dd_raw": np.zeros((nk, 6, N)),
                      "dD": np.zeros((nk, 6, N), dtype=np.complex128)})
            tout = TangentOut_C(*[o[n].ctypes.data_as(_dp) for n in ("dK", "dA", "dB", "db_overlap", "dd_raw", "dD")])
            check(L.sehf_scan_run_tangent(self._h, ks.ctypes.data_as(_dp), d.ctypes.data_as(_dp), nk, batch,
                                          seg_len, ctypes.byref(out), ctypes.byref(tout)), "sehf_scan_run_tangent")
        return self._results(o, ks, tangent is not None)

    def _outputs(self, nk):
        N, no = self.p.n_psi, self.p.n_occ
        o = {"K": np.zeros((nk, N, N)), "A": np.zeros((nk, N, N)), "B": np.zeros((nk, N, N)),
             "K_fit": np.zeros((nk, N, N)), "b_overlap": np.zeros((nk, no, N)), "d_raw": np.zeros((nk, 6, N)),
             "D_raw": np.zeros((nk, 6, N), dtype=np.complex128), "D": np.zeros((nk, 6, N), dtype=np.complex128),
             "d_corr": np.zeros((6, no)), "diag": np.zeros((nk, 16))}
        out = ScanOut_C(*[o[n].ctypes.data_as(_dp) for n in ("K", "A", "B", "K_fit", "b_overlap", "d_raw",
                                                             "D_raw", "D", "d_corr", "diag")])
        return o, out

    @staticmethod
    def _results(o, ks, tangent):
        res = []
        for e in range(len(ks)):
            # matrices come back column-major per energy
            r = {"k": float(ks[e]), "K": o["K"][e].T.copy(), "A": o["A"][e].T.copy(),
                 "B": o["B"][e].T.copy(), "K_fit": o["K_fit"][e].T.copy(),
                 "b_overlap": o["b_overlap"][e].T.copy(), "d_raw": o["d_raw"][e], "D_raw": o["D_raw"][e],
                 "D": o["D"][e], "d_corr": o["d_corr"], "diag": o["diag"][e]}
            if tangent:
                r.update({"dK": o["dK"][e].T.copy(), "dA": o["dA"][e].T.copy(), "dB": o["dB"][e].T.copy(),
                          "db_overlap": o["db_overlap"][e].T.copy(), "dd_raw": o["dd_raw"][e], "dD": o["dD"][e]})
            res.append(r)
        return res

    PSI_TAIL = 1   # SEHF_PSI_TAIL
    PSI_MAIN = 2   # SEHF_PSI_MAIN

    def set_psi_store(self, root, keep: int = 3):
        """Later whole scans write per energy <root>/psi_ik<NNNN>/ with the
        300-point asymptotic tail of psi (psi_asym.bin, keep & PSI_TAIL) and
        the dipole range [0, n_pts) (psi_main.bin, keep & PSI_MAIN); None
        turns it off (sehf_scan_set_psi_store).  `sehf.psi_files` reads them."""
        r = None if root is None else os.fsencode(os.fspath(root))
        check(load().sehf_scan_set_psi_store(self._h, r, int(keep) if r is not None else 0),
              "sehf_scan_set_psi_store")

    def dipole_only(self, ks, root, batch: int = 0):
        """The fit and the dipoles of energies ks from the psi files under
        root (written by a scan with set_psi_store over the same k list), with
        this handle's dipole tables; "K" is not formed and stays zero
        (sehf_scan_dipole_only)."""
        ks = np.ascontiguousarray(ks, dtype=np.float64)
        o, out = self._outputs(len(ks))
        check(load().sehf_scan_dipole_only(self._h, os.fsencode(os.fspath(root)), ks.ctypes.data_as(_dp), len(ks),
                                           int(batch), ctypes.byref(out)), "sehf_scan_dipole_only")
        return self._results(o, ks, False)

    def set_checkpoints(self, where: str):
        """Ratio-matrix checkpoints in "device" memory or pinned "host"
        memory (sehf_scan_set_checkpoints)."""
        check(load().sehf_scan_set_checkpoints(self._h, CHECKPOINTS[where]), "sehf_scan_set_checkpoints")

    def memory(self) -> dict:
        """Bytes held: energy-independent device tables, batch buffers, pinned
        host checkpoints (sehf_scan_memory)."""
        b = (ctypes.c_size_t * 3)()
        check(load().sehf_scan_memory(self._h, b), "sehf_scan_memory")
        return {"shared": b[0], "batch": b[1], "host": b[2]}

    DIPOLES_FORWARD = 0    # SEHF_DIPOLES_FORWARD
    DIPOLES_BACKWARD = 1   # SEHF_DIPOLES_BACKWARD

    def set_dipole_path(self, path: int):
        """How plain runs sum the dipole integrals: DIPOLES_FORWARD
        (default) carries the adjoint row block Ad_n = (Ad_{n-1} + rows_n P_n)
        Rinv_n along the forward recursion and back-propagates only over the
        fit window; DIPOLES_BACKWARD back-propagates over the whole grid.
        Tangent runs always back-propagate (sehf_scan_set_dipole_path)."""
        check(load().sehf_scan_set_dipole_path(self._h, int(path)), "sehf_scan_set_dipole_path")

    def layout(self, seg_len: int = 0) -> dict:
        """Segment layout of the radial grid: L points per segment, the number
        of segments, the fit window's first point and the first segment of
        the fit pre-pass."""
        lay = (ctypes.c_int * 4)()
        check(load().sehf_scan_run_segment(self._h, None, None, 0, -1, int(seg_len), None, lay),
              "sehf_scan_run_segment")
        return dict(zip(("L", "n_seg", "n_fit_lo", "prepass_seg_lo"), list(lay)))

    def run_segment(self, ks, seg: int, seg_len: int = 0, tangent=None) -> dict:
        """One segment cycle (forward part, recompute, back-propagation of
        radial segment `seg`) for all ks in one batch; returns its timing
        (sehf_scan_run_segment)."""
        ks = np.ascontiguousarray(ks, dtype=np.float64)
        nk = len(ks)
        d = None
        if tangent is not None:
            d = np.asarray(tangent, dtype=np.float64)
            d = np.ascontiguousarray(np.broadcast_to(d, (nk, 2)) if d.ndim == 1 else d)
        t = (ctypes.c_double * 4)()
        check(load().sehf_scan_run_segment(self._h, ks.ctypes.data_as(_dp), d.ctypes.data_as(_dp) if d is not None
                                           else None, nk, int(seg), int(seg_len), t, None), "sehf_scan_run_segment")
        return dict(zip(("host_tables", "forward", "recompute", "back"), list(t)))

    def autotune(self, ks, probe_steps: int = 8, tangent: bool = False) -> dict:
        """Pick the batch width, group count and GEMM tile policy for these
        energies by timing probe_steps radial steps per candidate
        (sehf_scan_autotune); the handle keeps the chosen group count and the
        process the chosen tile policy."""
        L = load()
        ks = np.ascontiguousarray(ks, dtype=np.float64)
        b, g, t, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        rep = (ctypes.c_double * 128)()
        check(L.sehf_scan_autotune(self._h, ks.ctypes.data_as(_dp), len(ks), int(probe_steps), 1 if tangent else 0,
                                   ctypes.byref(b), ctypes.byref(g), ctypes.byref(t), rep, ctypes.byref(n)),
              "sehf_scan_autotune")
        table = [{"batch": int(rep[4 * i]), "groups": int(rep[4 * i + 1]), "tiles": int(rep[4 * i + 2]),
                  "energy_steps_per_s": rep[4 * i + 3]} for i in range(n.value)]
        return {"batch": b.value, "groups": g.value, "tiles": t.value, "probes": table}

    def stats(self) -> dict:
        t = (ctypes.c_double * 8)()
        check(load().sehf_scan_stats(self._h, t))
        return dict(zip(("setup", "forward", "kmatrix", "backprop", "fit_dipole", "host_tables"), list(t)[:6]))

    def close(self):
        """Release the device tables and batch buffers (sehf_scan_destroy)."""
        if getattr(self, "_h", None):
            load().sehf_scan_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def solve_scan(problem, ks, batch: int = 0, seg_len: int = 0, ctx=None):
    return Scan(problem, ctx).run(ks, batch=batch, seg_len=seg_len)
