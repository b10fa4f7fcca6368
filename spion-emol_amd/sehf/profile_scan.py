"""Profiling harness: one batch of the scan without torch, for hardware-counter
runs (rocprofv3 --pmc) where a framework's own kernels and allocator would
only add noise.

    python -m sehf.profile_scan --config h2o_l30 --rmax 3.2 --batch 32 --calib-mib 1024

runs, in order: a device-to-device copy of --calib-mib MiB through the HIP
runtime (the known-size reference for the byte counters: it reads and writes
exactly that many bytes), then one scan batch of --batch energies on the
configuration's channels with the radial grid cut to --rmax (every radial
step launches the same kernels with the same shapes as the full grid).
Progress lines go to stderr.
"""
from __future__ import annotations

import argparse
import ctypes
import sys
import time

import numpy as np


def _hip():
    L = ctypes.CDLL("libamdhip64.so")
    L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    L.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    L.hipFree.argtypes = [ctypes.c_void_p]
    return L


def calibration_copy(mib: int) -> None:
    L = _hip()
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if L.hipMalloc(ctypes.byref(p), n) != 0:
            raise RuntimeError("hipMalloc failed")
    if L.hipMemset(a, 1, n) != 0 or L.hipDeviceSynchronize() != 0:
        raise RuntimeError("hipMemset failed")
    if L.hipMemcpy(b, a, n, 3) != 0 or L.hipDeviceSynchronize() != 0:   # 3 = device to device
        raise RuntimeError("hipMemcpy failed")
    L.hipFree(a)
    L.hipFree(b)


def main(argv=None) -> int:
    from . import scan, synth
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="h2o_l30")
    ap.add_argument("--rmax", type=float, default=3.2)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--calib-mib", type=int, default=1024)
    ap.add_argument("--tangent", action="store_true")
    a = ap.parse_args(argv)
    if a.calib_mib > 0:
        calibration_copy(a.calib_mib)
        print(f"[profile_scan] calibration copy: {a.calib_mib} MiB read + {a.calib_mib} MiB written",
              file=sys.stderr, flush=True)
    p, ks = synth.config(a.config)
    q = synth.h2o_like(p.l_cont, p.dr, a.rmax, p.n_occ, seed=1, allow_closed=True,
                       pol=(p.pol_alpha, p.pol_rc) if p.pol_kind == 2 else None)
    S = scan.Scan(q)
    kb = ks[(np.arange(a.batch) * 37) % len(ks)]
    t0 = time.time()
    res = S.run(kb, batch=a.batch, tangent=(1.0, 0.0) if a.tangent else None)
    ok = all(np.all(np.isfinite(r["D"])) for r in res)
    print(f"[profile_scan] {a.config} channels, Nr={q.Nr}, batch {a.batch}: {time.time() - t0:.2f} s, "
          f"finite={ok}", file=sys.stderr, flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
