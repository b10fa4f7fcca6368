#!/usr/bin/env python
"""Benchmark of the batched static-exchange photoionization scan on one MI355X.

A step is one pass of the whole energy-dependent path -- Schur inverses,
forward ratio recursion, K matrix, back-propagation, asymptotic fit and the
six dipole slices -- over one batch of energies of the l_cont = 30 H2O-like
scan (961 channels, N_total = 1566, 6001 radial points, k in [0.05, 2.0],
closed channels allowed).  The metric is energy points per second with the
energy-independent inputs already resident in HBM.

Prints one JSON line (rank 0).  `roofline` is for the dominant kernel, the
FP64 MFMA GEMM (every Gauss-Jordan block update and every back-propagation
product): algorithmic flop of the sampled launches over their summed
durations, timed with HIP events on the launch stream inside the timed steps.
`cpu_baseline` times the CPU oracle (the restatement of the reference's FP64
CPU path) on a bounded sample of radial steps and extrapolates to one energy.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "spion-emol_amd"))

PEAK_FP64_MFMA_TFLOPS = 78.6      # MI355X dense FP64 matrix peak (datasheet)

# HBM bytes of one k_dgemm launch of the top-level Gauss-Jordan update of the
# order-1566 inverse (783 x 1566 x 783, the largest single share of GEMM time
# at l_cont = 30), measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE on
# spion-emol_amd/_build/gemm_probe and calibrated by a 1 GiB device copy
# (profiles/k_dgemm_hbm_traffic_gemm_probe_b32.md).  batch -> (measured bytes,
# algorithmic bytes).  Counters cannot be collected inside the scan itself.
TOP_UPDATE_TRAFFIC = {32: (1.7831e+09, 1.0987e+09), 48: (2.7581e+09, 1.6480e+09)}


def model_flops_per_energy(Np, Nf, Nt, Nr, n_tr, n_start, n_occ):
    """Algorithmic FP64 flop of one energy point (the path's flop model):
    Sinv: sum_n 2Np^3 + [n<n_tr] 2Np^2 Nf
    FRP : sum_{n>=n_start} 2Nt^3 + [n<n_tr] 2 Np Nt (2Nf + Np)
    BP  : sum_n 2Nt^2 Np + 2Np^3 + [n<n_tr] 4 Np^2 Nf
    dipole: n_tr (6 + n_occ) 2 Np^2 ;  K, fit, M: 10 Nt^3."""
    ex = min(n_tr, Nr)
    sinv = Nr * 2 * Np ** 3 + ex * 2 * Np ** 2 * Nf
    frp = (Nr - n_start) * 2 * Nt ** 3 + max(0, ex - n_start) * 2 * Np * Nt * (2 * Nf + Np)
    bp = Nr * (2 * Nt ** 2 * Np + 2 * Np ** 3) + ex * 4 * Np ** 2 * Nf
    dip = ex * (6 + n_occ) * 2 * Np ** 2
    return float(sinv + frp + bp + dip + 10 * Nt ** 3)


def tangent_flops_per_energy(Np, Nf, Nt, Nr, n_tr, n_start, n_occ):
    """Algorithmic FP64 flop the forward-mode derivative adds per energy (the
    forward model's counting: no credit for the back-propagation recompute):
    ratio steps: sum_{n>=n_start} G G^T 2Nt^2 Np + two products 4Nt^3
    back-propagation: sum_n dRinv Z and Rinv(..) 4Nt^2 Np, Sinv psi and the
        W^{-1} apply of dZ 4Np^3, [n<n_tr] 4Np^2 Nf ; dipole: n_tr (6+n_occ) 2Np^2."""
    ex = min(n_tr, Nr)
    fwd = (Nr - n_start) * (2 * Nt ** 2 * Np + 4 * Nt ** 3)
    bp = Nr * (4 * Nt ** 2 * Np + 4 * Np ** 3) + ex * 4 * Np ** 2 * Nf
    return float(fwd + bp + ex * (6 + n_occ) * 2 * Np ** 2)


def cpu_baseline(p, ks, n_steps):
    """The CPU oracle (FP64 restatement of the reference CPU path) on the
    host: one energy-independent setup, then `cores` energies concurrently
    (one thread each) run n_steps radial steps of Sinv + FRP + BP at full
    size; extrapolated to Nr steps per energy."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O
    cores = max(1, min(os.cpu_count() or 1, 16))
    sample = np.asarray(ks[:cores], dtype=np.float64)
    t, wall = O.time_steps_multi(p, sample, n_steps)
    step_s = float(np.mean(t.sum(axis=1)))
    energy_s = step_s * p.Nr
    return {"value": cores / energy_s, "unit": "energy points/s", "cores": cores, "kind": "port",
            "sample": f"{n_steps} radial step(s) of Sinv+FRP+BP at full size (N_tot={p.n_tot}) for {cores} "
                      f"energies run concurrently, one thread each ({wall[1]:.1f} s wall after a {wall[0]:.1f} s "
                      f"energy-independent setup); {step_s:.2f} s per step per energy x {p.Nr} steps = "
                      f"{energy_s:.0f} s per energy (extrapolated)"}


def side_measurements(args, p, ks_all, B):
    """Short-grid ratios (same channel sizes, r_max = args.side_rmax; per-step
    costs do not depend on the grid length):
    * batched vs the plain per-ik path on this configuration: B energies in
      one batch against one energy alone (batch 1);
    * the forward-mode derivative's cost on the l_cont = 20 polarisation
      configuration: the same B energies with and without the tangent."""
    import torch
    from sehf import scan as sc, synth
    out = {"r_max": args.side_rmax}
    q = synth.h2o_like(p.l_cont, p.dr, args.side_rmax, p.n_occ, seed=1, allow_closed=True)
    S = sc.Scan(q)
    ks = ks_all[(np.arange(B) * 37) % len(ks_all)]

    def timed(f):
        torch.cuda.synchronize()
        t0 = time.time()
        f()
        torch.cuda.synchronize()
        return time.time() - t0

    S.run(ks[:1], batch=1)                                    # warm the per-ik buffers
    t_b = timed(lambda: S.run(ks, batch=B, groups=1))
    t_1 = timed(lambda: S.run(ks[:1], batch=1, groups=1))
    out["per_ik"] = {"workload": f"{args.config} channels, Nr={q.Nr}", "batched_energies_per_s": B / t_b,
                     "per_ik_energies_per_s": 1.0 / t_1, "batched_over_per_ik": (B / t_b) * t_1}
    print(f"[bench] side: batched {B / t_b:.3f} vs per-ik {1.0 / t_1:.3f} energies/s on Nr={q.Nr}",
          file=sys.stderr, flush=True)
    S.close()
    pc, kc = synth.config("h2o_l20_pol")
    q = synth.h2o_like(pc.l_cont, pc.dr, args.side_rmax, pc.n_occ, seed=1, pol=(pc.pol_alpha, pc.pol_rc),
                       allow_closed=True)
    S = sc.Scan(q)
    ks = kc[(np.arange(B) * 37) % len(kc)]
    S.run(ks, batch=B, tangent=(1.0, 0.0))                   # allocate the tangent buffers
    t_f = timed(lambda: S.run(ks, batch=B))
    t_t = timed(lambda: S.run(ks, batch=B, tangent=(1.0, 0.0)))
    out["tangent"] = {"workload": f"h2o_l20_pol channels (N_tot={q.n_tot}), Nr={q.Nr}, batch {B}",
                      "forward_energies_per_s": B / t_f, "with_tangent_energies_per_s": B / t_t,
                      "tangent_over_forward_cost": t_t / t_f}
    print(f"[bench] side: tangent/forward cost {t_t / t_f:.2f} on Nr={q.Nr}", file=sys.stderr, flush=True)
    S.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--batch", type=int, default=0,
                    help="energies per batch; 0 = batch width, groups and GEMM tiles chosen by the library's "
                         "autotuner (timed probes of radial steps, outside the timed region)")
    ap.add_argument("--groups", type=int, default=1,
                    help="concurrent energy groups (streams); a step is groups x batch energies")
    ap.add_argument("--config", default="h2o_l30")
    ap.add_argument("--seg-len", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=4, help="radial steps in the CPU oracle sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--prof-every", type=int, default=97)
    ap.add_argument("--per-ik", action="store_true",
                    help="also time the plain per-ik path (batch 1, one group) on one energy of the full grid")

    ap.add_argument("--tangent", action="store_true",
                    help="time the batched scan with its forward-mode derivative along E (batched + adjoint)")
    ap.add_argument("--side-rmax", type=float, default=5.0,
                    help="radial extent of the short-grid side measurements (per-ik vs batched on this config; "
                         "tangent vs plain on the l_cont=20 polarisation config); 0 skips them")
    ap.add_argument("--rmax", type=float, default=0.0,
                    help="profiling aid: shorten the radial grid (same channel sizes); not the metric workload")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="profiling aid: one device-to-device copy of this many MiB before the timed step, "
                         "the known-size reference for HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only; --gpus above 1 is not supported "
              "(the design shards energies across GPUs in contiguous k blocks; no collective is needed)", file=sys.stderr)
        sys.exit(2)

    import ctypes
    import torch
    from sehf import scan as sc, synth
    from sehf._lib import load

    torch.cuda.set_device(0)
    t0 = time.time()
    p, ks_all = synth.config(args.config)
    if args.rmax > 0:
        p = synth.h2o_like(p.l_cont, p.dr, args.rmax, p.n_occ, seed=1, allow_closed=True)
    t_gen = time.time() - t0
    print(f"[bench] {args.config}: N_psi={p.n_psi} N_f={p.n_f} N_tot={p.n_tot} Nr={p.Nr} "
          f"(inputs generated in {t_gen:.1f} s)", file=sys.stderr, flush=True)
    t0 = time.time()
    S = sc.Scan(p)
    t_setup = time.time() - t0
    info = S.info
    print(f"[bench] setup (V(r), exchange, dipole tables on device) {t_setup:.1f} s; info {info}",
          file=sys.stderr, flush=True)

    B = args.batch
    G = max(1, args.groups)
    tune = None
    if B <= 0:
        t0 = time.time()
        tune = S.autotune(ks_all[(np.arange(min(128, len(ks_all))) * 37) % len(ks_all)], probe_steps=8,
                          tangent=args.tangent)
        B, G = tune["batch"], tune["groups"]
        tune["seconds"] = time.time() - t0
        print(f"[bench] autotune: batch {B}, groups {G} ({tune['seconds']:.1f} s)", file=sys.stderr, flush=True)
    E_step = B * G
    nk = len(ks_all)

    def batch_k(i):
        # a spread of the scan's k values (closed and open channels mixed)
        idx = [((i * E_step + j) * 37) % nk for j in range(E_step)]
        return ks_all[idx]

    L = load()
    if args.calib_copy_mib > 0:
        n = args.calib_copy_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        print(f"[bench] calibration copy: {args.calib_copy_mib} MiB read + {args.calib_copy_mib} MiB written",
              file=sys.stderr, flush=True)
    tan = (1.0, 0.0) if args.tangent else None
    for w in range(args.warmup):
        S.run(batch_k(w), batch=B, seg_len=args.seg_len, groups=G, tangent=tan)
        print(f"[bench] warmup step {w} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    L.sehf_prof_begin(args.prof_every)
    t0 = time.time()
    outs = []
    for i in range(args.steps):
        outs.append(S.run(batch_k(args.warmup + i), batch=B, seg_len=args.seg_len, groups=G, tangent=tan))
        print(f"[bench] timed step {i} done ({time.time() - t0:.1f} s)", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.time() - t0
    prof = (ctypes.c_double * 16)()
    L.sehf_prof_end(prof)
    prof = list(prof)
    energies = E_step * args.steps
    value = energies / elapsed
    mflop = model_flops_per_energy(info["n_psi"], info["n_f"], info["n_tot"], info["Nr"],
                                   info["n_transition"], info["n_start"], p.n_occ)
    if args.tangent:
        mflop += tangent_flops_per_energy(info["n_psi"], info["n_f"], info["n_tot"], info["Nr"],
                                          info["n_transition"], info["n_start"], p.n_occ)
    gemm_tflops = prof[1] / (prof[0] * 1e-3) / 1e12 if prof[0] > 0 else 0.0
    avg_launch_us = prof[0] / max(prof[2], 1) * 1e3
    buckets = {}
    for i, nm in enumerate(("gemm_all", "gj_leaf", "gemm_K_lt_64", "gemm_K_64_256", "gemm_K_ge_256")):
        ms, fl, n = prof[3 * i: 3 * i + 3]
        if n:
            buckets[nm] = {"sampled": int(n), "avg_us": ms / n * 1e3, "tflops": fl / (ms * 1e-3) / 1e12,
                           "est_total_s": ms * args.prof_every / 1e3}
    stages = S.stats()
    for o in outs:
        for r in o:
            if not np.all(np.isfinite(r["D"])) or (args.tangent and not np.all(np.isfinite(r["dD"]))):
                raise SystemExit("non-finite dipole in bench output")

    traffic, traffic_note = None, None
    if args.config == "h2o_l30" and args.rmax == 0 and B in TOP_UPDATE_TRAFFIC:
        tb, ta = TOP_UPDATE_TRAFFIC[B]
        traffic = tb
        traffic_note = (f"HBM bytes of the 783x1566x783 Gauss-Jordan update launch at batch {B} "
                        f"(rocprofv3 --pmc on gemm_probe, calibrated); algorithmic {ta:.3e} B, ratio {tb / ta:.2f}")
    line = {
        "metric": "energy points/sec @l_cont=30 (batched vs per-ik, 1 GPU); FP64-MFMA %; adjoint/fwd cost",
        "value": value, "unit": "energy points/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{args.config}: H2O-like, l_cont={p.l_cont} (N_psi={p.n_psi}, N_tot={p.n_tot}), "
                               f"Nr={p.Nr}, dr={p.dr}, k in [{ks_all[0]}, {ks_all[-1]}] ({nk} pts), "
                               f"batch {B} energies per step, closed channels allowed"
                               + (", with the forward-mode derivative along E (batched + adjoint)" if args.tangent
                                  else ""),
                   "batch_energies": B, "groups": G, "energies_per_step": E_step, "tangent": bool(args.tangent),
                   "seg_len": args.seg_len or "ceil(sqrt(Nr))"},
        "roofline": {"bound": "mfma", "achieved": gemm_tflops, "peak": PEAK_FP64_MFMA_TFLOPS, "unit": "TFLOP/s",
                     "frac": gemm_tflops / PEAK_FP64_MFMA_TFLOPS, "traffic": traffic,
                     "traffic_launch": traffic_note,
                     "kernel": "k_dgemm", "sampled_launches": int(prof[2]),
                     "avg_launch_us": avg_launch_us, "event_pair_overhead_us": prof[15] * 1e3},
        "model_tflop_per_energy": mflop / 1e12,
        "path_tflops": mflop * energies / elapsed / 1e12,
        "path_frac_of_fp64_peak": mflop * energies / elapsed / 1e12 / PEAK_FP64_MFMA_TFLOPS,
        "stage_seconds": stages, "setup_seconds": t_setup,
        "sampled_kernels": buckets, "prof_every": args.prof_every,
    }
    if args.rmax > 0:
        line["config"]["workload"] += f" [profiling grid r_max={args.rmax}, not the metric workload]"
    if args.per_ik:
        t0 = time.time()
        S.run(ks_all[nk // 2: nk // 2 + 1], batch=1, seg_len=args.seg_len, groups=1)
        torch.cuda.synchronize()
        t1 = time.time() - t0
        line["per_ik"] = {"value": 1.0 / t1, "unit": "energy points/s", "seconds": t1,
                          "batched_over_per_ik": value * t1}
    if tune is not None:
        line["autotune"] = tune
    S.close()   # the side measurements build their own device tables
    if args.side_rmax > 0 and args.rmax == 0:
        line["side"] = side_measurements(args, p, ks_all, B)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(p, batch_k(0), args.cpu_steps)
        line["gpu_over_cpu"] = value / line["cpu_baseline"]["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
