#!/usr/bin/env python
"""Benchmark of the batched static-exchange photoionization scan on one MI355X.

Workload: the l_cont = 30 H2O-like scan (961 channels, N_total = 1566, 6001
radial points at dr = 0.01, k in [0.05, 2.0], closed channels allowed), with
the energy-independent inputs resident in HBM.

A scan takes each batch of B energies through the forward ratio recursion
over the whole radial grid (Schur inverses one step ahead on a side stream),
carrying the dipole integrals forward as an adjoint row block
(sehf_scan_set_dipole_path, SEHF_DIPOLES_FORWARD); then, over the top
segments only, two passes of recompute (the segment's ratio matrices from its
checkpoint) and back-propagation down to the fit window: the fit sums, then
the fit residual; plus per-batch work (host tables, K matrix, fit solve,
dipole normalisation).  With --dipole-path backward the dipoles are instead
back-propagated over the whole grid: recompute and back-propagation of every
segment after one fit-sum pass over the top ones.

A timed step is one segment cycle at full size (sehf_scan_run_segment): the
B energies run segment s of the forward recursion, its recompute and its
back-propagation -- L = ceil(sqrt(Nr-1)) = 78 radial points of each.  The
timed steps visit segments spread evenly over the grid (inside and beyond
the exchange cut-off), so their mean cost per radial point is the grid's.
`value` (energy points per second) is B over the modelled time of one whole
batch: the measured per-point forward cost times the grid's Nr-1 points, plus
the top-segment passes from the measured per-point recompute and
back-propagation costs, plus the per-batch fixed stages timed on a full
short-grid scan of the same channel sizes.  --full-scan also runs one whole
batch through the full grid and reports its measured rate beside the model.

Prints one JSON line.  `roofline` is for the dominant kernel, the FP64 MFMA
GEMM k_dgemm (every Gauss-Jordan block update and every back-propagation
product): the algorithmic flop of the sampled launches over their summed
durations, timed with HIP events on each launch's own stream inside the
timed steps.  `cpu_baseline` times the CPU oracle (the restatement of the
reference's FP64 CPU path) on a bounded sample of radial steps.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "spion-emol_amd"))

PEAK_FP64_MFMA_TFLOPS = 78.6      # MI355X dense FP64 matrix peak (datasheet)
METRIC = "energy points/sec @l_cont=30 (batched vs per-ik, 1 GPU); FP64-MFMA %; adjoint/fwd cost"

# HBM bytes of the main k_dgemm launch (768 rows, 128 x 128 tiles) of the
# top-level Gauss-Jordan update of the order-1566 inverse (783 x 1566 x 783,
# the largest single share of GEMM time at l_cont = 30): rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE per dispatch over `bench.py --kernels gemm`,
# scaled by the same counters on a 1 GiB device copy (FETCH x 2, WRITE x 1).
# batch -> (measured bytes, algorithmic bytes).  From separate counter runs of
# the kernel table at the scan's shapes, not from this process.
TOP_UPDATE_TRAFFIC = {64: (3.416e+09, 2.167e+09)}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def model_flops_per_energy(Np, Nf, Nt, Nr, n_tr, n_start, n_occ, n_pts, n_back):
    """Algorithmic FP64 flop of one energy point (the path's flop model):
    Sinv: sum_n 2Np^3 + [n<n_tr] 2Np^2 Nf
    FRP : sum_{n>=n_start} 2Nt^3 + [n<n_tr] 2 Np Nt (2Nf + Np)
    BP  : over the n_back back-propagated points 2Nt^2 Np + 2Np^3 + [n<n_tr] 4 Np^2 Nf
    dipole rows: sum_{n<n_pts} (6 + n_occ) 2 Np (Np + Nf), and with the
        forward-accumulated dipoles (n_back < Nr) the row block times Rinv_n,
        sum_n (6 + n_occ) 2 Nt^2;  K, fit, M: 10 Nt^3.
    Recomputed ratio matrices and the second (residual) pass over the fit
    window are not counted."""
    ex = min(n_tr, Nr)
    ns = 6 + n_occ
    sinv = Nr * 2 * Np ** 3 + ex * 2 * Np ** 2 * Nf
    frp = (Nr - n_start) * 2 * Nt ** 3 + max(0, ex - n_start) * 2 * Np * Nt * (2 * Nf + Np)
    bp = n_back * (2 * Nt ** 2 * Np + 2 * Np ** 3) + min(ex, n_back) * 4 * Np ** 2 * Nf
    dip = n_pts * ns * 2 * Np * (Np + Nf)
    if n_back < Nr:
        dip += Nr * ns * 2 * Nt ** 2
    return float(sinv + frp + bp + dip + 10 * Nt ** 3)


def spread_segments(n_seg, count, offset=0.5):
    """`count` segment indices spread evenly over [0, n_seg)."""
    return [min(n_seg - 1, int((i + offset) * n_seg / count)) for i in range(count)]


def seg_points(lay, N, s):
    L = lay["L"]
    return min((s + 1) * L, N) - s * L


def batch_seconds(lay, N, pts, t_fwd, t_rec, t_back, t_fixed, forward_dipoles=True):
    """Modelled seconds of one whole batch from segment-cycle timings.
    Forward-accumulated dipoles: the forward sweep over the N = Nr-1 points
    and two passes over the top segments (recompute from the checkpoint of
    the segment holding the fit window's first point, back-propagation down
    to that point): fit sums, then fit residual.  Back-propagated dipoles:
    all three sweeps over the N points plus one such pass (the fit sums).
    Plus the fixed per-batch stages.  The back-propagation of the top passes
    is priced at the cycle's (which also forms the dipole integrands)."""
    L, n_seg, nfs, s_pre = lay["L"], lay["n_seg"], lay["n_fit_lo"], lay["prepass_seg_lo"]
    P = float(np.sum(pts))
    f, r, b = np.sum(t_fwd) / P, np.sum(t_rec) / P, np.sum(t_back) / P
    pre_rec = N - s_pre * L          # points recomputed by one top pass
    pre_back = N - nfs               # points back-propagated by one top pass
    if forward_dipoles:
        sweeps = N * f
        top = 2 * (pre_rec * r + pre_back * b)
    else:
        sweeps = N * (f + r + b)
        top = pre_rec * r + pre_back * b
    return sweeps + top + t_fixed, {"per_point_s": {"forward": f, "recompute": r, "back": b},
                                    "sweeps_s": sweeps, "top_passes_s": top, "fixed_s": t_fixed}


def fixed_stage_seconds(p, ks, B):
    """Per-batch fixed stages (K matrix, fit solve, dipole normalisation,
    download) from the stage timers of one whole scan of the same channel
    sizes on a short grid (r_max = 3.2); they do not depend on the grid
    length.  (The host tables are part of every timed segment cycle.)"""
    from sehf import scan as sc, synth
    q = synth.h2o_like(p.l_cont, p.dr, 3.2, p.n_occ, seed=1, allow_closed=True)
    S = sc.Scan(q)
    try:
        t0 = time.time()
        res = S.run(ks[:B], batch=B, groups=1)
        wall = time.time() - t0
        d = S.stats()
        if not all(np.all(np.isfinite(r["D"])) for r in res):
            raise SystemExit("non-finite dipole in the short-grid scan")
    finally:
        S.close()
    fixed = d["kmatrix"] + d["fit_dipole"]
    return fixed, {"short_grid_Nr": q.Nr, "scan_s": wall, "kmatrix_s": d["kmatrix"], "fit_dipole_s": d["fit_dipole"]}


def cycle_rate(S, p, lay, ks, segs, warm, tangent=None, seg_len=0):
    """Segment cycles of the energies ks at the listed segments (after `warm`
    untimed ones); returns the modelled batch seconds and the timings."""
    N = p.Nr - 1
    for s in warm:
        S.run_segment(ks, s, seg_len=seg_len, tangent=tangent)
    tf, tr, tb, pts, walls = [], [], [], [], []
    for s in segs:
        t0 = time.time()
        t = S.run_segment(ks, s, seg_len=seg_len, tangent=tangent)
        walls.append(time.time() - t0)
        tf.append(t["forward"] + t["host_tables"]); tr.append(t["recompute"]); tb.append(t["back"])
        pts.append(seg_points(lay, N, s))
    return tf, tr, tb, pts, walls


def kernel_table(L, B, calib_mib, parts="all"):
    """Rates of the path's inner kernels alone at the scan's shapes
    (sehf_time_gemm, sehf_time_inverse, sehf_time_spd_inverse) and, first,
    one device copy of calib_mib MiB, the known-size reference of HBM byte
    counters.  parts: "all", "gemm" (the product shapes only, so that counter
    passes see no other launch of the same grid) or "inverse"."""
    import ctypes
    from sehf._lib import check
    out = {"batch": B}
    ms = ctypes.c_double()
    if calib_mib > 0:
        check(L.sehf_copy_probe(ctypes.c_size_t(calib_mib << 20), ctypes.byref(ms)), "sehf_copy_probe")
        out["calibration_copy"] = {"MiB": calib_mib, "ms": ms.value, "GB_per_s": 2 * (calib_mib << 20) / (ms.value * 1e-3) / 1e9}
    n, npsi = 1566, 961
    # (name, M, N, K, lda, ldb, ldc, beta) of the l_cont = 30 scan's products
    shapes = [("Gauss-Jordan update 783x1566x783", 783, n, 783, 783, n, n, 1),
              ("Gauss-Jordan update 392x1566x391", 392, n, 391, 392, n, n, 1),
              ("Gauss-Jordan update 196x1566x196", 196, n, 196, 196, n, n, 1),
              ("Gauss-Jordan update 98x1566x98", 98, n, 98, 98, n, n, 1),
              ("Gauss-Jordan update 48x1566x48", 48, n, 48, 48, n, n, 1),
              ("back-propagation Rinv Z 1566x961x1566", n, npsi, n, n, n, n, 0),
              ("back-propagation Sinv rhs 961^3", npsi, npsi, npsi, npsi, npsi, npsi, 0)]
    rows = []
    for name, M, N, K, lda, ldb, ldc, beta in (shapes if parts in ("all", "gemm") else []):
        check(L.sehf_time_gemm(M, N, K, lda, ldb, ldc, B, beta, 3, ctypes.byref(ms)), "sehf_time_gemm")
        fl = 2.0 * M * N * K * B
        by = 8.0 * B * (M * K + K * N + (2 if beta else 1) * M * N)
        rows.append({"shape": name, "us": ms.value * 1e3, "tflops": fl / (ms.value * 1e-3) / 1e12,
                     "algorithmic_bytes": by})
        log(f"kernels: {name} batch {B}: {ms.value * 1e3:.1f} us, {rows[-1]['tflops']:.1f} TFLOP/s")
    out["gemm"] = rows
    if parts == "gemm":
        return out
    inv = []
    res = ctypes.c_double()
    for order in (n, npsi):
        for b in (8, 16, 32, 48, 64):
            check(L.sehf_time_inverse(order, b, 3, ctypes.byref(ms), ctypes.byref(res)), "sehf_time_inverse")
            inv.append({"n": order, "batch": b, "ms": ms.value, "ms_per_matrix": ms.value / b,
                        "tflops_2n3": 2.0 * order ** 3 * b / (ms.value * 1e-3) / 1e12, "resid": res.value})
            log(f"kernels: inverse n={order} batch {b}: {ms.value:.2f} ms ({inv[-1]['tflops_2n3']:.1f} TFLOP/s, "
                f"|AX-I| {res.value:.1e})")
    out["inverse"] = inv
    spd = []
    for b in (16, 48, 64):
        check(L.sehf_time_spd_inverse(npsi, b, 3, ctypes.byref(ms), ctypes.byref(res)), "sehf_time_spd_inverse")
        spd.append({"n": npsi, "batch": b, "ms": ms.value, "ms_per_matrix": ms.value / b,
                    "tflops_4n3_3": 4.0 / 3.0 * npsi ** 3 * b / (ms.value * 1e-3) / 1e12, "resid": res.value})
        log(f"kernels: SPD inverse n={npsi} batch {b}: {ms.value:.2f} ms ({spd[-1]['tflops_4n3_3']:.1f} TFLOP/s of "
            f"4/3 n^3, |AX-I| {res.value:.1e})")
    out["spd_inverse"] = spd
    return out


def cpu_info():
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count() or 1


def cpu_baseline(p, ks, n_steps):
    """The CPU oracle (FP64 restatement of the reference CPU path) on the
    host: one energy-independent setup, then `cores` energies concurrently
    (one thread each) run n_steps radial steps of Sinv + FRP + BP at full
    size; extrapolated to Nr steps per energy.  The reference keeps every
    ratio matrix (disk or RAM), so its energy is Nr steps of each stage with
    no recompute.  K, fit and dipole are O(N^3) once per energy (under 0.1%)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O
    model, nproc = cpu_info()
    cores = max(1, min(nproc, 16))
    sample = np.asarray(ks[:cores], dtype=np.float64)
    t, wall = O.time_steps_multi(p, sample, n_steps)
    per_stage = t.mean(axis=0)                     # seconds per radial step per energy
    step_s = float(per_stage.sum())
    energy_s = step_s * p.Nr
    rate_hoisted = cores / energy_s
    rate_per_ik_setup = cores / (energy_s + wall[0])
    return {"value": rate_hoisted, "unit": "energy points/s", "cores": cores, "kind": "port",
            "cpu_model": model, "nproc": nproc,
            "sample": f"{n_steps} radial step(s) of Sinv+FRP+BP at full size (N_tot={p.n_tot}) for {cores} "
                      f"energies run concurrently, one thread each ({wall[1]:.1f} s wall after a {wall[0]:.1f} s "
                      f"energy-independent setup); {step_s:.2f} s per step per energy x {p.Nr} steps = "
                      f"{energy_s:.0f} s per energy (extrapolated)",
            "stage_seconds_per_energy": {"sinv": per_stage[0] * p.Nr, "frp": per_stage[1] * p.Nr,
                                         "bp": per_stage[2] * p.Nr},
            "setup_seconds": wall[0],
            "setup_hoisted_value": rate_hoisted,
            "setup_per_ik_value": rate_per_ik_setup}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=0,
                    help="energies per batch; 0 = the device storage planner's choice (sehf_plan_storage: the "
                         "widest batch the HBM holds, at most 64); -1 = run the library's autotuner first (timed "
                         "probes, outside the timed region)")
    ap.add_argument("--tiles", type=int, default=0, help="GEMM tile policy (sehf_set_gemm_tiles); -1 = library default")
    ap.add_argument("--config", default="h2o_l30")
    ap.add_argument("--cpu-steps", type=int, default=4, help="radial steps in the CPU oracle sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-side", action="store_true", help="skip the per-ik, derivative and fixed-stage runs")
    ap.add_argument("--prof-every", type=int, default=97, help="HIP-event sampling of every n-th GEMM/leaf "
                                                                "launch; 0 = off")
    ap.add_argument("--kernels", nargs="?", const="all", default=None, choices=("all", "gemm", "inverse"),
                    help="instead of the scan: the inner kernels alone at the scan's shapes (GEMM table, inverse "
                         "scaling), after a calibration copy of --calib-mib MiB; 'gemm' or 'inverse' for one part")
    ap.add_argument("--calib-mib", type=int, default=1024)
    ap.add_argument("--dipole-path", choices=("forward", "backward"), default="forward",
                    help="forward: dipoles accumulated along the forward recursion (default); backward: "
                         "back-propagated over the whole grid")
    ap.add_argument("--full-scan", action="store_true",
                    help="also run one whole batch through the full grid (minutes) and report its measured rate "
                         "beside the modelled one")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only; --gpus above 1 is not supported "
              "(the design shards energies across GPUs in contiguous k blocks; no collective is needed)",
              file=sys.stderr)
        sys.exit(2)

    import ctypes
    from sehf import scan as sc, synth
    from sehf._lib import check, load

    L = load()
    if args.tiles >= 0:
        check(L.sehf_set_gemm_tiles(args.tiles), "sehf_set_gemm_tiles")
    if args.kernels:
        print(json.dumps({"kernels": kernel_table(L, max(args.batch, 1), args.calib_mib, args.kernels)}), flush=True)
        return
    t0 = time.time()
    p, ks_all = synth.config(args.config)
    log(f"{args.config}: N_psi={p.n_psi} N_f={p.n_f} N_tot={p.n_tot} Nr={p.Nr} "
        f"(inputs generated in {time.time() - t0:.1f} s)")
    fwd_dip = args.dipole_path == "forward"
    dpath = sc.Scan.DIPOLES_FORWARD if fwd_dip else sc.Scan.DIPOLES_BACKWARD
    nk = len(ks_all)
    plan = sc.plan_storage(p, nk, dipole_path=dpath)
    log("storage plan:\n" + plan["report"])
    t0 = time.time()
    S = sc.Scan(p, tables=plan["tables"])
    S.set_dipole_path(dpath)
    S.set_checkpoints(plan["checkpoints"])
    t_setup = time.time() - t0
    info = S.info
    log(f"setup (V(r), exchange, dipole tables on device) {t_setup:.1f} s; info {info}")
    B, tune = (plan["batch"] if args.batch == 0 else args.batch), None
    if B < 0:
        t0 = time.time()
        tune = S.autotune(ks_all[(np.arange(min(128, nk)) * 37) % nk], probe_steps=8)
        B = tune["batch"]
        tune["seconds"] = time.time() - t0
        log(f"autotune: batch {B}, groups {tune['groups']}, tiles {tune['tiles']} ({tune['seconds']:.1f} s)")
    # one batch: a spread of the scan's k values (closed and open channels mixed)
    ks = ks_all[(np.arange(B) * 37) % nk]
    L_plan = plan["seg_len"]
    lay = S.layout(L_plan)
    N = p.Nr - 1
    warm = spread_segments(lay["n_seg"], args.warmup, offset=0.25)
    segs = spread_segments(lay["n_seg"], args.steps)
    log(f"layout {lay}; warmup segments {warm}; timed segments {segs}")

    for i, s in enumerate(warm):
        S.run_segment(ks, s, seg_len=L_plan)
        log(f"warmup step {i} (segment {s}) done")
    if args.prof_every > 0:
        check(L.sehf_prof_begin(args.prof_every), "sehf_prof_begin")
    tf, tr, tb, pts, walls = [], [], [], [], []
    t_start = time.time()
    for i, s in enumerate(segs):
        t = S.run_segment(ks, s, seg_len=L_plan)
        tf.append(t["forward"] + t["host_tables"]); tr.append(t["recompute"]); tb.append(t["back"])
        pts.append(seg_points(lay, N, s))
        log(f"timed step {i} (segment {s}, {pts[-1]} points) {t['host_tables'] + t['forward'] + t['recompute'] + t['back']:.2f} s "
            f"({time.time() - t_start:.1f} s)")
    elapsed = time.time() - t_start
    prof = (ctypes.c_double * 16)()
    if args.prof_every > 0:
        check(L.sehf_prof_end(prof), "sehf_prof_end")
    prof = list(prof)

    gemm_tflops = prof[1] / (prof[0] * 1e-3) / 1e12 if prof[0] > 0 else 0.0
    avg_launch_us = prof[0] / max(prof[2], 1) * 1e3
    buckets = {}
    for i, nm in enumerate(("gemm_all", "gj_leaf", "gemm_K_lt_64", "gemm_K_64_256", "gemm_K_ge_256")):
        ms, fl, n = prof[3 * i: 3 * i + 3]
        if n:
            buckets[nm] = {"sampled": int(n), "avg_us": ms / n * 1e3, "tflops": fl / (ms * 1e-3) / 1e12}

    side = {}
    if not args.no_side:
        # plain per-ik GPU path on the same configuration and handle: one
        # energy per batch, the same segment cycles
        k1 = ks[B // 2: B // 2 + 1]
        few = spread_segments(lay["n_seg"], 4)
        f1, r1, b1, p1, _ = cycle_rate(S, p, lay, k1, few, warm=[few[1]], seg_len=L_plan)
        t1, _ = batch_seconds(lay, N, p1, f1, r1, b1, 0.0, fwd_dip)
        side["per_ik"] = {"unit": "energy points/s", "segments": few, "modelled_energy_seconds": t1,
                          "value_without_fixed_stages": 1.0 / t1}
        log(f"per-ik: {t1:.1f} s per energy (without the fixed stages)")
    full = None
    if args.full_scan:
        t0 = time.time()
        res = S.run(ks, batch=B, seg_len=L_plan, groups=1)
        t_full = time.time() - t0
        if not all(np.all(np.isfinite(r["D"])) for r in res):
            raise SystemExit("non-finite dipole in the full scan")
        full = {"seconds": t_full, "value": B / t_full, "unit": "energy points/s",
                "stage_seconds": S.stats()}
        log(f"full scan of the batch: {t_full:.1f} s ({B / t_full:.4f} energy points/s)")
    S.close()   # the side runs build their own device tables

    t_fixed, fixed_info = 0.0, None
    if not args.no_side:
        t_fixed, fixed_info = fixed_stage_seconds(p, ks, B)
        log(f"fixed per-batch stages {t_fixed:.2f} s ({fixed_info})")
    t_batch, model = batch_seconds(lay, N, pts, tf, tr, tb, t_fixed, fwd_dip)
    value = B / t_batch
    log(f"modelled batch {t_batch:.1f} s -> {value:.4f} energy points/s")
    if "per_ik" in side:
        tb0, _ = batch_seconds(lay, N, pts, tf, tr, tb, 0.0, fwd_dip)
        side["per_ik"]["batched_over_per_ik"] = (B / tb0) * side["per_ik"]["modelled_energy_seconds"]

    mflop = model_flops_per_energy(info["n_psi"], info["n_f"], info["n_tot"], info["Nr"],
                                   info["n_transition"], info["n_start"], p.n_occ, info["n_pts"],
                                   (N - lay["n_fit_lo"] + 1) if fwd_dip else info["Nr"])
    if not args.no_side:
        # the forward-mode derivative on the l_cont = 20 polarisation
        # configuration: the same segment cycles with and without the tangent
        pc, kc = synth.config("h2o_l20_pol")
        S = sc.Scan(pc)
        S.set_dipole_path(dpath)
        lay2 = S.layout()
        kb = kc[(np.arange(B) * 37) % len(kc)]
        few2 = spread_segments(lay2["n_seg"], 4)
        ff, rf, bf, pf_, _ = cycle_rate(S, pc, lay2, kb, few2, warm=[few2[1]])
        ft, rt, bt, pt_, _ = cycle_rate(S, pc, lay2, kb, few2, warm=[few2[1]], tangent=(1.0, 0.0))
        S.close()
        tb_f, _ = batch_seconds(lay2, pc.Nr - 1, pf_, ff, rf, bf, 0.0, fwd_dip)
        tb_t, _ = batch_seconds(lay2, pc.Nr - 1, pt_, ft, rt, bt, 0.0, fwd_dip)
        side["tangent"] = {"workload": f"h2o_l20_pol (l_cont={pc.l_cont}, N_tot={pc.n_tot}, Nr={pc.Nr}), batch {B}, "
                                       f"segments {few2}, without the fixed stages",
                           "forward_energies_per_s": B / tb_f, "with_tangent_energies_per_s": B / tb_t,
                           "tangent_over_forward_cost": tb_t / tb_f}
        log(f"tangent / forward cost {tb_t / tb_f:.2f} on h2o_l20_pol")

    molecule = {"h2o": "H2O-like", "ring": "16-centre ring", "h2": "H2-like"}.get(args.config.split("_")[0], p.name)
    workload = (f"{args.config}: {molecule}, l_cont={p.l_cont} (N_psi={p.n_psi}, N_tot={p.n_tot}), Nr={p.Nr}, dr={p.dr}, "
                f"k in [{ks_all[0]}, {ks_all[-1]}] ({nk} pts), "
                + ("closed channels allowed" if p.allow_closed else "all channels open")
                + f"; batch of {B} energies; dipoles "
                + ("accumulated along the forward recursion" if fwd_dip else "back-propagated over the whole grid")
                + f".  A step is one full-size segment cycle: the batch runs {lay['L']} radial points of forward "
                  f"ratio steps, their recompute and their back-propagation at segment s (timed segments spread "
                  f"over the {lay['n_seg']} segments of the grid).  value = {B} / T_batch, T_batch = (Nr-1) x the "
                  f"measured per-point cost of the "
                + ("forward sweep + two recompute/back-propagation passes over the top segments (fit sums, fit "
                   "residual)" if fwd_dip else "three sweeps + one recompute/back-propagation pass over the top "
                                               "segments (fit sums)")
                + " at the same per-point costs"
                + (" (per-batch fixed stages not timed: --no-side)" if args.no_side else
                   " + the per-batch fixed stages (timed on a short-grid scan)")
                + f"; a {nk}-energy scan takes {nk} / value seconds")
    line = {
        "metric": METRIC,
        "value": value, "unit": "energy points/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / max(args.steps, 1) * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": workload, "batch_energies": B, "segment_points": lay["L"], "n_segments": lay["n_seg"],
                   "timed_segments": segs, "tiles": args.tiles, "dipole_path": args.dipole_path,
                   "storage": {"tables": plan["tables"], "checkpoints": plan["checkpoints"],
                               "planned_batch": plan["batch"], "batch_source": "plan" if args.batch == 0 else
                               ("autotune" if args.batch < 0 else "argument")}},
        "roofline": {"bound": "mfma", "achieved": gemm_tflops, "peak": PEAK_FP64_MFMA_TFLOPS, "unit": "TFLOP/s",
                     "frac": gemm_tflops / PEAK_FP64_MFMA_TFLOPS,
                     "traffic": TOP_UPDATE_TRAFFIC.get(B, (None,))[0] if args.config == "h2o_l30" else None,
                     "traffic_note": "HBM bytes of the 768-row main launch of the 783x1566x783 Gauss-Jordan update "
                                     "at this batch (algorithmic %s), from calibrated rocprofv3 --pmc runs of "
                                     "bench.py --kernels gemm (profiles/k_dgemm_counters_kernel_table_b64.md), not "
                                     "this process" % (TOP_UPDATE_TRAFFIC.get(B, (None, None))[1],),
                     "kernel": "k_dgemm", "sampled_launches": int(prof[2]), "avg_launch_us": avg_launch_us,
                     "event_pair_overhead_us": prof[15] * 1e3},
        "model": model, "fixed_stages": fixed_info,
        "energy_radial_steps_per_s": B * float(np.sum(pts)) / float(np.sum(tf) + np.sum(tr) + np.sum(tb)),
        "model_tflop_per_energy": mflop / 1e12,
        "path_tflops": mflop * value / 1e12,
        "path_frac_of_fp64_peak": mflop * value / 1e12 / PEAK_FP64_MFMA_TFLOPS,
        "setup_seconds": t_setup, "sampled_kernels": buckets, "prof_every": args.prof_every,
    }
    if full is not None:
        line["full_scan"] = full
    if side:
        line["side"] = side
    if tune is not None:
        line["autotune"] = tune
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(p, ks, args.cpu_steps)
        line["gpu_over_cpu"] = value / line["cpu_baseline"]["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
