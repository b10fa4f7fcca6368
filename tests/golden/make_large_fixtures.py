"""Writes the full-size oracle fixtures of the l_cont = 20 and 30 configurations.

One fixture per (configuration, k[, tangent direction]): the CPU oracle's
outputs for one energy of the named benchmark configuration
(sehf.synth.config, deterministic from its seeds), kept compact:

  D          reduced dipoles d_lm, 6 x N_psi complex (both gauges)
  d_raw      raw dipole integrals, 6 x N_psi
  b_overlap  orthogonality integrals, N_psi x n_occ
  K_diag, K_rows, K_row_idx    diagonal and three full rows of the K matrix
  Kfit_diag, A_diag, B_diag    diagonals of the fitted K, A and B
  diag       the oracle's 16 diagnostics (fit residuals, K_fit symmetry,
             |A - I|, n_transition, n_start, fit window, shift counts,
             closed-l count, unitarity errors)
  dD, dd_raw, dK_diag  (tangent fixtures) derivatives along (t_E, t_alpha)

ORC_INVERSE_MODE=32 runs the oracle with correctly rounded inverses (every
dense inverse by 80-bit LU rounded once to FP64, all else unchanged; the
parity referee of the energy-amplified outputs) and writes
large_<config>_ik<ik>_mode32.npz.

Usage (one energy per process; ORC_THREADS threads inside the solve):
    ORC_THREADS=4 python tests/golden/make_large_fixtures.py h2o_l30 0
    ORC_THREADS=2 python tests/golden/make_large_fixtures.py h2o_l20_pol 31 1.0 0.5
    ORC_INVERSE_MODE=32 ORC_THREADS=2 python tests/golden/make_large_fixtures.py h2o_l20 52
The second argument is the index into the configuration's k grid.  The
oracle keeps sqrt(Nr)-spaced ratio-matrix checkpoints, so one l_cont = 30
energy needs about 6 GB of RAM; it runs for hours on a few cores.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "spion-emol_amd"), os.path.join(ROOT, "oracle")]


def fixture_name(config, ik, tangent=None, mode=0):
    tag = "" if tangent is None else f"_tan{tangent[0]:g}_{tangent[1]:g}"
    if mode:
        tag += f"_mode{mode}"
    return os.path.join(HERE, f"large_{config}_ik{ik}{tag}.npz")


def main():
    import oracle_py as O
    from sehf import synth
    config, ik = sys.argv[1], int(sys.argv[2])
    tangent = (float(sys.argv[3]), float(sys.argv[4])) if len(sys.argv) > 4 else None
    O.set_threads(int(os.environ.get("ORC_THREADS", "1")))
    mode = int(os.environ.get("ORC_INVERSE_MODE", "0"))
    if mode:
        O.lib().sehf_oracle_set_inverse_mode(mode)
    p, ks = synth.config(config)
    k = float(ks[ik])
    t0 = time.time()
    r = O.solve(p, k, tangent=tangent)
    N = p.n_psi
    rows = np.array([0, N // 2, N - 1])
    out = {"config": config, "ik": ik, "k": k, "seconds": time.time() - t0, "inverse_mode": mode,
           "D": r["D"], "d_raw": r["d_raw"], "b_overlap": r["b_overlap"],
           "K_diag": np.diag(r["K_kme"]).copy(), "K_rows": r["K_kme"][rows].copy(), "K_row_idx": rows,
           "Kfit_diag": np.diag(r["K_fit"]).copy(), "A_diag": np.diag(r["A"]).copy(),
           "B_diag": np.diag(r["B"]).copy(), "diag": r["diag"]}
    if tangent is not None:
        out.update({"tangent": np.array(tangent), "dD": r["dD"], "dd_raw": r["dd_raw"],
                    "dK_diag": np.diag(r["dK"]).copy()})
    path = fixture_name(config, ik, tangent, mode)
    np.savez_compressed(path, **out)
    print(f"{path}: k={k} in {out['seconds']:.0f} s", flush=True)


if __name__ == "__main__":
    main()
