"""psi files and the dipole-only replay (sehf_scan_set_psi_store,
sehf_scan_dipole_only).

The reference keeps the back-propagated psi_n of each energy on disk, the
dipole range [0, n_transition) and the 300-point asymptotic tail
(src/main.cpp:1259-1345, src/scatt/BackPropagator.cpp:27-101), and its
--dipole-only mode (main.cpp:167-189, 1278-1320) redoes only the fit and the
dipoles from them.  Checked here:
* the files: header fields of the reference's psi_asym.bin layout, the ranges,
  and every stored psi_n against the CPU oracle's psi_n at the same point;
* writing them changes no output of the scan;
* the replay gives bitwise the outputs of a back-propagated run (same psi,
  same kernels in the same order), also with a different initial state than
  the run that wrote the files, and fails clearly on a missing directory or
  an energy mismatch;
* the tail alone can be written by the default forward-accumulated path,
  bitwise the same tail as the back-propagated run's.
"""
import dataclasses

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

EXACT = ("A", "B", "K_fit", "b_overlap", "d_raw", "D_raw", "D", "d_corr")


@pytest.fixture(scope="module")
def ctx():
    from sehf.steppers import GpuContext
    return GpuContext(True)


def _problem():
    from sehf import synth
    # exchange cut-off inside the grid (n_transition < Nr), so the dipole
    # range and the tail are separate files with a gap between them
    return synth.h2_like(l_cont=2, dr=0.05, r_max=45.0, seed=0)


KS = np.array([0.5, 0.9, 1.3])


def _same(a, b, keys, label):
    for x, y in zip(a, b):
        for key in keys:
            assert np.array_equal(x[key], y[key]), (label, key, x["k"])
        assert np.array_equal(x["diag"][:11], y["diag"][:11]), (label, "diag", x["k"])


def test_psi_files_layout_and_oracle(ctx, tmp_path):
    import oracle_py as O
    from sehf import psi_files, scan
    p = _problem()
    S = scan.Scan(p, ctx)
    info = S.info
    assert info["n_transition"] < info["Nr"]
    S.set_psi_store(tmp_path, keep=scan.Scan.PSI_TAIL | scan.Scan.PSI_MAIN)
    got = S.run(KS, batch=2)
    S.close()
    Nr, Np, n_pts = info["Nr"], p.n_psi, info["n_pts"]
    n_asym = min(300, Nr)
    for ik, k in enumerate(KS):
        rec = psi_files.load_energy(tmp_path, ik, mmap=False)
        m = rec["marker"]
        assert m["k"] == k and m["Nr"] == Nr and m["N_psi"] == Np
        assert (m["n_asym"], m["asym_offset"], m["n_main"]) == (n_asym, Nr - n_asym, n_pts)
        h, tail = psi_files.read(psi_files.energy_dir(tmp_path, ik) + "/psi_asym.bin", "asym", mmap=False)
        assert (h["n"], h["first"], h["n_psi"]) == (n_asym, Nr - n_asym, Np)
        raw = open(psi_files.energy_dir(tmp_path, ik) + "/psi_asym.bin", "rb").read(32)
        assert raw[:8] == b"PSI_ASYM" and int.from_bytes(raw[8:12], "little") == 1
        assert raw[24:32] == bytes(8)
        first_m, main = rec["main"]
        assert first_m == 0 and main.shape == (n_pts, Np, Np)
        # the oracle's psi_n at points of both ranges, both ends of each
        pts_t = [Nr - n_asym, Nr - n_asym + 1, Nr - 150, Nr - 4, Nr - 1]
        pts_m = [0, 1, n_pts // 2, n_pts - 1]
        o = O.solve(p, float(k), dump_psi=pts_t + pts_m)
        errs = []
        for i, n in enumerate(pts_t):
            errs.append(rel_err(tail[n - (Nr - n_asym)], o["psi"][i]))
        for i, n in enumerate(pts_m):
            if n == 0:
                # psi_0 = Rinv_0 Z_1 with Rinv_0 = 0
                assert not np.any(main[0])
                continue
            errs.append(rel_err(main[n], o["psi"][len(pts_t) + i]))
        print(f"k={k}: stored psi vs oracle psi, worst {max(errs):.1e}")
        assert max(errs) < 1e-9, errs
        assert np.max(np.abs(tail[-1])) > 0.1   # psi_N ~ Riccati functions at the boundary
    # writing the files changes no output: a back-propagated run without a store
    S2 = scan.Scan(p, ctx)
    S2.set_dipole_path(scan.Scan.DIPOLES_BACKWARD)
    plain = S2.run(KS, batch=2)
    S2.close()
    _same(got, plain, EXACT + ("K",), "store vs plain")


def test_dipole_only_replay_bitwise(ctx, tmp_path):
    from sehf import scan
    p = _problem()
    S = scan.Scan(p, ctx)
    S.set_psi_store(tmp_path)
    ref = S.run(KS, batch=2)
    S.set_psi_store(None)
    rep = S.dipole_only(KS, tmp_path, batch=3)
    _same(ref, rep, EXACT, "replay")
    assert all(not r["K"].any() for r in rep)   # K is not formed
    # another initial state (stored raw): the HOMO's r F with a faster decay
    r = np.arange(p.Nr) * p.dr
    p2 = dataclasses.replace(p, init_psi_lm=(r * np.exp(-0.3 * r))[None, :] * p.psi_lm[p.n_occ - 1])
    S2 = scan.Scan(p2, ctx)
    S2.set_dipole_path(scan.Scan.DIPOLES_BACKWARD)
    ref2 = S2.run(KS, batch=2)
    rep2 = S2.dipole_only(KS, tmp_path, batch=2)
    _same(ref2, rep2, EXACT, "replay, other initial state")
    for a, b in zip(ref, ref2):
        assert rel_err(b["d_raw"], a["d_raw"]) > 1e-3   # the initial state matters
    # errors: a missing directory, an energy that is not the one on file
    with pytest.raises(RuntimeError, match="__SUCCESS__"):
        S.dipole_only(KS, tmp_path / "nothing")
    with pytest.raises(RuntimeError, match="energy mismatch"):
        S.dipole_only(KS[::-1].copy(), tmp_path)
    # the handle still runs after the failures
    again = S.dipole_only(KS[:1], tmp_path)
    _same(ref[:1], again, EXACT, "replay after errors")
    S.close()
    S2.close()


def test_tail_only_on_the_forward_path(ctx, tmp_path):
    """With only the tail kept, the default forward-accumulated path writes
    it (its fit pre-pass extended to the tail's first point) and its outputs
    are unchanged; the tail is bitwise the back-propagated run's."""
    from sehf import psi_files, scan
    p = _problem()
    S = scan.Scan(p, ctx)
    plain = S.run(KS, batch=3)
    S.set_psi_store(tmp_path / "tail", keep=scan.Scan.PSI_TAIL)
    tail_run = S.run(KS, batch=3)
    _same(plain, tail_run, EXACT + ("K",), "tail store, forward path")
    S.set_psi_store(tmp_path / "both", keep=scan.Scan.PSI_TAIL | scan.Scan.PSI_MAIN)
    S.run(KS, batch=3)
    S.close()
    for ik in range(len(KS)):
        a = psi_files.load_energy(tmp_path / "tail", ik, mmap=False)
        b = psi_files.load_energy(tmp_path / "both", ik, mmap=False)
        assert "main" not in a and a["marker"]["n_main"] == 0
        assert np.array_equal(a["asym"][1], b["asym"][1])
