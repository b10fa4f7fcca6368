"""A device scan written in the reference's DipoleWriter layout and read back.

The files hold exactly the scan's outputs (bit-equal), the manifest carries
the problem's grid, orbitals and atoms as src/main.cpp:765-841 fills it, and
the cross sections recomputed from the files equal those of the in-memory
results.
"""
import numpy as np
import pytest
import torch

from sehf import dipole_io as dio
from sehf import observables, scan, synth

pytestmark = pytest.mark.gpu


def test_scan_to_files_round_trip(tmp_path):
    p = synth.h2_like(l_cont=3, dr=0.02, r_max=14.0, seed=1)
    dk = 0.05
    iks = [10, 14, 19]
    res = scan.Scan(p).run(np.array(iks) * dk)
    meta = dio.meta_from_problem(p, dk, 1, 41, molecule_name=p.name)
    d = str(tmp_path / "scan")
    print(f"[dipole_io] {p.name}, ik {iks}", flush=True)
    assert dio.write_scan(d, meta, res) == iks
    m = dio.read_manifest(d)
    assert m["N_grid"] == p.Nr and m["r_min"] == p.dr and m["l_max_continuum"] == p.l_cont
    assert m["E_HOMO"] == float(p.orb_energies[p.n_occ - 1])
    assert np.array_equal(m["atom_Z"], p.Z) and np.array_equal(m["atom_xyz"], p.xyz)
    for ik, r in zip(iks, res):
        f = dio.read_energy(d, ik)
        assert f["k"] == ik * dk
        for key in ("A", "B", "b_overlap", "d_raw", "D_raw", "D"):
            assert np.array_equal(f[key], r[key]), (ik, key)
        assert np.array_equal(f["d_corr"], r["d_corr"])
        assert f["fit_residual_rel"] == r["diag"][0] and f["K_symmetry_err"] == r["diag"][2]
        assert np.allclose(f["partial_sigma"], np.sum(np.abs(r["D"]) ** 2, axis=1), rtol=1e-13, atol=0)
    # cross sections from the files equal those from the results
    E_homo = m["E_HOMO"]
    kf = np.array([dio.read_energy(d, ik)["k"] for ik in iks])
    Df = np.array([dio.read_energy(d, ik)["D"] for ik in iks])
    Dr = np.array([r["D"] for r in res])
    a = observables.from_scan([{"k": k, "D": D} for k, D in zip(kf, Df)], E_homo)
    b = observables.from_scan([{"k": r["k"], "D": r["D"]} for r in res], E_homo)
    for key in ("sigma_L", "sigma_V"):
        assert torch.equal(a[key], b[key]), key
    assert np.array_equal(Df, Dr)
