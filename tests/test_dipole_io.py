"""Per-energy HDF5 output in the reference's DipoleWriter layout (CPU).

Follows the reference's round-trip test src/tests/test_dipole_io.cpp:60-208
(same scan metadata, random payload, bit-equal read-back, no temporary files
left, a resumed scan keeps its manifest, tag padding), and pins the stored
layout against src/scatt/DipoleIO.cpp:141-335: object paths, HDF5 type class
and size of every dataset and attribute, and matrix orientation.  The reader
used here calls the HDF5 C library directly and shares no code with the
writer (libsehf_io.so).
"""
import os
import time

import numpy as np
import pytest

from sehf import dipole_io as dio

pytestmark = pytest.mark.skipif(not os.path.exists(dio.IO_LIB_PATH) or not os.path.exists(dio.HDF5_LIB),
                                reason="libsehf_io.so or the HDF5 C library is not present")

F64, I32, I64 = (dio.H5T_FLOAT, 8, None), (dio.H5T_INTEGER, 4, None), (dio.H5T_INTEGER, 8, None)


def ref_meta(**kw):
    # src/tests/test_dipole_io.cpp:60-93
    m = dict(r_min=0.0, dr=0.005, N_grid=4001, l_max_continuum=8, dk=0.02, ik_min=1, ik_max=6, E_HOMO=-0.493,
             occ_energies=[-19.25, -1.34, -0.72, -0.58, -0.493], occ_spin_factors=[2.0, 2.0, 2.0, 2.0, 1.0],
             atoms=[(8, 0.0, 0.0, 0.0), (1, 0.757, 0.0, 0.587), (1, -0.757, 0.0, 0.587)],
             molecule_name="H2O_test", git_hash="deadbeef", iso_date_utc="2026-04-23T00:00:00Z",
             psi_dir_prefix="/scratch/di35ker/psi_H2O_test")
    m.update(kw)
    return dio.ScanMeta(**m)


def payload(n_ch, n_occ, k, seed=0x1234):
    """A Scan.run-shaped result with uniform(-1, 1) entries."""
    g = np.random.default_rng(seed)
    U = lambda *s: g.uniform(-1.0, 1.0, s)
    diag = np.zeros(16)
    diag[0], diag[2] = 1.2e-7, 3.4e-9
    return {"k": k, "A": U(n_ch, n_ch), "B": U(n_ch, n_ch), "b_overlap": U(n_ch, n_occ), "d_raw": U(6, n_ch),
            "D_raw": U(6, n_ch) + 1j * U(6, n_ch), "D": U(6, n_ch) + 1j * U(6, n_ch), "d_corr": U(6, n_occ),
            "diag": diag}


def seq_sigma(D):
    s = 0.0
    for z in D:
        s += z.real * z.real + z.imag * z.imag
    return s


def test_round_trip_reference_case(tmp_path):
    meta = ref_meta()
    scan = str(tmp_path / "dipole_testscan")
    n_ch, n_occ = 81, 5
    p = payload(n_ch, n_occ, k=3 * meta.dk)
    dio.open_scan(scan, meta)
    assert not dio.has_energy(scan, 3)
    dio.write_energy(scan, meta, 3, p)
    assert dio.has_energy(scan, 3)
    dio.finalize(scan)
    names = sorted(os.listdir(scan))
    assert names == ["__SUCCESS__", "ik0003.h5", "manifest.h5"], names     # no temporary files left
    assert os.path.getsize(os.path.join(scan, "__SUCCESS__")) == 0

    # a resumed scan keeps its manifest
    t0 = os.stat(os.path.join(scan, "manifest.h5")).st_mtime_ns
    time.sleep(0.01)
    dio.open_scan(scan, meta)
    assert os.stat(os.path.join(scan, "manifest.h5")).st_mtime_ns == t0
    assert dio.has_energy(scan, 3)

    m = dio.read_manifest(scan)
    for key in ("r_min", "dr", "N_grid", "l_max_continuum", "E_HOMO", "dk", "ik_min", "ik_max", "molecule_name",
                "git_hash", "iso_date_utc", "psi_dir_prefix"):
        assert m[key] == getattr(meta, key), key
    assert np.array_equal(m["occ_energies"], meta.occ_energies)
    assert np.array_equal(m["occ_spin_factors"], meta.occ_spin_factors)
    assert m["atom_Z"].tolist() == [a[0] for a in meta.atoms]
    assert np.array_equal(m["atom_xyz"], np.array([a[1:] for a in meta.atoms]))
    assert dio.available_ik(scan) == [3]

    r = dio.read_energy(scan, 3)
    assert r["ik"] == 3
    k = 3 * meta.dk
    assert r["k"] == k and r["E"] == 0.5 * k * k and r["omega"] == 0.5 * k * k - meta.E_HOMO
    assert r["fit_residual_rel"] == 1.2e-7 and r["K_symmetry_err"] == 3.4e-9
    for key in ("A", "B", "b_overlap", "d_raw", "D_raw", "D"):
        assert np.array_equal(r[key], p[key]), key
    assert np.array_equal(r["d_corr"], p["d_corr"])
    for s in range(6):   # partial_sigma = sum |D|^2 of the orthogonalised slice
        assert r["partial_sigma"][s] == seq_sigma(p["D"][s])
        assert abs(r["partial_sigma"][s] - np.sum(np.abs(p["D"][s]) ** 2)) <= 1e-13 * r["partial_sigma"][s]


def test_stored_layout(tmp_path):
    meta = ref_meta()
    scan = str(tmp_path / "s")
    n_ch, n_occ = 9, 5
    p = payload(n_ch, n_occ, k=2 * meta.dk, seed=7)
    p["A"] = np.arange(n_ch * n_ch, dtype=float).reshape(n_ch, n_ch)   # orientation: A[i, j] = 9 i + j
    dio.write_scan(scan, meta, [p])
    m = dio.read_manifest(scan)
    t = m["types"]
    for a in ("r_min", "dr", "E_HOMO"):
        assert t["/grid@" + a] == (F64, ()), a
    for a in ("N_grid", "l_max_continuum"):
        assert t["/grid@" + a] == (I64, ()), a
    assert t["/kgrid@dk"] == (F64, ())
    assert t["/kgrid@ik_min"] == (I64, ()) and t["/kgrid@ik_max"] == (I64, ())
    assert t["/occ/energies"] == (F64, (5,)) and t["/occ/spin_factors"] == (F64, (5,))
    assert t["/atoms/Z"] == (I32, (3,)) and t["/atoms/xyz_bohr"] == (F64, (3, 3))
    # fixed-length, null-padded strings of the text's own length (H5T_STR_NULLPAD = 1)
    for a, v in m["conventions"].items():
        assert t["/conventions@" + a] == ((dio.H5T_STRING, len(v), 1), ()), a
    assert m["conventions"]["real_Ylm_q_map"] == "x=+1, y=-1, z=0"
    assert m["conventions"]["psi_norm"] == "incoming-wave Psi- = (A - iB)^(-dagger)"
    assert m["conventions"]["u_convention"] == "chi_lm(r) = r * F_lm(r)"
    assert t["/run@molecule_name"] == ((dio.H5T_STRING, len("H2O_test"), 1), ())

    r = dio.read_energy(scan, 2)
    t = r["types"]
    assert t["/meta@ik"] == (I64, ())
    for a in ("k", "E", "omega", "fit_residual_rel", "K_symmetry_err"):
        assert t["/meta@" + a] == (F64, ()), a
    assert t["/A"] == (F64, (n_ch, n_ch)) and t["/B"] == (F64, (n_ch, n_ch))
    assert t["/b_overlap"] == (F64, (n_ch, n_occ))
    assert r["A"][2, 5] == 23.0 and r["A"][5, 2] == 47.0      # row index first, as h5py reads it
    for g in ("length", "velocity"):
        for q in ("x", "y", "z"):
            b = f"/dipole/{g}/{q}"
            for d in ("D_ortho_re", "D_ortho_im", "D_raw_re", "D_raw_im", "d_raw"):
                assert t[f"{b}/{d}"] == (F64, (n_ch,)), (b, d)
            assert t[f"{b}/d_correction"] == (F64, (n_occ,))
            assert t[f"{b}@partial_sigma"] == (F64, ())
    # slot order L x, L y, L z, V x, V y, V z (DipoleIO.hpp:132-139)
    assert np.array_equal(r["d_raw"], p["d_raw"])


def test_tags_resume_and_errors(tmp_path):
    meta = ref_meta(ik_max=20000)
    scan = str(tmp_path / "t")
    res = [payload(4, 2, k=ik * meta.dk, seed=ik) for ik in (7, 1234, 12345)]
    assert dio.write_scan(scan, meta, res) == [7, 1234, 12345]
    assert sorted(n for n in os.listdir(scan) if n.startswith("ik")) == ["ik0007.h5", "ik1234.h5", "ik12345.h5"]
    assert dio.available_ik(scan) == [7, 1234, 12345]
    assert dio.write_scan(scan, meta, res) == []                      # resumed: nothing rewritten
    with pytest.raises(ValueError, match="not on the grid"):
        dio.write_scan(scan, meta, [payload(4, 2, k=0.013)])
    with pytest.raises(RuntimeError, match="dk must be > 0"):
        dio.open_scan(str(tmp_path / "u"), ref_meta(dk=0.0))
    with pytest.raises(RuntimeError, match="ik < 0"):
        dio.write_energy(scan, meta, -1, res[0])
    with pytest.raises(ValueError):
        dio.write_energy(scan, meta, 8, {**res[0], "d_raw": np.zeros((5, 4))})


def test_no_occupied_orbitals_and_no_atoms(tmp_path):
    meta = ref_meta(occ_energies=[], occ_spin_factors=[], atoms=[], molecule_name="")
    scan = str(tmp_path / "v")
    p = payload(4, 0, k=meta.dk, seed=3)
    dio.write_scan(scan, meta, [p])
    m = dio.read_manifest(scan)
    assert "occ_energies" not in m and "atom_Z" not in m
    assert m["molecule_name"] == ""
    assert m["types"]["/run@molecule_name"] == ((dio.H5T_STRING, 1, 1), ())
    r = dio.read_energy(scan, 1)
    assert r["b_overlap"].shape == (4, 0) and r["d_corr"].shape == (6, 0)
    assert np.array_equal(r["D"], p["D"])


def test_library_exports_every_declared_symbol():
    import ctypes
    import re
    from conftest import ROOT
    hdr = open(os.path.join(ROOT, "include", "sehf_io.h")).read()
    names = sorted(set(re.findall(r"\b(sehf_io_\w+)\s*\(", hdr)))
    assert names == ["sehf_io_finalize", "sehf_io_has_energy", "sehf_io_last_error", "sehf_io_open",
                     "sehf_io_write_energy"]
    L = ctypes.CDLL(dio.IO_LIB_PATH)
    for n in names:
        assert hasattr(L, n), n


def test_idx_to_lm():
    seen = []
    for l in range(12):
        for m in range(-l, l + 1):
            seen.append((l, m))
    assert [dio.idx_to_lm(mu) for mu in range(len(seen))] == seen


def test_gather_dat_files(tmp_path):
    """The per-channel files of postprocessing/gather_dipoles.py: one per
    (gauge, polarisation, channel); the columns reproduce k, E, omega and D
    to the 15 digits of %.14e, |D|^2 and arg D."""
    meta = ref_meta(l_max_continuum=2)
    scan = str(tmp_path / "g")
    iks = (5, 2, 9)
    res = [payload(9, 5, k=ik * meta.dk, seed=ik) for ik in iks]
    dio.write_scan(scan, meta, res)
    out = str(tmp_path / "gathered")
    assert dio.gather(scan, out) == 6 * 9
    order = np.argsort(iks)
    D = np.stack([res[i]["D"] for i in order])
    for s, name in enumerate(["len_homo_x", "len_homo_y", "len_homo_z", "vel_homo_x", "vel_homo_y", "vel_homo_z"]):
        for mu in range(9):
            f = os.path.join(out, f"dipole_{name}_{mu}.dat")
            head = [ln for ln in open(f) if ln.startswith("#")]
            l, m = dio.idx_to_lm(mu)
            assert head[-1].strip() == f"# mu = {mu}  (l={l}, m={m:+d})"
            assert "dipole = ORTHO" in head[1]
            c = np.loadtxt(f)
            k = np.array(sorted(iks)) * meta.dk
            assert np.allclose(c[:, 0], k, rtol=1e-14, atol=0)
            assert np.allclose(c[:, 1], 0.5 * k * k, rtol=1e-14, atol=0)
            assert np.allclose(c[:, 2], 0.5 * k * k - meta.E_HOMO, rtol=1e-14, atol=0)
            d = D[:, s, mu]
            assert np.allclose(c[:, 3] + 1j * c[:, 4], d, rtol=1e-14, atol=0)
            assert np.allclose(c[:, 5], d.real ** 2 + d.imag ** 2, rtol=1e-13, atol=0)
            assert np.allclose(c[:, 6], np.arctan2(d.imag, d.real), rtol=1e-13, atol=0)
    _, raw = dio.load_results(scan, use_raw=True)
    assert np.array_equal(raw[0]["D"], res[1]["D_raw"])      # ik 2 first


def test_gathered_files_give_the_same_cross_sections(tmp_path):
    """The .dat files, loaded the way postprocessing/cross_section_delay.py
    loads them (np.loadtxt per channel, D = Re + i Im, channels of x, y, z
    side by side; :167-192), give the cross sections of the numpy
    restatement equal to sehf.observables on the in-memory results."""
    import sys
    import torch
    from conftest import ROOT
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import observables_ref as OR
    from sehf import observables
    meta = ref_meta(l_max_continuum=2)
    scan = str(tmp_path / "x")
    res = [payload(9, 5, k=ik * meta.dk, seed=10 + ik) for ik in range(3, 9)]
    dio.write_scan(scan, meta, res)
    out = str(tmp_path / "gx")
    dio.gather(scan, out)

    def load(prefix):
        cols = [np.loadtxt(os.path.join(out, f"{prefix}_{pol}_{mu}.dat")) for pol in "xyz" for mu in range(9)]
        return cols[0][:, 0], cols[0][:, 2], np.stack([c[:, 3] + 1j * c[:, 4] for c in cols], axis=1)

    kL, omL, DL = load("dipole_len_homo")
    kV, omV, DV = load("dipole_vel_homo")
    sL = OR.sigma_from_dipole(kL, DL, omL, "length")
    sV = OR.sigma_from_dipole(kV, DV, omV, "velocity")
    got = observables.from_scan(res, meta.E_HOMO, device="cpu")
    assert np.allclose(got["sigma_L"].numpy(), sL, rtol=1e-13, atol=0)
    assert np.allclose(got["sigma_V"].numpy(), sV, rtol=1e-13, atol=0)
    assert isinstance(got["sigma_L"], torch.Tensor)


def test_failed_write_leaves_no_file(tmp_path):
    """A write that fails inside HDF5 leaves neither the target nor its
    temporary file (here the temporary name is taken by an empty directory,
    so H5Fcreate fails), and a payload of inconsistent shapes is refused
    before any file is created."""
    meta = ref_meta()
    scan = str(tmp_path / "f")
    dio.open_scan(scan, meta)
    os.makedirs(os.path.join(scan, "ik0004.h5.tmp.pid%d" % os.getpid()))
    with pytest.raises(RuntimeError, match="H5Fcreate"):
        dio.write_energy(scan, meta, 4, payload(4, 2, k=4 * meta.dk))
    assert not dio.has_energy(scan, 4)
    assert sorted(os.listdir(scan)) == ["manifest.h5"]
    with pytest.raises(ValueError):
        dio.write_energy(scan, meta, 4, {**payload(4, 2, k=4 * meta.dk), "d_corr": np.zeros((6, 3))})
    assert sorted(os.listdir(scan)) == ["manifest.h5"]


def test_gather_command_line(tmp_path, capsys):
    """python -m sehf.dipole_io <scan_dir> [--output-dir] [--raw], the
    replacement of postprocessing/gather_dipoles.py's command line."""
    meta = ref_meta(l_max_continuum=1)
    scan = str(tmp_path / "c")
    res = [payload(4, 5, k=ik * meta.dk, seed=ik) for ik in (2, 3)]
    dio.write_scan(scan, meta, res)
    out = str(tmp_path / "o")
    dio._main([scan, "--output-dir", out, "--raw"])
    assert "wrote 24 .dat files" in capsys.readouterr().out
    f = os.path.join(out, "dipole_vel_homo_z_3.dat")
    head = [ln for ln in open(f) if ln.startswith("#")]
    assert "dipole = RAW (pre-ortho)" in head[1]
    c = np.loadtxt(f)
    assert np.allclose(c[:, 3] + 1j * c[:, 4], [r["D_raw"][5, 3] for r in res], rtol=1e-14, atol=0)
