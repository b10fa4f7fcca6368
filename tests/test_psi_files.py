"""The psi file reader (sehf.psi_files) on files built byte by byte from the
layout of the reference's psi_asym.bin (src/scatt/BackPropagator.cpp:27-40):
8-byte magic, int32 version, int32 count, int32 first index, int32 N_psi, 8
reserved bytes, then column-major matrices."""
import struct

import numpy as np
import pytest


def _write(path, magic, mats, first, version=1):
    n, Np, _ = mats.shape
    with open(path, "wb") as f:
        f.write(magic + struct.pack("<iiii", version, n, first, Np) + bytes(8))
        for m in mats:
            f.write(np.asarray(m, dtype="<f8").tobytes(order="F"))   # column-major


def test_reader_orientation_and_header(tmp_path):
    from sehf import psi_files
    rng = np.random.default_rng(0)
    mats = rng.standard_normal((5, 3, 3))
    p = tmp_path / "psi_asym.bin"
    _write(p, b"PSI_ASYM", mats, first=95)
    for mm in (True, False):
        h, a = psi_files.read(p, "asym", mmap=mm)
        assert h == {"kind": "asym", "n": 5, "first": 95, "n_psi": 3}
        assert np.array_equal(np.asarray(a), mats)   # a[i][mu, j]
    assert p.stat().st_size == 32 + 5 * 9 * 8


def test_reader_rejects_bad_files(tmp_path):
    from sehf import psi_files
    mats = np.ones((2, 2, 2))
    bad = tmp_path / "x.bin"
    _write(bad, b"PSI_XXXX", mats, 0)
    with pytest.raises(ValueError, match="not a psi file"):
        psi_files.read(bad)
    _write(bad, b"PSI_ASYM", mats, 0, version=2)
    with pytest.raises(ValueError, match="not a psi file"):
        psi_files.read(bad)
    _write(bad, b"PSI_MAIN", mats, 0)
    with pytest.raises(ValueError, match="holds 'main'"):
        psi_files.read(bad, "asym")
    with open(bad, "r+b") as f:
        f.truncate(32 + 8 * 7)
    with pytest.raises(ValueError, match="shorter"):
        psi_files.read(bad)


def test_load_energy(tmp_path):
    from sehf import psi_files
    d = tmp_path / "psi_ik0003"
    d.mkdir()
    _write(d / "psi_asym.bin", b"PSI_ASYM", np.zeros((3, 2, 2)), first=7)
    with pytest.raises(FileNotFoundError):
        psi_files.load_energy(tmp_path, 3)
    (d / "__SUCCESS__").write_text("k 0.5\nNr 10\nN_psi 2\nn_asym 3\nasym_offset 7\nn_main 0\n")
    rec = psi_files.load_energy(tmp_path, 3)
    assert rec["marker"]["k"] == 0.5 and rec["marker"]["asym_offset"] == 7
    assert rec["asym"][0] == 7 and "main" not in rec
