"""Readers of the per-energy psi files written by `Scan.set_psi_store`
(sehf_scan_set_psi_store), for consumers of the continuum wave function
(the fit and dipole replay, continuum-continuum dipoles).

<root>/psi_ik<NNNN>/psi_asym.bin holds the 300-point asymptotic tail in the
reference's single-file layout (src/scatt/BackPropagator.cpp:27-40): a
32-byte header (magic b"PSI_ASYM", int32 version 1, int32 n_asym, int32 first
grid index, int32 N_psi, 8 reserved bytes) followed by n_asym N_psi x N_psi
matrices, column-major, in ascending grid index.  psi_main.bin holds
[0, n_main) in the same layout with the magic b"PSI_MAIN".  __SUCCESS__ is
written last and records k and the sizes.
"""
from __future__ import annotations

import os
import struct

import numpy as np

HEADER = struct.Struct("<8siiiiq")
MAGIC = {"asym": b"PSI_ASYM", "main": b"PSI_MAIN"}
FILES = {"asym": "psi_asym.bin", "main": "psi_main.bin"}


def energy_dir(root, ik: int) -> str:
    return os.path.join(os.fspath(root), f"psi_ik{ik:04d}")


def read_marker(root, ik: int) -> dict:
    """The completed directory's record: k (float) and the sizes (int).
    Raises FileNotFoundError when the directory is incomplete."""
    out = {}
    with open(os.path.join(energy_dir(root, ik), "__SUCCESS__")) as f:
        for line in f:
            key, val = line.split()
            out[key] = float(val) if key == "k" else int(val)
    return out


def read(path, which: str | None = None, mmap: bool = True):
    """(header dict, psi) of one psi file; psi[i] is the matrix of grid
    index first + i as psi[i][mu, j] (channel mu, solution j).  With mmap the
    matrices stay on disk until touched."""
    with open(path, "rb") as f:
        raw = f.read(HEADER.size)
    if len(raw) != HEADER.size:
        raise ValueError(f"{path}: shorter than the 32-byte header")
    magic, version, n, first, n_psi, _ = HEADER.unpack(raw)
    kinds = [k for k, m in MAGIC.items() if m == magic]
    if not kinds or version != 1:
        raise ValueError(f"{path}: not a psi file (magic {magic!r}, version {version})")
    if which is not None and kinds[0] != which:
        raise ValueError(f"{path}: holds {kinds[0]!r}, not {which!r}")
    shape = (n, n_psi, n_psi)
    need = HEADER.size + n * n_psi * n_psi * 8
    if os.path.getsize(path) < need:
        raise ValueError(f"{path}: shorter than its header says")
    if mmap:
        a = np.memmap(path, dtype="<f8", mode="r", offset=HEADER.size, shape=shape)
    else:
        a = np.fromfile(path, dtype="<f8", count=n * n_psi * n_psi, offset=HEADER.size).reshape(shape)
    # column-major per matrix: element (mu, j) at mu + j N_psi
    return {"kind": kinds[0], "n": n, "first": first, "n_psi": n_psi}, a.transpose(0, 2, 1)


def load_energy(root, ik: int, mmap: bool = True) -> dict:
    """Everything stored for energy ik: the marker and, where present, the
    tail and the main range as (first index, matrices)."""
    rec = {"marker": read_marker(root, ik)}
    for kind, name in FILES.items():
        p = os.path.join(energy_dir(root, ik), name)
        if os.path.exists(p):
            h, a = read(p, kind, mmap)
            rec[kind] = (h["first"], a)
    return rec
