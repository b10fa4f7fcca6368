"""Per-energy output of a scan in the reference's DipoleWriter layout.

Writer: ctypes over spion-emol_amd/_build/libsehf_io.so (include/sehf_io.h),
the host-only HDF5 writer that mirrors src/scatt/DipoleIO.cpp:111-340:

    <scan_dir>/manifest.h5, <scan_dir>/ik<NNNN>.h5, <scan_dir>/__SUCCESS__

Reader: read_manifest / read_energy mirror DipoleReader
(src/scatt/DipoleIO.cpp:427-533) and also report each object's stored type
and shape.  They call the HDF5 C library directly (h5py is not needed), so
they share no code with the writer.

The reference names energies by an index on a uniform k grid,
k(ik) = ik * dk (src/scatt/EnergyGrid.hpp:24-29); write_scan maps each
result's k to its ik and refuses a k that is not on the grid.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
IO_LIB_PATH = os.path.join(_PKG_ROOT, "_build", "libsehf_io.so")
HDF5_LIB = os.path.join(os.environ.get("HDF5_ROOT", "/opt/conda"), "lib", "libhdf5.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class Meta_C(ctypes.Structure):
    _fields_ = [("r_min", ctypes.c_double), ("dr", ctypes.c_double), ("N_grid", ctypes.c_longlong),
                ("l_max_continuum", ctypes.c_int), ("dk", ctypes.c_double), ("ik_min", ctypes.c_int),
                ("ik_max", ctypes.c_int), ("E_HOMO", ctypes.c_double), ("n_occ", ctypes.c_int),
                ("occ_energies", _dp), ("occ_spin_factors", _dp), ("n_atoms", ctypes.c_int),
                ("atom_Z", _ip), ("atom_xyz", _dp), ("molecule_name", ctypes.c_char_p),
                ("git_hash", ctypes.c_char_p), ("iso_date_utc", ctypes.c_char_p),
                ("psi_dir_prefix", ctypes.c_char_p)]


@dataclass
class ScanMeta:
    """DipoleScanMeta (src/scatt/DipoleIO.hpp:42-79)."""
    r_min: float
    dr: float
    N_grid: int
    l_max_continuum: int
    dk: float
    ik_min: int
    ik_max: int                      # exclusive
    E_HOMO: float = 0.0
    occ_energies: Sequence[float] = ()
    occ_spin_factors: Sequence[float] = ()
    atoms: Sequence[tuple] = ()      # (Z, x, y, z) in bohr
    molecule_name: str = ""
    git_hash: str = ""
    iso_date_utc: str = ""
    psi_dir_prefix: str = ""
    _keep: list = field(default_factory=list, repr=False)

    def c(self) -> Meta_C:
        oe = np.ascontiguousarray(self.occ_energies, dtype=np.float64)
        osf = np.ascontiguousarray(self.occ_spin_factors, dtype=np.float64)
        if oe.shape != osf.shape:
            raise ValueError("occ_energies and occ_spin_factors differ in length")
        Z = np.ascontiguousarray([int(a[0]) for a in self.atoms], dtype=np.int32)
        xyz = np.ascontiguousarray([[float(v) for v in a[1:4]] for a in self.atoms], dtype=np.float64).reshape(-1, 3)
        self._keep = [oe, osf, Z, xyz]
        return Meta_C(self.r_min, self.dr, int(self.N_grid), int(self.l_max_continuum), self.dk, int(self.ik_min),
                      int(self.ik_max), self.E_HOMO, int(oe.size), oe.ctypes.data_as(_dp), osf.ctypes.data_as(_dp),
                      int(Z.size), Z.ctypes.data_as(_ip), xyz.ctypes.data_as(_dp), self.molecule_name.encode(),
                      self.git_hash.encode(), self.iso_date_utc.encode(), self.psi_dir_prefix.encode())


def meta_from_problem(p, dk: float, ik_min: int, ik_max: int, **strings) -> ScanMeta:
    """The manifest the reference's driver writes for a problem
    (src/main.cpp:765-841): solver r_min = dr (WavefunctionSetup.cpp:327),
    N_grid = Nr, E_HOMO = the initial-state energy when the problem carries
    one, else the last occupied orbital's; spin factor 2 per orbital."""
    e_homo = p.init_energy if p.init_psi_lm is not None else float(p.orb_energies[p.n_occ - 1])
    return ScanMeta(r_min=p.dr, dr=p.dr, N_grid=p.Nr, l_max_continuum=p.l_cont, dk=dk, ik_min=ik_min,
                    ik_max=ik_max, E_HOMO=e_homo, occ_energies=[float(e) for e in p.orb_energies[:p.n_occ]],
                    occ_spin_factors=[2.0] * p.n_occ,
                    atoms=[(int(z), *map(float, x)) for z, x in zip(p.Z, p.xyz)], **strings)


_io = None


def io_lib():
    """The writer library; raises if it was not built (no HDF5 found)."""
    global _io
    if _io is None:
        if not os.path.exists(IO_LIB_PATH):
            raise RuntimeError(f"{IO_LIB_PATH} is missing: build it with make -C spion-emol_amd/csrc "
                               "(needs the HDF5 C library, HDF5_ROOT)")
        L = ctypes.CDLL(IO_LIB_PATH)
        L.sehf_io_last_error.restype = ctypes.c_char_p
        L.sehf_io_open.argtypes = [ctypes.c_char_p, ctypes.POINTER(Meta_C)]
        L.sehf_io_has_energy.argtypes = [ctypes.c_char_p, ctypes.c_int]
        L.sehf_io_write_energy.argtypes = [ctypes.c_char_p, ctypes.POINTER(Meta_C), ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int] + [_dp] * 7 + [ctypes.c_double, ctypes.c_double]
        L.sehf_io_finalize.argtypes = [ctypes.c_char_p]
        _io = L
    return _io


def _check(rc: int, what: str) -> int:
    if rc < 0:
        raise RuntimeError(f"{what}: {io_lib().sehf_io_last_error().decode()}")
    return rc


def open_scan(scan_dir: str, meta: ScanMeta) -> None:
    """Create the directory and manifest.h5 (kept if it exists)."""
    m = meta.c()
    _check(io_lib().sehf_io_open(scan_dir.encode(), ctypes.byref(m)), "sehf_io_open")


def has_energy(scan_dir: str, ik: int) -> bool:
    return _check(io_lib().sehf_io_has_energy(scan_dir.encode(), int(ik)), "sehf_io_has_energy") == 1


def finalize(scan_dir: str) -> None:
    _check(io_lib().sehf_io_finalize(scan_dir.encode()), "sehf_io_finalize")


def _c64(a, shape) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def write_energy(scan_dir: str, meta: ScanMeta, ik: int, res: dict) -> None:
    """Write one energy from a Scan.run result (matrices as returned there:
    A, B (n_ch, n_ch); b_overlap (n_ch, n_occ); d_raw (6, n_ch); D_raw, D
    complex (6, n_ch); d_corr (6, n_occ); diag[0] = fit_residual_rel,
    diag[2] = K_fit symmetry error)."""
    N = res["A"].shape[0]
    no = res["b_overlap"].shape[1]
    A = _c64(res["A"].T, (N, N))                  # column-major A
    B = _c64(res["B"].T, (N, N))
    bo = _c64(res["b_overlap"].T, (no, N))        # [n_occ][n_ch]
    dr = _c64(res["d_raw"], (6, N))
    DR = _c64(np.ascontiguousarray(res["D_raw"], dtype=np.complex128).view(np.float64), (6, 2 * N))
    D = _c64(np.ascontiguousarray(res["D"], dtype=np.complex128).view(np.float64), (6, 2 * N))
    dc = _c64(res["d_corr"], (6, no))
    diag = np.asarray(res["diag"], dtype=np.float64)
    m = meta.c()
    p = [x.ctypes.data_as(_dp) for x in (A, B, bo, dr, DR, D, dc)]
    _check(io_lib().sehf_io_write_energy(scan_dir.encode(), ctypes.byref(m), int(ik), N, no, *p,
                                         float(diag[0]), float(diag[2])), "sehf_io_write_energy")


def ik_of(k: float, dk: float) -> int:
    ik = int(round(k / dk))
    if abs(ik * dk - k) > 1e-12 * max(1.0, abs(k)):
        raise ValueError(f"k = {k} is not on the grid k = ik * {dk}")
    return ik


def write_scan(scan_dir: str, meta: ScanMeta, results: Sequence[dict], skip_existing: bool = True) -> list:
    """Manifest, one file per energy (skipping those already written, as a
    resumed reference scan does, src/main.cpp:847-877) and the __SUCCESS__
    marker.  Returns the energy indices written."""
    open_scan(scan_dir, meta)
    done = []
    for r in results:
        ik = ik_of(r["k"], meta.dk)
        if skip_existing and has_energy(scan_dir, ik):
            continue
        write_energy(scan_dir, meta, ik, r)
        done.append(ik)
    finalize(scan_dir)
    return done


# ---------------------------------------------------------------------------
# Reader: the HDF5 C library through ctypes.
# ---------------------------------------------------------------------------
_h5 = None
_hid = ctypes.c_int64
H5T_INTEGER, H5T_FLOAT, H5T_STRING = 0, 1, 3


def _hdf5():
    global _h5
    if _h5 is None:
        L = ctypes.CDLL(HDF5_LIB)
        L.H5open()
        for fn, res, args in [
            ("H5Fopen", _hid, [ctypes.c_char_p, ctypes.c_uint, _hid]), ("H5Fclose", ctypes.c_int, [_hid]),
            ("H5Gopen2", _hid, [_hid, ctypes.c_char_p, _hid]), ("H5Gclose", ctypes.c_int, [_hid]),
            ("H5Dopen2", _hid, [_hid, ctypes.c_char_p, _hid]), ("H5Dclose", ctypes.c_int, [_hid]),
            ("H5Dget_space", _hid, [_hid]), ("H5Dget_type", _hid, [_hid]),
            ("H5Dread", ctypes.c_int, [_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p]),
            ("H5Sget_simple_extent_ndims", ctypes.c_int, [_hid]),
            ("H5Sget_simple_extent_dims", ctypes.c_int, [_hid, ctypes.POINTER(ctypes.c_uint64), ctypes.c_void_p]),
            ("H5Sclose", ctypes.c_int, [_hid]), ("H5Aopen", _hid, [_hid, ctypes.c_char_p, _hid]),
            ("H5Aget_type", _hid, [_hid]), ("H5Aread", ctypes.c_int, [_hid, _hid, ctypes.c_void_p]),
            ("H5Aclose", ctypes.c_int, [_hid]), ("H5Tget_class", ctypes.c_int, [_hid]),
            ("H5Tget_size", ctypes.c_size_t, [_hid]), ("H5Tget_strpad", ctypes.c_int, [_hid]),
            ("H5Tclose", ctypes.c_int, [_hid]), ("H5Lexists", ctypes.c_int, [_hid, ctypes.c_char_p, _hid]),
            ("H5Eset_auto2", ctypes.c_int, [_hid, ctypes.c_void_p, ctypes.c_void_p]),
        ]:
            f = getattr(L, fn)
            f.restype, f.argtypes = res, args
        L.H5Eset_auto2(0, None, None)
        _h5 = L
    return _h5


def _native(name: str) -> int:
    return _hid.in_dll(_hdf5(), name + "_g").value


def _ok(h: int, what: str) -> int:
    if h < 0:
        raise RuntimeError(f"HDF5: {what} failed")
    return h


def _type_info(L, t: int) -> tuple:
    cls, size = L.H5Tget_class(t), L.H5Tget_size(t)
    pad = L.H5Tget_strpad(t) if cls == H5T_STRING else None
    return cls, int(size), pad


def _read_dset(L, loc: int, path: str):
    ds = _ok(L.H5Dopen2(loc, path.encode(), 0), path)
    try:
        sp = L.H5Dget_space(ds)
        nd = L.H5Sget_simple_extent_ndims(sp)
        dims = (ctypes.c_uint64 * max(nd, 1))()
        L.H5Sget_simple_extent_dims(sp, dims, None)
        L.H5Sclose(sp)
        shape = tuple(int(dims[i]) for i in range(nd))
        t = L.H5Dget_type(ds)
        info = _type_info(L, t)
        L.H5Tclose(t)
        if info[0] == H5T_INTEGER:
            out, mt = np.zeros(shape, dtype=np.int32), _native("H5T_NATIVE_INT")
        else:
            out, mt = np.zeros(shape, dtype=np.float64), _native("H5T_NATIVE_DOUBLE")
        if out.size:
            _ok(L.H5Dread(ds, mt, 0, 0, 0, out.ctypes.data), path)
        return out, info
    finally:
        L.H5Dclose(ds)


def _read_attr(L, loc: int, name: str):
    at = _ok(L.H5Aopen(loc, name.encode(), 0), name)
    try:
        t = L.H5Aget_type(at)
        info = _type_info(L, t)
        if info[0] == H5T_STRING:
            buf = ctypes.create_string_buffer(info[1])
            _ok(L.H5Aread(at, t, buf), name)
            val = buf.raw.rstrip(b"\0").decode()
        elif info[0] == H5T_INTEGER:
            v = ctypes.c_longlong()
            _ok(L.H5Aread(at, _native("H5T_NATIVE_LLONG"), ctypes.byref(v)), name)
            val = v.value
        else:
            v = ctypes.c_double()
            _ok(L.H5Aread(at, _native("H5T_NATIVE_DOUBLE"), ctypes.byref(v)), name)
            val = v.value
        L.H5Tclose(t)
        return val, info
    finally:
        L.H5Aclose(at)


class _File:
    def __init__(self, path: str):
        self.L = _hdf5()
        self.f = _ok(self.L.H5Fopen(path.encode(), 0, 0), "H5Fopen " + path)
        self.types = {}

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.L.H5Fclose(self.f)

    def exists(self, path: str) -> bool:
        return self.L.H5Lexists(self.f, path.encode(), 0) > 0

    def dset(self, path: str):
        v, info = _read_dset(self.L, self.f, path)
        self.types[path] = (info, v.shape)
        return v

    def attrs(self, group: str, names: Sequence[str]) -> dict:
        g = _ok(self.L.H5Gopen2(self.f, group.encode(), 0), group)
        try:
            out = {}
            for n in names:
                out[n], info = _read_attr(self.L, g, n)
                self.types[group + "@" + n] = (info, ())
            return out
        finally:
            self.L.H5Gclose(g)


def read_manifest(scan_dir: str) -> dict:
    """DipoleReader::read_manifest_ (DipoleIO.cpp:431-475); 'types' maps each
    object to ((class, size, string padding), shape) as stored."""
    with _File(os.path.join(scan_dir, "manifest.h5")) as F:
        m = {}
        m.update(F.attrs("/grid", ["r_min", "dr", "N_grid", "l_max_continuum", "E_HOMO"]))
        m.update(F.attrs("/kgrid", ["dk", "ik_min", "ik_max"]))
        if F.exists("/occ"):
            m["occ_energies"] = F.dset("/occ/energies")
            m["occ_spin_factors"] = F.dset("/occ/spin_factors")
        if F.exists("/atoms"):
            m["atom_Z"] = F.dset("/atoms/Z")
            m["atom_xyz"] = F.dset("/atoms/xyz_bohr")
        m["conventions"] = F.attrs("/conventions", ["real_Ylm_q_map", "psi_norm", "u_convention",
                                                    "gauge_cross_section"])
        if F.exists("/run"):
            m.update(F.attrs("/run", ["molecule_name", "git_hash", "iso_date_utc", "psi_dir_prefix"]))
        m["types"] = F.types
    return m


def available_ik(scan_dir: str) -> list:
    """DipoleReader::available_ik (DipoleIO.cpp:477-490)."""
    out = []
    for n in os.listdir(scan_dir):
        if n.startswith("ik") and n.endswith(".h5") and len(n) > 5:
            try:
                out.append(int(n[2:-3]))
            except ValueError:
                pass
    return sorted(out)


def read_energy(scan_dir: str, ik: int) -> dict:
    """DipoleReader::read_energy (DipoleIO.cpp:492-533): /meta attributes,
    A, B, b_overlap and the six slices in slot order L x, L y, L z, V x, V y,
    V z (D, D_raw complex; d_raw; d_correction; partial_sigma)."""
    with _File(os.path.join(scan_dir, "ik%04d.h5" % ik)) as F:
        r = F.attrs("/meta", ["ik", "k", "E", "omega", "fit_residual_rel", "K_symmetry_err"])
        r["A"], r["B"], r["b_overlap"] = F.dset("/A"), F.dset("/B"), F.dset("/b_overlap")
        D, DR, dr, dc, ps = [], [], [], [], []
        for g in ("length", "velocity"):
            for p in ("x", "y", "z"):
                base = f"/dipole/{g}/{p}"
                D.append(F.dset(base + "/D_ortho_re") + 1j * F.dset(base + "/D_ortho_im"))
                DR.append(F.dset(base + "/D_raw_re") + 1j * F.dset(base + "/D_raw_im"))
                dr.append(F.dset(base + "/d_raw"))
                dc.append(F.dset(base + "/d_correction"))
                ps.append(F.attrs(base, ["partial_sigma"])["partial_sigma"])
        r.update({"D": np.array(D), "D_raw": np.array(DR), "d_raw": np.array(dr), "d_corr": np.array(dc),
                  "partial_sigma": np.array(ps), "types": F.types})
    return r


# ---------------------------------------------------------------------------
# Post-processing front end: the per-channel .dat files of
# postprocessing/gather_dipoles.py, and the files as Scan.run-shaped results.
# ---------------------------------------------------------------------------
def idx_to_lm(mu: int) -> tuple:
    """Channel index of the real-Y basis, mu = l^2 + l + m -> (l, m)."""
    l = int(np.floor(np.sqrt(mu)))
    while (l + 1) * (l + 1) <= mu:
        l += 1
    while l * l > mu:
        l -= 1
    return l, mu - l * l - l


def load_results(scan_dir: str, use_raw: bool = False) -> tuple:
    """(manifest, results): one dict per written energy in ascending ik with
    k, E, omega and D (6, n_ch) complex (D_raw instead when use_raw), the
    input sehf.observables.from_scan takes."""
    meta = read_manifest(scan_dir)
    out = []
    for ik in available_ik(scan_dir):
        r = read_energy(scan_dir, ik)
        out.append({"ik": ik, "k": r["k"], "E": r["E"], "omega": r["omega"], "D": r["D_raw"] if use_raw else r["D"]})
    if not out:
        raise RuntimeError(f"no ikNNNN.h5 files under {scan_dir}")
    n = out[0]["D"].shape[1]
    for r in out:
        if r["D"].shape[1] != n:
            raise RuntimeError(f"channel count differs between ik files (ik={r['ik']}: {r['D'].shape[1]} vs {n})")
    return meta, out


def gather(scan_dir: str, output_dir: str, use_raw: bool = False) -> int:
    """Write dipole_{len,vel}_homo_{x,y,z}_{mu}.dat, one file per gauge,
    polarisation and channel, with columns k_elec, E_kin, omega, Re D, Im D,
    |D|^2, arg D over the energies (postprocessing/gather_dipoles.py:148-209),
    so its consumer postprocessing/cross_section_delay.py reads them
    unchanged.  Returns the number of files written."""
    meta, res = load_results(scan_dir, use_raw)
    os.makedirs(output_dir, exist_ok=True)
    k = np.array([r["k"] for r in res])
    E = np.array([r["E"] for r in res])
    om = np.array([r["omega"] for r in res])
    D = np.stack([r["D"] for r in res])          # [n_e, 6, n_ch]
    head = (f"scan_dir: {scan_dir}\n"
            f"dk = {meta['dk']:.12g} au   E_HOMO = {meta['E_HOMO']:.6f} au   "
            f"Ip = {-meta['E_HOMO']:.6f} au   dipole = {'RAW (pre-ortho)' if use_raw else 'ORTHO'}\n"
            "Columns: k_elec[au]  E_kin[au]  omega[au]  Re(D)  Im(D)  |D|^2  arg(D)")
    n = 0
    for s, (prefix, pol) in enumerate([(g, p) for g in ("dipole_len_homo", "dipole_vel_homo") for p in "xyz"]):
        for mu in range(D.shape[2]):
            l, m = idx_to_lm(mu)
            d = D[:, s, mu]
            cols = np.column_stack([k, E, om, d.real, d.imag, np.abs(d) ** 2, np.angle(d)])
            np.savetxt(os.path.join(output_dir, f"{prefix}_{pol}_{mu}.dat"), cols,
                       header=head + f"\nmu = {mu}  (l={l}, m={m:+d})", fmt="%.14e")
            n += 1
    return n


def _main(argv=None) -> None:
    import argparse
    ap = argparse.ArgumentParser(description="Gather per-channel dipole files from a scan directory.")
    ap.add_argument("scan_dir")
    ap.add_argument("--output-dir", default=None, help="default: ./gathered_<scan_dir name>")
    ap.add_argument("--raw", action="store_true", help="D_raw (before orthogonalisation) instead of D_ortho")
    a = ap.parse_args(argv)
    out = a.output_dir or os.path.join(os.getcwd(), "gathered_" + os.path.basename(os.path.normpath(a.scan_dir)))
    n = gather(a.scan_dir, out, a.raw)
    print(f"wrote {n} .dat files to {out}")


if __name__ == "__main__":
    _main()
