// sehf_io.cpp -- per-energy HDF5 output of a scan in the reference's
// DipoleWriter layout (src/scatt/DipoleIO.cpp:111-340).  Host code over the
// HDF5 C library; no device code and no dependency on libsehf.
//
// Every HDF5 handle is owned by a small RAII wrapper, every call's status is
// checked, and each file is written under a per-process temporary name and
// renamed into place, so a reader never sees a partial file and two writers
// of the same scan directory do not collide (DipoleIO.cpp:142-156, 270-275).
#include "../../include/sehf_io.h"

#include <hdf5.h>
#include <unistd.h>

#include <cstdio>
#include <filesystem>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace fs = std::filesystem;

namespace {

thread_local std::string g_err;

struct H {   // an HDF5 identifier and the call that releases it
    hid_t id;
    herr_t (*close)(hid_t);
    H(hid_t i, herr_t (*c)(hid_t), const std::string& what) : id(i), close(c) {
        if (id < 0) throw std::runtime_error("HDF5: " + what + " failed");
    }
    H(const H&) = delete;
    H& operator=(const H&) = delete;
    ~H() { if (id >= 0) close(id); }
};

void ok(herr_t st, const std::string& what) {
    if (st < 0) throw std::runtime_error("HDF5: " + what + " failed");
}

std::string str(const char* s) { return s ? std::string(s) : std::string(); }

std::string tag(int ik) {   // "ik0007" (EnergyGrid::tag, EnergyGrid.hpp:33-38)
    if (ik < 0) throw std::invalid_argument("energy index ik < 0");
    std::string s = std::to_string(ik);
    if (s.size() < 4) s.insert(0, 4 - s.size(), '0');
    return "ik" + s;
}

void attr_f64(hid_t loc, const char* name, double v) {
    H sp(H5Screate(H5S_SCALAR), H5Sclose, name);
    H at(H5Acreate2(loc, name, H5T_NATIVE_DOUBLE, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose, name);
    ok(H5Awrite(at.id, H5T_NATIVE_DOUBLE, &v), name);
}

void attr_i64(hid_t loc, const char* name, long long v) {
    H sp(H5Screate(H5S_SCALAR), H5Sclose, name);
    H at(H5Acreate2(loc, name, H5T_NATIVE_LLONG, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose, name);
    ok(H5Awrite(at.id, H5T_NATIVE_LLONG, &v), name);
}

// Fixed-length, null-padded string attribute of the string's own length
// (length 1 for the empty string).
void attr_str(hid_t loc, const char* name, const std::string& s) {
    H tp(H5Tcopy(H5T_C_S1), H5Tclose, name);
    ok(H5Tset_size(tp.id, s.empty() ? 1 : s.size()), name);
    ok(H5Tset_strpad(tp.id, H5T_STR_NULLPAD), name);
    H sp(H5Screate(H5S_SCALAR), H5Sclose, name);
    H at(H5Acreate2(loc, name, tp.id, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose, name);
    const char zero = '\0';
    ok(H5Awrite(at.id, tp.id, s.empty() ? &zero : s.data()), name);
}

void dset(hid_t loc, const char* path, hid_t type, int rank, const hsize_t* dims, const void* data) {
    H sp(H5Screate_simple(rank, dims, nullptr), H5Sclose, path);
    H ds(H5Dcreate2(loc, path, type, sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose, path);
    ok(H5Dwrite(ds.id, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), path);
}

void vec_f64(hid_t loc, const char* path, const double* v, size_t n) {
    const hsize_t d[1] = {n};
    dset(loc, path, H5T_NATIVE_DOUBLE, 1, d, v);
}

// A matrix stored (rows, cols) in row order, the orientation numpy and h5py
// read (DipoleIO.cpp:77-93); `at(i, j)` gives element (i, j).
template <class F>
void mat_f64(hid_t loc, const char* path, size_t rows, size_t cols, F at) {
    std::vector<double> buf(rows * cols);
    for (size_t i = 0; i < rows; ++i)
        for (size_t j = 0; j < cols; ++j) buf[i * cols + j] = at(i, j);
    const hsize_t d[2] = {rows, cols};
    dset(loc, path, H5T_NATIVE_DOUBLE, 2, d, buf.data());
}

hid_t group(hid_t loc, const char* path) { return H5Gcreate2(loc, path, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT); }

// Write `body` into `final_path` through a per-process temporary file.  If
// the rename fails but another writer has produced the file, that copy is
// kept (every writer of a scan directory writes the same content).
template <class F>
void write_atomic(const std::string& final_path, F body) {
    const std::string tmp = final_path + ".tmp.pid" + std::to_string(static_cast<long long>(::getpid()));
    try {
        H f(H5Fcreate(tmp.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose, "H5Fcreate " + tmp);
        body(f.id);
    } catch (...) {   // no partial file is left behind
        std::error_code rm;
        fs::remove(tmp, rm);
        throw;
    }
    std::error_code ec;
    fs::rename(tmp, final_path, ec);
    if (ec) {
        std::error_code rm;
        fs::remove(tmp, rm);
        if (!fs::exists(final_path))
            throw std::runtime_error("cannot rename " + tmp + " -> " + final_path + ": " + ec.message());
    }
}

void check_meta(const sehf_io_meta* m) {
    if (!m) throw std::invalid_argument("meta is NULL");
    if (!(m->dk > 0.0)) throw std::invalid_argument("meta.dk must be > 0");
    if (m->ik_min < 0) throw std::invalid_argument("meta.ik_min < 0");
    if (m->ik_max <= m->ik_min) throw std::invalid_argument("meta.ik_max <= meta.ik_min");
    if (m->n_occ < 0 || (m->n_occ > 0 && (!m->occ_energies || !m->occ_spin_factors)))
        throw std::invalid_argument("meta: n_occ > 0 needs occ_energies and occ_spin_factors");
    if (m->n_atoms < 0 || (m->n_atoms > 0 && (!m->atom_Z || !m->atom_xyz)))
        throw std::invalid_argument("meta: n_atoms > 0 needs atom_Z and atom_xyz");
}

void manifest(hid_t f, const sehf_io_meta& m) {
    {
        H g(group(f, "/grid"), H5Gclose, "/grid");
        attr_f64(g.id, "r_min", m.r_min);
        attr_f64(g.id, "dr", m.dr);
        attr_i64(g.id, "N_grid", m.N_grid);
        attr_i64(g.id, "l_max_continuum", m.l_max_continuum);
        attr_f64(g.id, "E_HOMO", m.E_HOMO);
    }
    {
        H g(group(f, "/kgrid"), H5Gclose, "/kgrid");
        attr_f64(g.id, "dk", m.dk);
        attr_i64(g.id, "ik_min", m.ik_min);
        attr_i64(g.id, "ik_max", m.ik_max);
    }
    if (m.n_occ > 0) {
        H g(group(f, "/occ"), H5Gclose, "/occ");
        vec_f64(g.id, "energies", m.occ_energies, static_cast<size_t>(m.n_occ));
        vec_f64(g.id, "spin_factors", m.occ_spin_factors, static_cast<size_t>(m.n_occ));
    }
    if (m.n_atoms > 0) {
        H g(group(f, "/atoms"), H5Gclose, "/atoms");
        const hsize_t dz[1] = {static_cast<hsize_t>(m.n_atoms)};
        dset(g.id, "Z", H5T_NATIVE_INT, 1, dz, m.atom_Z);
        const hsize_t dx[2] = {static_cast<hsize_t>(m.n_atoms), 3};
        dset(g.id, "xyz_bohr", H5T_NATIVE_DOUBLE, 2, dx, m.atom_xyz);
    }
    {
        H g(group(f, "/conventions"), H5Gclose, "/conventions");
        attr_str(g.id, "real_Ylm_q_map", "x=+1, y=-1, z=0");
        attr_str(g.id, "psi_norm", "incoming-wave Psi- = (A - iB)^(-dagger)");
        attr_str(g.id, "u_convention", "chi_lm(r) = r * F_lm(r)");
        attr_str(g.id, "gauge_cross_section",
                 "sigma^L = (4 pi^2 / c) omega Sum_mu |D^L|^2;  sigma^V = (4 pi^2 / c) (1/omega) Sum_mu |D^V|^2");
    }
    {
        H g(group(f, "/run"), H5Gclose, "/run");
        attr_str(g.id, "molecule_name", str(m.molecule_name));
        attr_str(g.id, "git_hash", str(m.git_hash));
        attr_str(g.id, "iso_date_utc", str(m.iso_date_utc));
        attr_str(g.id, "psi_dir_prefix", str(m.psi_dir_prefix));
    }
}

template <class F>
int guarded(F f) {
    try {
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);   // errors come back as status codes
        return f();
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

}  // namespace

extern "C" {

const char* sehf_io_last_error(void) { return g_err.c_str(); }

int sehf_io_open(const char* scan_dir, const sehf_io_meta* meta) {
    return guarded([&] {
        if (!scan_dir || !*scan_dir) throw std::invalid_argument("scan_dir is empty");
        check_meta(meta);
        std::error_code ec;
        fs::create_directories(scan_dir, ec);
        if (ec) throw std::runtime_error(std::string("cannot create ") + scan_dir + ": " + ec.message());
        const std::string mpath = std::string(scan_dir) + "/manifest.h5";
        if (!fs::exists(mpath)) write_atomic(mpath, [&](hid_t f) { manifest(f, *meta); });
        return 0;
    });
}

int sehf_io_has_energy(const char* scan_dir, int ik) {
    return guarded([&] {
        if (!scan_dir) throw std::invalid_argument("scan_dir is NULL");
        return fs::exists(std::string(scan_dir) + "/" + tag(ik) + ".h5") ? 1 : 0;
    });
}

int sehf_io_write_energy(const char* scan_dir, const sehf_io_meta* meta, int ik, int n_ch, int n_occ,
                         const double* A, const double* B, const double* b_overlap, const double* d_raw,
                         const double* D_raw, const double* D, const double* d_corr, double fit_residual_rel,
                         double K_symmetry_err) {
    return guarded([&] {
        if (!scan_dir) throw std::invalid_argument("scan_dir is NULL");
        check_meta(meta);
        if (n_ch <= 0 || n_occ < 0) throw std::invalid_argument("n_ch must be > 0 and n_occ >= 0");
        if (!A || !B || !d_raw || !D_raw || !D || (n_occ > 0 && (!b_overlap || !d_corr)))
            throw std::invalid_argument("a payload array is NULL");
        const std::string path = std::string(scan_dir) + "/" + tag(ik) + ".h5";
        const size_t N = static_cast<size_t>(n_ch), no = static_cast<size_t>(n_occ);
        const double k = ik * meta->dk, E = 0.5 * k * k;
        write_atomic(path, [&](hid_t f) {
            {
                H g(group(f, "/meta"), H5Gclose, "/meta");
                attr_i64(g.id, "ik", ik);
                attr_f64(g.id, "k", k);
                attr_f64(g.id, "E", E);
                attr_f64(g.id, "omega", E - meta->E_HOMO);
                attr_f64(g.id, "fit_residual_rel", fit_residual_rel);
                attr_f64(g.id, "K_symmetry_err", K_symmetry_err);
            }
            mat_f64(f, "/A", N, N, [&](size_t i, size_t j) { return A[i + j * N]; });
            mat_f64(f, "/B", N, N, [&](size_t i, size_t j) { return B[i + j * N]; });
            mat_f64(f, "/b_overlap", N, no, [&](size_t i, size_t j) { return b_overlap[j * N + i]; });
            { H g(group(f, "/dipole"), H5Gclose, "/dipole"); }
            static const char* gauge[2] = {"length", "velocity"};
            static const char* pol[3] = {"x", "y", "z"};
            std::vector<double> re(N), im(N);
            for (int gi = 0; gi < 2; ++gi) {
                const std::string gp = std::string("/dipole/") + gauge[gi];
                { H g(group(f, gp.c_str()), H5Gclose, gp); }
                for (int pi = 0; pi < 3; ++pi) {
                    const int slot = 3 * gi + pi;
                    const std::string pp = gp + "/" + pol[pi];
                    H g(group(f, pp.c_str()), H5Gclose, pp);
                    const double* Dz = D + 2 * N * slot;
                    const double* Rz = D_raw + 2 * N * slot;
                    double sigma = 0.0;
                    for (size_t i = 0; i < N; ++i) {
                        re[i] = Dz[2 * i];
                        im[i] = Dz[2 * i + 1];
                        sigma += re[i] * re[i] + im[i] * im[i];
                    }
                    vec_f64(g.id, "D_ortho_re", re.data(), N);
                    vec_f64(g.id, "D_ortho_im", im.data(), N);
                    for (size_t i = 0; i < N; ++i) { re[i] = Rz[2 * i]; im[i] = Rz[2 * i + 1]; }
                    vec_f64(g.id, "D_raw_re", re.data(), N);
                    vec_f64(g.id, "D_raw_im", im.data(), N);
                    vec_f64(g.id, "d_raw", d_raw + N * slot, N);
                    vec_f64(g.id, "d_correction", no ? d_corr + no * slot : re.data(), no);   // (a valid pointer for 0 elements)
                    attr_f64(g.id, "partial_sigma", sigma);
                }
            }
        });
        return 0;
    });
}

int sehf_io_finalize(const char* scan_dir) {
    return guarded([&] {
        if (!scan_dir) throw std::invalid_argument("scan_dir is NULL");
        const std::string p = std::string(scan_dir) + "/__SUCCESS__";
        std::ofstream o(p);
        if (!o) throw std::runtime_error("cannot write " + p);
        return 0;
    });
}

}  // extern "C"
