/* sehf_io.h -- persistent per-energy output of a scan in the layout of the
 * reference's DipoleWriter (src/scatt/DipoleIO.hpp:1-25, DipoleIO.cpp:111-340):
 *
 *   <scan_dir>/manifest.h5   scan-level metadata, written once
 *   <scan_dir>/ik<NNNN>.h5   one energy: /meta attributes, /A, /B, /b_overlap,
 *                            /dipole/{length,velocity}/{x,y,z}/{D_ortho_re,
 *                            D_ortho_im, D_raw_re, D_raw_im, d_raw,
 *                            d_correction} + attribute partial_sigma
 *   <scan_dir>/__SUCCESS__   zero-byte marker written by sehf_io_finalize
 *
 * Host-only (HDF5 C library); built as libsehf_io.so beside libsehf.so.
 * Every entry point returns 0 on success and -1 on failure, with the reason
 * in sehf_io_last_error() (the reference throws std::runtime_error there). */
#ifndef SEHF_IO_H
#define SEHF_IO_H

#ifdef __cplusplus
extern "C" {
#endif

/* DipoleScanMeta (DipoleIO.hpp:42-79).  Energy grid k(ik) = ik * dk,
 * E(ik) = k^2 / 2 (EnergyGrid.hpp:28-29).  Any string may be NULL (empty). */
typedef struct {
    double r_min, dr;
    long long N_grid;
    int l_max_continuum;
    double dk;
    int ik_min, ik_max;           /* ik_max exclusive */
    double E_HOMO;
    int n_occ;                    /* 0: no /occ group */
    const double* occ_energies;   /* [n_occ] */
    const double* occ_spin_factors;
    int n_atoms;                  /* 0: no /atoms group */
    const int* atom_Z;            /* [n_atoms] */
    const double* atom_xyz;       /* [n_atoms][3], bohr */
    const char* molecule_name;
    const char* git_hash;
    const char* iso_date_utc;
    const char* psi_dir_prefix;
} sehf_io_meta;

/* Creates scan_dir and writes manifest.h5 unless it exists (a resumed scan
 * keeps its manifest): DipoleWriter::DipoleWriter, DipoleIO.cpp:111-131,
 * write_manifest_ DipoleIO.cpp:141-252. */
int sehf_io_open(const char* scan_dir, const sehf_io_meta* meta);

/* 1 if ik<NNNN>.h5 exists, 0 if not, -1 on a bad argument
 * (DipoleWriter::has_energy, DipoleIO.cpp:137-139). */
int sehf_io_has_energy(const char* scan_dir, int ik);

/* One energy (DipoleEnergyPayload, DipoleIO.hpp:96-112) in the layouts that
 * sehf_scan_out holds for one energy (n_ch = N_psi):
 *   A, B       [n_ch*n_ch] column-major
 *   b_overlap  [n_occ][n_ch]       (b_overlap[a][beta] = <phi_a|psi_beta>)
 *   d_raw      [6][n_ch]
 *   D_raw, D   [6][n_ch] complex interleaved (D = the orthogonalised d_lm)
 *   d_corr     [6][n_occ]
 * slot order L x, L y, L z, V x, V y, V z (DipoleWriter::slice_index,
 * DipoleIO.hpp:132-139).  partial_sigma of each slice is sum |D|^2
 * (DipoleMatrixElement.cpp:930-947).  Written to a per-process temporary
 * file and renamed (DipoleWriter::write_energy, DipoleIO.cpp:254-335). */
int sehf_io_write_energy(const char* scan_dir, const sehf_io_meta* meta, int ik, int n_ch, int n_occ,
                         const double* A, const double* B, const double* b_overlap, const double* d_raw,
                         const double* D_raw, const double* D, const double* d_corr, double fit_residual_rel,
                         double K_symmetry_err);

/* Writes the __SUCCESS__ marker (DipoleWriter::finalize, DipoleIO.cpp:337-339). */
int sehf_io_finalize(const char* scan_dir);

const char* sehf_io_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SEHF_IO_H */
