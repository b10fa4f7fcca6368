/* sehf.h -- C ABI of the MI355X static-exchange photoionization solver.
 *
 * Two layers, both plain C with caller-owned host buffers, opaque handles and
 * int status codes (0 = success; otherwise sehf_last_error() holds the
 * message, which the C++ shim in INTEGRATION.md turns into std::runtime_error
 * as the reference's CHECK macros do, src/scatt/gpu/cuda/GpuPropagate.cu:66-99).
 *
 *  1. Per-step compatibility layer: one entry point per method of the
 *     reference's vendor-neutral GPU stepper classes, so its host code runs
 *     unchanged with a third backend next to sycl/ and cuda/.
 *  2. Stage-level scan layer: energy-independent inputs are uploaded once and
 *     a whole k-scan advances through the radial grid on the device; only the
 *     per-energy results (K, A, B, b_overlap, the six dipole slices) return.
 *
 * All matrices are column-major FP64 (Eigen's default layout in the
 * reference).  Complex outputs are interleaved (re, im).
 */
#ifndef SEHF_H
#define SEHF_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Molecule + grid description: the contents of the preprocessing HDF5 that
 * the reference reads into io::PreprocData (src/io/HDF5Reader.hpp:42-86,
 * layouts /grid, /geometry, /potential/V_H, /orbitals/psi_lm,
 * /initial_state/psi_lm) plus the continuum cutoff and the two run options
 * on this path (polarisation potential, --allow-closed-channels). */
typedef struct {
    int     l_cont;            /* Parameters::l_max_continuum                           */
    int     n_occ;             /* occupied orbitals in exchange / orthogonalisation      */
    int     Nr;                /* radial points, r_k = k dr, k in [0, Nr)                */
    double  dr;
    int     n_atoms;
    const int*    Z;           /* [n_atoms]                                              */
    const double* xyz;         /* [n_atoms*3] bohr, SCE origin at 0                      */
    int     nlm_vh;            /* rows of V_H, >= (2 l_cont + 1)^2                       */
    const double* V_H;         /* [nlm_vh][Nr] row-major (/potential/V_H)                */
    int     n_orb;             /* orbitals stored, >= n_occ                              */
    int     nlm_orb;           /* SCE channels per orbital, >= (l_cont+min(l_cont,10)+1)^2 */
    const double* psi_lm;      /* [n_orb*nlm_orb][Nr] row-major (/orbitals/psi_lm)       */
    int     has_init;          /* 1: initial state from init_psi_lm, else the HOMO        */
    const double* init_psi_lm; /* [nlm_orb][Nr] row-major (/initial_state/psi_lm)        */
    int     pol_kind;          /* 0 none; 1 -a/2 (1-e^{-(r/1.5)^6})^2/r^4 (Potentials.cpp:111-129);
                                  2 -a/(2 (r^2+rc^2)^2)                                   */
    double  pol_alpha;
    double  pol_rc;
    int     allow_closed;      /* AsymptoticAmplitudes::Config::allow_closed_channels     */
} sehf_problem;

typedef struct sehf_ctx  sehf_ctx;
typedef struct sehf_frp  sehf_frp;
typedef struct sehf_sinv sehf_sinv;
typedef struct sehf_bp   sehf_bp;
typedef struct sehf_dme  sehf_dme;
typedef struct sehf_scan sehf_scan;

/* Last error message of the calling thread ("" if none). */
const char* sehf_last_error(void);

/* ---- device / context -------------------------------------------------
 * Replaces GpuContext (src/scatt/GpuPropagate.hpp:52-80): static
 * gpu_available(), ctor GpuContext(prefer_gpu), info(). */
int  sehf_gpu_available(void);
int  sehf_ctx_create(int prefer_gpu, sehf_ctx** out);
void sehf_ctx_destroy(sehf_ctx* ctx);
int  sehf_ctx_info(const sehf_ctx* ctx, char* device_name, int name_len, size_t* global_mem_bytes,
                   int* is_gpu);

/* ---- forward stepper --------------------------------------------------
 * Replaces GpuForwardStepper (src/scatt/GpuPropagate.hpp:88-150):
 *   Rinv_n = 1/2 (X + X^T),  X = (12 W_inv - 10 I - Rinv_{n-1})^{-1}.
 * step() keeps Rinv on the device as the next step's Rinv_{n-1}; sym_err
 * (may be NULL) receives max|Rinv - Rinv^T| (step()'s return value). */
int  sehf_frp_create(sehf_ctx* ctx, int N_total, sehf_frp** out);
int  sehf_frp_init_R_prev_inv(sehf_frp* s, const double* R);           /* N_total^2 */
int  sehf_frp_step(sehf_frp* s, const double* W_inv, double* Rinv_out, double* sym_err);
/* enable_symmetric_inverse(bool) request (GpuPropagate.hpp:113-125): the
 * HIP backend always takes the LU-equivalent path, so the request is
 * recorded and is_symmetric_inverse_active() reports 0. */
int  sehf_frp_enable_symmetric_inverse(sehf_frp* s, int enable);
int  sehf_frp_is_symmetric_inverse_active(const sehf_frp* s);
void sehf_frp_destroy(sehf_frp* s);

/* ---- Schur-complement inverse stepper ----------------------------------
 * Replaces GpuSinvStepper (src/scatt/GpuPropagate.hpp:178-227):
 *   Sinv = sym( (A - B diag(Dinv) B^T)^{-1} ),  A: N_psi^2, B: N_psi x N_f. */
int  sehf_sinv_create(sehf_ctx* ctx, int N_psi, int N_f, sehf_sinv** out);
int  sehf_sinv_step(sehf_sinv* s, const double* A, const double* B, const double* Dinv,
                    double* Sinv_out, double* sym_err);
int  sehf_sinv_step_inverse_only(sehf_sinv* s, const double* A, double* Sinv_out, double* sym_err);
void sehf_sinv_destroy(sehf_sinv* s);

/* ---- back stepper -------------------------------------------------------
 * Replaces GpuBackStepper (src/scatt/GpuPropagate.hpp:235-279):
 *   Z := Rinv Z;  Y := W_inv Z;  psi = Y[0:N_psi, :],  f = Y[N_psi:, :].
 * Z is N_total x N_psi and stays on the device between steps. */
int  sehf_bp_create(sehf_ctx* ctx, int N_total, int N_psi, int N_f, sehf_bp** out);
int  sehf_bp_init_Z(sehf_bp* s, const double* Z);
int  sehf_bp_step(sehf_bp* s, const double* Rinv, const double* W_inv, double* psi_out,
                  double* f_out, int compute_f);
int  sehf_bp_get_Z(const sehf_bp* s, double* Z_out);
void sehf_bp_destroy(sehf_bp* s);

/* ---- dipole engine ------------------------------------------------------
 * Replaces GpuDipoleEngine (src/scatt/GpuDipoleEngine.hpp:100-125):
 *   result (n_slots x N_psi) = V (n_slots x N_psi) * psi_ir (N_psi x N_psi). */
int  sehf_dme_create(sehf_ctx* ctx, int N_psi, int n_slots, sehf_dme** out);
int  sehf_dme_step(sehf_dme* s, const double* V, const double* psi_ir, double* result_out);
void sehf_dme_destroy(sehf_dme* s);

#ifdef __cplusplus
}
#endif
#endif /* SEHF_H */
