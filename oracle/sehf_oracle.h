/* sehf_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * C entry points of the CPU FP64 oracle: a restatement of the reference's
 * energy-dependent static-exchange photoionization path (Potentials ->
 * SchurInverter -> ForwardRPropagator -> KMatrixExtractor -> BackPropagator
 * -> AsymptoticAmplitudes -> DipoleMatrixElement).  Only tests/, the smoke
 * check and bench.py's cpu_baseline leg load this library, as the checker.
 * The product (spion-emol_amd/) never links, imports or calls it.
 *
 * Conventions: all matrices are column-major FP64 (Eigen's default, as in
 * the reference); "row-major" inputs are named as such.
 */
#ifndef SEHF_ORACLE_H
#define SEHF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Molecule + grid description (mirrors io::PreprocData, HDF5Reader.hpp:42-86). */
typedef struct {
    int     l_cont;          /* continuum l_max (Parameters::l_max_continuum)         */
    int     n_occ;           /* occupied orbitals used in exchange + orthogonalisation */
    int     Nr;              /* radial points, r_k = k*dr, k in [0,Nr)                 */
    double  dr;
    int     n_atoms;
    const int*    Z;         /* [n_atoms]                                              */
    const double* xyz;       /* [n_atoms*3] bohr, SCE origin at 0                      */
    int     nlm_vh;          /* rows of V_H  (>= (2*l_cont+1)^2)                       */
    const double* V_H;       /* [nlm_vh][Nr] row-major                                 */
    int     n_orb;           /* orbitals stored (>= n_occ)                             */
    int     nlm_orb;         /* SCE channels stored per orbital                        */
    const double* psi_lm;    /* [n_orb*nlm_orb][Nr] row-major, F (no r factor)         */
    int     has_init;        /* 1: initial state from init_psi_lm (no r factor)        */
    const double* init_psi_lm; /* [nlm_orb][Nr] row-major                              */
    int     pol_kind;        /* 0 none, 1 damped -a/2 (1-e^{-(r/1.5)^6})^2/r^4,
                                2 soft -a/(2 (r^2+rc^2)^2)                              */
    double  pol_alpha;
    double  pol_rc;
    int     allow_closed;    /* AsymptoticAmplitudes::Config::allow_closed_channels    */
} sehf_oracle_problem;

/* Per-energy outputs.  Any pointer may be NULL.  N = (l_cont+1)^2. */
typedef struct {
    double* K_kme;        /* [N*N]  KMatrixExtractor K (symmetrised)                  */
    double* A;            /* [N*N]  AsymptoticAmplitudes A                            */
    double* B;            /* [N*N]                                                    */
    double* K_fit;        /* [N*N]  B*A^{-1} (not symmetrised)                        */
    double* b_overlap;    /* [N*n_occ]                                                */
    double* d_raw;        /* [6*N]  slot-major (L x,y,z, V x,y,z)                     */
    double* D_raw;        /* [6*N*2] complex interleaved                              */
    double* D;            /* [6*N*2] complex interleaved: the reduced dipole d_lm     */
    double* d_corr;       /* [6*n_occ]                                                */
    double* diag;         /* [16]: fit_residual_rel, fit_residual_max, K_fit_sym_err,
                             A_minus_I, n_tr, n_start, n_fit_lo, n_fit_hi, shiftsA,
                             shiftsS, n_closed_l, unitarity_err_kme, S_fit_unit_err  */
    /* optional stage dumps: indices + buffers (N_psi^2 / N_tot^2 per index) */
    int  n_dump_sinv; const int* dump_sinv_n; double* dump_sinv;
    int  n_dump_rinv; const int* dump_rinv_n; double* dump_rinv;
    int  n_dump_psi;  const int* dump_psi_n;  double* dump_psi;
    double* t_stage;      /* [8] seconds: setup, sinv, frp, kme, bp, fit, dipole, total */
    /* tangent along (E, alpha): derivative of every output below with respect
     * to E times tan_E plus with respect to the polarisation strength alpha
     * times tan_alpha (computed only when tangent != 0) */
    int     tangent;
    double  tan_E, tan_alpha;
    double* dK;           /* [N*N]  of K_kme                                        */
    double* dA;           /* [N*N]                                                  */
    double* dB;           /* [N*N]                                                  */
    double* db_overlap;   /* [N*n_occ]                                              */
    double* dd_raw;       /* [6*N]                                                  */
    double* dD;           /* [6*N*2] complex interleaved                            */
    double* dump_dpsi;    /* [n_dump_psi*N*N] dpsi at dump_psi_n                    */
    double* dump_drinv;   /* [n_dump_rinv*N_tot^2] dRinv at dump_rinv_n             */
} sehf_oracle_out;

/* Solve one energy point E = k^2/2.  Returns 0 on success, else writes a
 * message retrievable with sehf_oracle_last_error(). */
int sehf_oracle_solve(const sehf_oracle_problem* p, double k, sehf_oracle_out* out);

/* Time a bounded sample: runs the per-n bodies of Sinv/FRP/BP for n_steps
 * radial points at the full problem size (no storage of the whole grid) and
 * writes per-step seconds: t[0]=Sinv, t[1]=FRP, t[2]=BP(incl. W^-1 apply). */
int sehf_oracle_time_steps(const sehf_oracle_problem* p, double k, int n_steps, double* t);

/* The same sample for nk energies at once: one energy-independent setup,
 * then the energies run concurrently (one OpenMP thread each).  t[3*e+i] as
 * above per energy; wall[0] = setup seconds, wall[1] = wall seconds of the
 * concurrent steps. */
int sehf_oracle_time_steps_multi(const sehf_oracle_problem* p, const double* ks, int nk, int n_steps,
                                 double* t, double* wall);

const char* sehf_oracle_last_error(void);

/* Rounding-order probes, bit flags (0 = the reference's order, default):
 * bit 0: Gauss-Jordan with partial pivoting in place of partial-pivot LU for
 *        the dense inverses;
 * bit 1: the asymptotic-fit window sums accumulated from the outer end
 *        inwards (the order of a back-propagation sweep);
 * bits 2, 3, 4: Gauss-Jordan for the Schur inverses, the forward ratio
 *        steps or the K-matrix inverses only;
 * bit 5: the partial-pivot LU in 80-bit extended precision;
 * bit 6: the W^{-1} application's product sums in reverse order.
 * Each is a second, equally valid rounding order of the reference algorithm,
 * used by tests only to measure its rounding noise floor.  Process-wide; not
 * thread-safe against concurrent solves. */
void sehf_oracle_set_inverse_mode(int mode);

/* Threads (default 1) for the loops over independent output columns inside
 * one solve (dense products, the LU update and the LU solves).  Each
 * column's operations are unchanged, so results are bitwise those of one
 * thread.  Process-wide. */
void sehf_oracle_set_threads(int threads);

/* ---- primitives exposed for golden-vector tests ---- */
double sehf_oracle_w3j(int j1, int j2, int j3, int m1, int m2, int m3);
double sehf_oracle_gaunt_real(int l1, int m1, int l2, int m2, int l3, int m3);
double sehf_oracle_real_ylm(int l, int m, double theta, double phi);
double sehf_oracle_riccati_j(int l, double x);
double sehf_oracle_riccati_y(int l, double x);
int    sehf_oracle_inverse_general(int n, const double* A, double* Ainv);
double sehf_oracle_min_eig_sym(int n, const double* A);
void   sehf_oracle_five_point_derivative(const double* f, int n, double dr, double* df);
double sehf_oracle_simpson(const double* f, int n_pts, double h);
/* V(ir) (N_psi^2) and Q(ir) (N_psi x N_f) as the reference assembles them. */
int    sehf_oracle_potential(const sehf_oracle_problem* p, int ir, double* V);
int    sehf_oracle_exchange_Q(const sehf_oracle_problem* p, int ir, double* Q);
/* One forward Numerov ratio step (test_gpu_steppers.cpp:85-95 shape):
 * Rinv = sym(inv(12*Winv - 10*I - Rprev)). */
int    sehf_oracle_frp_step(int n, const double* Winv, const double* Rprev, double* Rinv);

#ifdef __cplusplus
}
#endif
#endif
