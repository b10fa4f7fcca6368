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
#include <stdint.h>

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

/* Accumulated per-phase wall time of a stepper's step() calls, in
 * nanoseconds, and the number of steps; cleared by the matching
 * sehf_*_reset_stats.  Field for field the Stats structs of the reference's
 * steppers: GpuForwardStepper::Stats (src/scatt/GpuPropagate.hpp:133-139),
 * GpuSinvStepper::Stats (:211-217), GpuBackStepper::Stats (:263-269) and
 * GpuDipoleEngine::Stats (src/scatt/GpuDipoleEngine.hpp:132-137). */
typedef struct {
    uint64_t t_upload_ns;      /* W_inv host -> device                      */
    uint64_t t_u_combine_ns;   /* U = 12 W_inv - 10 I - Rinv_{n-1}          */
    uint64_t t_inverse_ns;     /* inverse + symmetrisation                  */
    uint64_t t_download_ns;    /* Rinv device -> host (+ symmetry check)    */
    uint64_t n_steps;
} sehf_frp_stats_t;
typedef struct {
    uint64_t t_upload_ns;      /* A, B, Dinv host -> device                 */
    uint64_t t_schur_ns;       /* S = A - B D^{-1} B^T                      */
    uint64_t t_inverse_ns;     /* inverse + symmetrisation                  */
    uint64_t t_download_ns;    /* Sinv device -> host (+ symmetry check)    */
    uint64_t n_steps;
} sehf_sinv_stats_t;
typedef struct {
    uint64_t t_upload_ns;      /* Rinv, W_inv host -> device                */
    uint64_t t_gemm_z_ns;      /* Z := Rinv Z                               */
    uint64_t t_gemm_y_ns;      /* Y := W_inv Z                              */
    uint64_t t_extract_ns;     /* psi, f rows device -> host                */
    uint64_t n_steps;
} sehf_bp_stats_t;
typedef struct {
    uint64_t t_upload_ns;
    uint64_t t_gemm_ns;
    uint64_t t_download_ns;
    uint64_t n_steps;
} sehf_dme_stats_t;

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
/* Forward step without device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Runs the
 * kernels of sehf_frp_step on caller-owned DEVICE buffers (column-major
 * N_total^2 FP64 each): Rinv_out = 1/2 (X + X^T), X = (12 W_inv - 10 I -
 * R_prev)^{-1}; no handle, no state.  The workspace is a fixed static device
 * array (orders up to SEHF_FRP_NOALLOC_MAX_N, about 39 MiB held on each
 * device the library runs on until the process ends); larger orders are
 * rejected.  The call runs on the default stream and returns after it has
 * finished (inputs must be complete on the device before the call).  All
 * calls on one device share one workspace: calls that overlap on one device
 * race and must be serialised by the caller.  sym_err (host, may be NULL)
 * receives max|Rinv - Rinv^T|. */
#define SEHF_FRP_NOALLOC_MAX_N 1600
int  sehf_frp_step_noalloc(int N_total, const double* W_inv, const double* R_prev, double* Rinv_out,
                           double* sym_err);
int  sehf_frp_is_symmetric_inverse_active(const sehf_frp* s);
int  sehf_frp_stats(const sehf_frp* s, sehf_frp_stats_t* out);   /* stats() */
int  sehf_frp_reset_stats(sehf_frp* s);                          /* reset_stats() */
void sehf_frp_destroy(sehf_frp* s);

/* ---- Schur-complement inverse stepper ----------------------------------
 * Replaces GpuSinvStepper (src/scatt/GpuPropagate.hpp:178-227):
 *   Sinv = sym( (A - B diag(Dinv) B^T)^{-1} ),  A: N_psi^2, B: N_psi x N_f. */
int  sehf_sinv_create(sehf_ctx* ctx, int N_psi, int N_f, sehf_sinv** out);
int  sehf_sinv_step(sehf_sinv* s, const double* A, const double* B, const double* Dinv,
                    double* Sinv_out, double* sym_err);
int  sehf_sinv_step_inverse_only(sehf_sinv* s, const double* A, double* Sinv_out, double* sym_err);
int  sehf_sinv_stats(const sehf_sinv* s, sehf_sinv_stats_t* out);
int  sehf_sinv_reset_stats(sehf_sinv* s);
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
int  sehf_bp_stats(const sehf_bp* s, sehf_bp_stats_t* out);
int  sehf_bp_reset_stats(sehf_bp* s);
void sehf_bp_destroy(sehf_bp* s);

/* ---- dipole engine ------------------------------------------------------
 * Replaces GpuDipoleEngine (src/scatt/GpuDipoleEngine.hpp:100-125):
 *   result (n_slots x N_psi) = V (n_slots x N_psi) * psi_ir (N_psi x N_psi). */
int  sehf_dme_create(sehf_ctx* ctx, int N_psi, int n_slots, sehf_dme** out);
int  sehf_dme_step(sehf_dme* s, const double* V, const double* psi_ir, double* result_out);
int  sehf_dme_stats(const sehf_dme* s, sehf_dme_stats_t* out);
int  sehf_dme_reset_stats(sehf_dme* s);
void sehf_dme_destroy(sehf_dme* s);

/* ---- stage-level batched scan ----------------------------------------
 * Replaces the per-ik loop of the scattering driver (src/main.cpp:847-877)
 * and the stage chain inside run_one_energy (main.cpp:1081-1623):
 * SchurInverter -> ForwardRPropagator -> KMatrixExtractor -> BackPropagator
 * -> AsymptoticAmplitudes -> DipoleMatrixElement::compute_six.
 * sehf_scan_create uploads the energy-independent inputs once (V(r), the
 * exchange coupling, dipole tables); sehf_scan_run advances `batch` energies
 * at a time through the radial grid (batch <= 0: all at once) with R
 * checkpoints every `seg_len` points (<= 0: ceil(sqrt(Nr))).
 * Per-energy outputs mirror DipoleEnergyPayload (src/scatt/DipoleIO.hpp:82-112);
 * N = (l_cont+1)^2, slot order L x, L y, L z, V x, V y, V z. */
typedef struct {
    double* K;          /* [nk][N*N]  K from the two-point match (symmetrised)     */
    double* A;          /* [nk][N*N]  fitted amplitude of j-hat                    */
    double* B;          /* [nk][N*N]  fitted amplitude of y-hat                    */
    double* K_fit;      /* [nk][N*N]  B A^{-1}                                     */
    double* b_overlap;  /* [nk][n_occ][N]  int phi_a . psi_beta                    */
    double* d_raw;      /* [nk][6][N]                                              */
    double* D_raw;      /* [nk][6][N] complex interleaved (M d_raw)                */
    double* D;          /* [nk][6][N] complex interleaved: the reduced dipole d_lm */
    double* d_corr;     /* [6][n_occ] energy independent                           */
    double* diag;       /* [nk][16]: fit_residual_rel, fit_residual_max,
                           K_fit symmetry error, max|A - I|, n_transition, n_start,
                           fit window lo, hi, A shifts, S shifts, closed l count  */
} sehf_scan_out;

int  sehf_scan_create(sehf_ctx* ctx, const sehf_problem* p, sehf_scan** out);
/* info[8] = N_psi, N_f, N_total, n_transition, n_start, dipole points, Nr, l_exch */
int  sehf_scan_info(const sehf_scan* s, int* info);
int  sehf_scan_run(sehf_scan* s, const double* k, int nk, int batch, int seg_len, sehf_scan_out* out);

/* Derivatives of the per-energy outputs along a direction t_E d/dE +
 * t_alpha d/dalpha (E = k^2/2 in Hartree, alpha the polarisation strength
 * pol_alpha of sehf_problem).  The reference has no derivative: it obtains
 * dd/dE for the Wigner delay by finite differences of d over the k grid
 * (postprocessing/cross_section_delay.py:227-263), and has no d/dalpha.  The
 * layouts are those of sehf_scan_out; any pointer may be NULL. */
typedef struct {
    double* dK;          /* [nk][N*N]  of K                                         */
    double* dA;          /* [nk][N*N]  of A                                         */
    double* dB;          /* [nk][N*N]  of B                                         */
    double* db_overlap;  /* [nk][n_occ][N]                                          */
    double* dd_raw;      /* [nk][6][N]                                              */
    double* dD;          /* [nk][6][N] complex interleaved: of the reduced d_lm     */
} sehf_scan_tangent_out;

/* sehf_scan_run plus the forward-mode derivative (tangent) of the whole stage
 * chain: the ratio matrices' derivatives are propagated with the forward
 * recursion and checkpointed with them, the back-propagation carries
 * dZ = d(Rinv Z) beside Z, and the K match, fit and dipole normalisation are
 * differentiated in closed form.  dir[2*i], dir[2*i+1] = (t_E, t_alpha) of
 * energy i; the same k may appear twice with two directions. */
int  sehf_scan_run_tangent(sehf_scan* s, const double* k, const double* dir, int nk, int batch, int seg_len,
                           sehf_scan_out* out, sehf_scan_tangent_out* tout);
/* One segment cycle of the scan at full size, for timing: the nk energies
 * (one batch) run the forward ratio steps of radial segment `seg` (points
 * [seg*L, min((seg+1)*L, Nr-1)), L = seg_len or ceil(sqrt(Nr-1)), with the
 * Schur inverses one step ahead on the side stream), that segment's
 * recompute from its start matrix, and its back-propagation with the fused
 * dipole and fit work -- the per-segment work of sehf_scan_run, whose
 * energies go through all segments once in each of the three sweeps.  The
 * forward part and the recompute start from the analytic start matrix, the
 * back-propagation from Z = [I; 0]; nothing is returned but the timing.
 * With dir (as in sehf_scan_run_tangent) the tangent runs alongside.
 * seconds[4] = host tables, forward part, recompute, back-propagation (may be
 * NULL); layout[4] = L, number of segments, first point of the fit window,
 * first segment of the fit pre-pass (may be NULL).  seg < 0 only fills
 * layout.  Replaces no reference interface (the reference runs whole
 * energies). */
int  sehf_scan_run_segment(sehf_scan* s, const double* k, const double* dir, int nk, int seg, int seg_len,
                           double* seconds, int* layout);
/* Autotuner of the batch width, the number of concurrent groups and the
 * GEMM tile policy.  For each candidate (batch in 1, 2, 4, ... 128 with
 * groups * batch <= nk, groups in 1..2) that fits in device memory, the batch
 * buffers are allocated and `probe_steps` radial steps of the forward
 * recursion (Schur inverses and ratio steps at full size, with the tangent
 * when `tangent` is set) are timed, all groups starting together; a
 * candidate list stops growing once the rate gains less than 3%.  At the
 * fastest batch width the other GEMM tile policies (sehf_set_gemm_tiles) are
 * probed as well.  The fastest combination in energy-steps per second is
 * returned; the handle's group count and the process-wide tile policy are
 * set to it.  report[4*i .. 4*i+3] = (batch, groups, tile policy,
 * energy-steps/s) of each probe (at most 32); NULL allowed.  The batch
 * buffers are released on return.  Replaces no reference interface: the
 * reference runs one energy at a time. */
int  sehf_scan_autotune(sehf_scan* s, const double* k, int nk, int probe_steps, int tangent, int* batch_out,
                        int* groups_out, int* tiles_out, double* report, int* n_report);
/* Process-wide tile policy of the FP64 GEMM: 0 (default) by shape, 128 x 128
 * tiles over 8 waves where M > 64 and K >= 256, 64 x 128 over 8 waves where
 * M > 64 and K < 256, 16 x 128 or 32 x 128 over 4 waves where M <= 16 or 32,
 * else 64 x 64 over 4 waves; 1 = 64 x 64 everywhere; 2 = 128 x 128 wherever
 * M > 64, else as 0.  The tile changes which
 * lanes accumulate, not the order of the K sum, so results are the same for
 * every policy. */
int  sehf_set_gemm_tiles(int policy);
/* Process-wide pivot order of the Gauss-Jordan inverses (the ratio matrices,
 * the K match, the fit): 0 (default) tournament pivoting over blocks of 256
 * rows (the panel step of communication-avoiding LU); 1 = partial pivoting,
 * the pivot rows LU with partial pivoting chooses (LapackInverse.hpp:53-92
 * uses Eigen's PartialPivLU), for orders up to 2048 (larger orders keep the
 * tournament), one workgroup per matrix and about 1.3x the inverse time at
 * order 1566.  $SEHF_GJ_LEAF=partial sets 1 at load.  Replaces no reference
 * interface. */
int  sehf_set_gj_pivoting(int mode);
/* t[8]: cumulative seconds of setup, forward sweep, K matrix, back-propagation,
 * fit + dipole, per-batch host tables. */
int  sehf_scan_stats(const sehf_scan* s, double* t);
/* Number of energy groups run concurrently by sehf_scan_run (1..4, default
 * 1 or $SEHF_SCAN_GROUPS): batch i of a run goes to group i % groups; each
 * group owns a HIP stream and its own batch buffers (device memory grows
 * linearly with groups) and is driven by its own host thread.  With the
 * batching it replaces the reference's serial per-ik loop
 * (src/main.cpp:845-877). */
int  sehf_scan_set_groups(sehf_scan* s, int groups);
/* How sehf_scan_run sums the dipole integrals d = sum_n rows_n psi_n.
 * The reference back-propagates Z from the outer boundary over the whole
 * grid, keeps psi_n (BackPropagator.cpp:553-, the loop n = N-1 ... 0) and
 * sums rows_n psi_n over the radial points afterwards
 * (DipoleMatrixElement.cpp:584-720, compute_six).  Since psi_n =
 * P_n Rinv_n ... Rinv_{N-1} Z_N with P_n = [Sinv_n | -Sinv_n B D^{-1}], the
 * same sum is
 *   d = Ad_{N-1} Z_N,   Ad_n = (Ad_{n-1} + rows_n P_n) Rinv_n,  Ad_{-1} = 0,
 * a row block of 6 + n_occ rows carried along the forward recursion.
 *   SEHF_DIPOLES_FORWARD (default): the forward accumulation; the only
 *     back-propagation left is over the fit window at the top of the grid
 *     (the fit sums, then the fit residual), from the checkpoints of the top
 *     segments only.
 *   SEHF_DIPOLES_BACKWARD: the back-propagation over the whole grid, with
 *     checkpoints of every segment and their recompute.
 * Both give the same outputs to rounding (the sums differ in order).  The
 * setting applies to sehf_scan_run; sehf_scan_run_tangent always
 * back-propagates: the tangent of the forward sums carries a ratio matrix's
 * near-pole growth twice (|dRinv_n| ~ |Rinv_n|^2) where the back-propagated
 * dZ carries it once, and measures up to 4x less accurate.  A new handle
 * starts with $SEHF_DIPOLE_PATH (0 or 1) when it is set. */
#define SEHF_DIPOLES_FORWARD 0
#define SEHF_DIPOLES_BACKWARD 1
int  sehf_scan_set_dipole_path(sehf_scan* s, int path);

/* ---- psi files and the dipole-only replay --------------------------------
 * The reference keeps the back-propagated psi_n of every energy on disk
 * (main.cpp:1259-1345): the range the dipole integrals read, [0, n_main)
 * with n_main = min(Nr, max(2, n_transition)), and the 300-point asymptotic
 * tail [Nr - 300, Nr - 1] that the fit reads (BackPropagator.cpp:27-101,
 * psi_asym.bin); --dipole-only (main.cpp:167-189, 1278-1320) then redoes only
 * the fit and the dipoles from them, for instance with another initial
 * state, and the continuum-continuum dipole codes read them.
 * sehf_scan_set_psi_store(s, root, keep) makes every later whole scan
 * (sehf_scan_run, sehf_scan_run_tangent) write, per energy, the directory
 * <root>/psi_ik<NNNN> (NNNN: the energy's index in that call's k list) with
 *   psi_asym.bin  (keep & SEHF_PSI_TAIL) the tail, in the reference's
 *                 single-file layout: 32-byte header 'PSI_ASYM', int32
 *                 version 1, int32 n_asym = min(300, Nr), int32 first index
 *                 Nr - n_asym, int32 N_psi, 8 reserved bytes; then n_asym
 *                 N_psi x N_psi matrices, column-major, ascending n;
 *   psi_main.bin  (keep & SEHF_PSI_MAIN) [0, n_main) in the same layout with
 *                 the magic 'PSI_MAIN' (the reference holds this range in
 *                 its chunked store);
 *   __SUCCESS__   written last ("k <value>" and the sizes, one per line).
 * Files are written under temporary names and renamed when complete.  The
 * tail extends the fit pre-pass down to its first point; SEHF_PSI_MAIN makes
 * the run back-propagate the dipoles (as SEHF_DIPOLES_BACKWARD).  Each
 * point costs a device-to-host copy of the batch's psi_n through a pinned
 * ring and a file write per energy.  root NULL or keep 0 turns it off.
 * sehf_scan_dipole_only(s, root, k, nk, batch, out) reads the files of
 * energies 0..nk-1 under root (each directory's recorded k must equal k[i],
 * N_psi and the tail range must match this problem, psi_main.bin must cover
 * this problem's dipole range) and runs only the fit and the dipoles, with
 * this handle's dipole tables (its initial state): A, B, K_fit, d_raw, D,
 * D_raw, b_overlap, d_corr and diag are written as by sehf_scan_run, bitwise
 * equal to a back-propagated run of this handle; out->K is left untouched
 * (K is not formed).  A missing or incomplete directory fails the call. */
#define SEHF_PSI_TAIL 1
#define SEHF_PSI_MAIN 2
int  sehf_scan_set_psi_store(sehf_scan* s, const char* root, int keep);
int  sehf_scan_dipole_only(sehf_scan* s, const char* root, const double* k, int nk, int batch,
                           sehf_scan_out* out);
void sehf_scan_destroy(sehf_scan* s);

/* ---- device storage plan ------------------------------------------------
 * The device counterpart of the reference's StoragePlanner
 * (src/scatt/StoragePlanner.hpp:106-117, plan_storage), which places the
 * four big per-radial-point objects (pot, sinv, rinv, psi) in RAM or in
 * chunked disk files under a memory budget.  On the device:
 *   pot   the energy-independent S0_n = I - h^2/6 V_n - B_n D_n^{-1} B_n^T
 *         and B_n D_n^{-1}: SEHF_TABLES_RESIDENT keeps them for every n in
 *         HBM (N_psi^2 Nr + N_psi N_f n_transition doubles);
 *         SEHF_TABLES_STREAMED keeps only their compact inputs and rebuilds
 *         them per radial step on the device, on the side stream, with the
 *         resident build's kernels (bitwise the same values);
 *   sinv, rinv  are never stored for every n: ratio matrices are kept at
 *         checkpoints every seg_len points and recomputed per segment
 *         (seg_len x (N_total^2 + N_psi^2) doubles per energy of segment
 *         buffers); the checkpoints live in HBM (SEHF_CKPT_DEVICE) or in
 *         pinned host memory (SEHF_CKPT_HOST), the reference's DISK level;
 *   psi   is not kept on the device (forward-accumulated dipole rows, or
 *         integrands formed during back-propagation); it goes to per-energy
 *         files only when a psi store is set (sehf_scan_set_psi_store).
 * sehf_plan_storage chooses tables, checkpoint level, segment length and
 * batch width: greedily, tables resident while that does not cost batch
 * width, checkpoints on the device while a batch fits, the segment length
 * that minimises bytes per energy when the default one leaves no batch,
 * and a batch of at most min(n_energies, batch_cap) energies (batch_cap <= 0:
 * 64, where the l_cont = 30 scan's throughput stops rising on MI355X).
 * device_bytes = 0 queries the current device's free memory.  Fails when
 * not even one energy fits. */
#define SEHF_TABLES_RESIDENT 0
#define SEHF_TABLES_STREAMED 1
#define SEHF_CKPT_DEVICE 0
#define SEHF_CKPT_HOST 1
typedef struct {
    int    Nr;                /* radial points                                      */
    int    n_psi, n_f;        /* channel blocks (N_total = n_psi + n_f)             */
    int    n_transition;      /* exchange cut-off index                             */
    int    n_occ;             /* occupied orbitals                                  */
    int    n_exp, n_lambda;   /* table inputs per point: V_s terms, chi per orbital */
    int    tangent;           /* derivative runs (sehf_scan_run_tangent)            */
    int    dipole_path;       /* SEHF_DIPOLES_FORWARD / _BACKWARD                   */
    int    n_energies;        /* energies to run                                    */
    int    batch_cap;         /* <= 0: 64                                           */
    size_t device_bytes;      /* HBM available; 0 = query the device                */
    size_t host_bytes;        /* pinned host memory for checkpoints; 0 = none       */
    size_t runtime_bytes;     /* kept for the runtime and libraries                 */
    double reserve_fraction;  /* of what remains, kept free (the reference: 0.10)   */
} sehf_plan_in;
typedef struct {
    int    tables;            /* SEHF_TABLES_*                                      */
    int    ckpt;              /* SEHF_CKPT_*                                        */
    int    seg_len;           /* checkpoint spacing L                               */
    int    n_ckpt;            /* checkpoints per energy                             */
    int    batch;             /* energies per batch                                 */
    size_t budget_bytes;      /* (device - runtime) x (1 - reserve)                 */
    size_t shared_bytes;      /* energy-independent device bytes                    */
    size_t per_energy_bytes;  /* device bytes per energy of a batch                 */
    size_t host_per_energy_bytes; /* pinned host bytes per energy (host checkpoints) */
    char   report[1536];      /* multi-line summary                                 */
} sehf_plan;
int  sehf_plan_storage(const sehf_plan_in* in, sehf_plan* out);
/* sehf_scan_create with the energy-independent tables resident or streamed
 * (SEHF_TABLES_*; sehf_scan_create is RESIDENT, or STREAMED when
 * $SEHF_TABLES=streamed). */
int  sehf_scan_create_with(sehf_ctx* ctx, const sehf_problem* p, int tables, sehf_scan** out);
/* Where the scan keeps its ratio-matrix checkpoints (SEHF_CKPT_*, default
 * device); takes effect when the batch buffers are next allocated. */
int  sehf_scan_set_checkpoints(sehf_scan* s, int where);
/* Device bytes the handle holds: [0] energy-independent tables and inputs,
 * [1] batch buffers of all groups, [2] pinned host bytes (checkpoints, and
 * the staging ring of the psi files once a psi store has been used). */
int  sehf_scan_memory(const sehf_scan* s, size_t* bytes);

/* Kernel timing: between begin and end, every `every`-th FP64 GEMM and
 * Gauss-Jordan leaf launch is bracketed by HIP events on its own stream.
 * out[16] = five triples {ms, flop, launches} over the sampled launches:
 * all GEMMs, leaves, GEMMs with K < 64, 64 <= K < 256, K >= 256; out[15] =
 * the event-pair cost of an empty launch on the sampled stream (median of
 * 32 pairs), already subtracted from every sampled duration. */
int  sehf_prof_begin(int every);
int  sehf_prof_end(double* out);

/* ---- timing probes of the inner kernels ---------------------------------
 * Synthetic operands on the default stream, HIP events around the timed
 * repetitions (after one untimed run); for the benchmark's kernel table and
 * hardware-counter runs.  No reference counterpart.
 *   sehf_time_gemm: `batch` column-major products C = A B (+ C when
 *     beta_one) of M x K by K x N with the given leading dimensions, average
 *     ms per batched launch sequence (the scan's own GEMM launcher and tile
 *     policy);
 *   sehf_time_inverse: the batched Gauss-Jordan inverse of `batch` random
 *     order-n matrices (diagonal raised by 2), average ms; resid_out (may be
 *     NULL) = sampled max |A X - I| of the first matrix;
 *   sehf_time_spd_inverse: the batched inverse of `batch` symmetric positive
 *     definite order-n matrices (I + a small symmetric perturbation, the
 *     Schur matrices' conditioning) by block elimination without pivoting,
 *     the scan's Schur inverse; resid_out as above, of the last matrix;
 *   sehf_copy_probe: one device-to-device copy of `bytes` (reads and writes
 *     exactly that many bytes: the reference of HBM byte counters). */
int  sehf_time_gemm(int M, int N, int K, int lda, int ldb, int ldc, int batch, int beta_one, int reps,
                    double* ms_out);
int  sehf_time_inverse(int n, int batch, int reps, double* ms_out, double* resid_out);
int  sehf_time_spd_inverse(int n, int batch, int reps, double* ms_out, double* resid_out);
int  sehf_copy_probe(size_t bytes, double* ms_out);

/* ---- host-only inputs (no device needed) -------------------------------
 * The energy-independent matrices and scalar tables the scan uploads, for
 * checking and for callers that stage their own inputs:
 *   V(ir)  N x N, Potentials assembly (src/scatt/Potentials.cpp:324-365);
 *   B(ir)  N x N_f = h^2/12 Q_psi_f (src/scatt/ExchangeCoupling.cpp:66-136);
 *   dipole rows (n_pts x (6+n_occ) x N, Simpson-weighted) and d_corr (6 x n_occ);
 *   sizes: N_psi, N_f, N_total, n_transition, n_start, Nr, l_exch, n_lambda;
 *   Riccati-Bessel / -Neumann, real Gaunt, smallest symmetric eigenvalue. */
int    sehf_host_potential(const sehf_problem* p, int ir, double* V);
int    sehf_host_exchange_B(const sehf_problem* p, int ir, double* B);
int    sehf_host_sizes(const sehf_problem* p, int* info);
int    sehf_host_dipole_tables(const sehf_problem* p, double* rows, double* dcorr, int* n_pts);
double sehf_host_ric_j(int l, double x);
double sehf_host_ric_y(int l, double x);
double sehf_host_real_gaunt(int l1, int m1, int l2, int m2, int l3, int m3);
double sehf_host_min_eig(int n, const double* a);

#ifdef __cplusplus
}
#endif
#endif /* SEHF_H */
