// sehf_scan_kernels.h -- launch wrappers of the scan's elementwise kernels.
// Batched over energies; all matrices column-major with their natural
// leading dimension unless an ld argument says otherwise.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sehf {

void init_U(double* U, const double* R, const double* Sinv, const double* Dinv, int Np, int Nt,
            int batch, hipStream_t s);
void add_psi_f_blocks(double* U, const double* X, int64_t sX, double a, int Np, int Nf, int Nt, int batch,
                      hipStream_t s);
void shift_copy(double* Y, const double* M0, const double* shift, int n, int batch, hipStream_t s);
void sinv_matrix(double* Y, const double* S0, const double* cvd, const double* bdb, const double* cE,
                 const double* dA, const double* dS, int n, int batch, hipStream_t s);
void row_dot(double* out, const double* X, const double* Y, int r, int c, int batch, hipStream_t s);
void diag_matrix(double* R, const double* d, int n, int batch, hipStream_t s);
void fit_accum(double* YJ, double* YY, double* win, int64_t sWin, const double* psi, const double* Jc,
               const double* Yc, int64_t sC, int Np, int batch, hipStream_t s);
void fit_solve(double* A, double* B, const double* YJ, const double* YY, const double* coef, int Np,
               int batch, hipStream_t s);
void fit_resid(const double* win, const double* A, const double* B, const double* Jc, const double* Yc,
               const double* coef, int Np, int nfit, int64_t sWin, int64_t sC, double* out, int batch,
               hipStream_t s);
void symmetrize(double* M, int n, int batch, hipStream_t s);
void get_block(double* Y, const double* X, int64_t ldx, int64_t sX, int i0, int j0, int r, int c,
               int batch, hipStream_t s);
void kme_mats(double* Am, double* Bm, const double* S, const double* tab, int Np, int batch, hipStream_t s);
void psi_boundary(double* P, const double* K, const double* jy, int Np, int batch, hipStream_t s);
void row_scale(double* Y, const double* X, const double* d, int r, int c, int batch, hipStream_t s);
void z_boundary(double* Z, const double* P, const double* Vpsi, const double* Bf, const double* cE,
                double c, int Np, int Nt, int batch, hipStream_t s);
void w2_assemble(double* W2, const double* A, const double* B, int Np, int batch, hipStream_t s);
void dipole_final(double* Dre, double* Dim, const double* DRre, const double* DRim, const double* ore,
                  const double* oim, const double* dcorr, int Np, int n_occ, int batch, hipStream_t s);
void sumsq_block(const double* X, int64_t ldx, int64_t sX, int i0, int j0, int r, int c, double* out,
                 int batch, hipStream_t s);

// ---- tangent along (E, alpha) ----
void axpy_b(double* Y, int64_t ldy, int64_t sY, const double* X, int64_t ldx, int64_t sX,
            const double* sc, double a, int r, int c, int batch, hipStream_t s);
void kme_dmats(double* dAm, double* dBm, const double* dS, const double* S, const double* tab,
               const double* dtab, int Np, int batch, hipStream_t s);
void dpsi_boundary(double* dP, const double* K, const double* dK, const double* jy, const double* djy,
                   int Np, int batch, hipStream_t s);
void fit_accum_t(double* dYJ, double* dYY, const double* psi, const double* dpsi, const double* Jc,
                 const double* Yc, const double* dJc, const double* dYc, int64_t sC, int Np, int batch,
                 hipStream_t s);
void fit_dsolve(double* dA, double* dB, const double* YJ, const double* YY, const double* dYJ,
                const double* dYY, const double* coef, const double* dcoef, int Np, int batch, hipStream_t s);

}  // namespace sehf
