// sehf_setup_kernels.h -- device-side build of the energy-independent tables
// (coupling potential V(r_n) and exchange coupling B(r_n)) from compact
// inputs; `count` consecutive radial points per call, matrices column-major,
// one per point at strides Np^2 and Np*Nf.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sehf {

// V(n) from V_s(r_n) (vs: count x n_exp) and the pair-CSR Gaunt table, then
// the diagonal terms vp[n] (polarisation) and l(l+1)/2 * inv_r2[n]; live[n] = 0
// gives the zero matrix (r = 0).
void build_V(double* V, const double* vs, int n_exp, const int* pi, const int* pj, const int64_t* vptr,
             const int* vsig, const double* vval, int64_t npairs, int Np, const int* l_mu, const double* vp,
             const double* inv_r2, const unsigned char* live, int count, hipStream_t s);

// B(n) = h2_12 * fac[n] * (row-CSR Gaunt table . chi(n)); fac[n] = 0 gives zeros.
void build_B(double* B, const double* chi, int n_occ, int n_lambda, const int64_t* xptr, const int* xlam,
             const double* xval, int Np, int n_sigma, const double* fac, double h2_12, int count, hipStream_t s);

// out[n][i] = c * V(n)_ii
void scaled_diag(double* out, const double* V, double c, int Np, int count, hipStream_t s);

}  // namespace sehf
