// sehf_setup_kernels.hip -- the energy-independent tables of the scan built on
// the device from compact inputs, instead of on the host and uploaded.
//
//   V(r_n)(a, b) = sum_s G_ab,s V_s(r_n)            (GauntSparse.hpp:212-222)
//   V_aa        += V_pol(r_n) + l_a (l_a + 1) / (2 r_n^2)   (Potentials.cpp:324-365)
//   B(r_n)(mu, j n_sigma + s) = h^2/12 (2 sqrt(2 pi) / r_n) sum_lambda G_(mu s),lambda chi_j,lambda(r_n)
//                                                    (ExchangeCoupling.cpp:66-136)
// The host uploads V_s(r) (n_exp per point), chi(r) (n_occ x n_lambda per
// point) and the two sparse Gaunt tables; the dense N_psi^2 and N_psi x N_f
// matrices per radial point (tens of GB at l_cont = 30) are formed in HBM.
// Every element is computed with the host's operations in the host's order
// and without contraction into fused multiply-adds, so the tables are
// bitwise those of the host build (sehf_setup.cpp).
#include "sehf_setup_kernels.h"

#pragma clang fp contract(off)

namespace sehf {

// One thread per (pair q, radial point): both triangle entries of the pair.
__global__ void k_build_V(double* V, const double* vs, int n_exp, const int* pi, const int* pj,
                          const int64_t* vptr, const int* vsig, const double* vval, int64_t npairs, int Np,
                          const unsigned char* live) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int n = blockIdx.y;
    if (q >= npairs) return;
    double* Vn = V + static_cast<int64_t>(n) * Np * Np;
    double acc = 0.0;
    if (live[n]) {
        const double* v = vs + static_cast<int64_t>(n) * n_exp;
        for (int64_t t = vptr[q]; t < vptr[q + 1]; ++t) acc += vval[t] * v[vsig[t]];
    }
    const int a = pi[q], b = pj[q];
    Vn[a + static_cast<int64_t>(b) * Np] = acc;
    Vn[b + static_cast<int64_t>(a) * Np] = acc;
}

// Diagonal terms after the pair sums: polarisation, then centrifugal.
__global__ void k_V_diag(double* V, const int* l_mu, const double* vp, const double* inv_r2, int Np,
                         const unsigned char* live) {
    const int mu = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = blockIdx.y;
    if (mu >= Np || !live[n]) return;
    double& d = V[static_cast<int64_t>(n) * Np * Np + mu + static_cast<int64_t>(mu) * Np];
    double x = d;
    if (vp[n] != 0.0) x += vp[n];
    x += l_mu[mu] * (l_mu[mu] + 1.0) * 0.5 * inv_r2[n];
    d = x;
}

// One thread per (mu, s) row and radial point, all occupied orbitals j.
__global__ void k_build_B(double* B, const double* chi, int n_occ, int n_lambda, const int64_t* xptr,
                          const int* xlam, const double* xval, int Np, int n_sigma, const double* fac,
                          double h2_12) {
    const int rr = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = blockIdx.y;
    if (rr >= Np * n_sigma) return;
    const int mu = rr / n_sigma, s = rr % n_sigma;
    const int Nf = n_occ * n_sigma;
    double* Bn = B + static_cast<int64_t>(n) * Np * Nf;
    const double f = fac[n];
    for (int j = 0; j < n_occ; ++j) {
        double out = 0.0;
        if (f != 0.0) {
            const double* c = chi + (static_cast<int64_t>(n) * n_occ + j) * n_lambda;
            double acc = 0.0;
            for (int64_t t = xptr[rr]; t < xptr[rr + 1]; ++t) acc += xval[t] * c[xlam[t]];
            out = h2_12 * (f * acc);
        }
        Bn[mu + static_cast<int64_t>(j * n_sigma + s) * Np] = out;
    }
}

// cvd[n][i] = c * V(n)_ii
__global__ void k_scaled_diag(double* out, const double* V, double c, int Np) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = blockIdx.y;
    if (i >= Np) return;
    out[static_cast<int64_t>(n) * Np + i] = c * V[static_cast<int64_t>(n) * Np * Np + i + static_cast<int64_t>(i) * Np];
}

void build_V(double* V, const double* vs, int n_exp, const int* pi, const int* pj, const int64_t* vptr,
             const int* vsig, const double* vval, int64_t npairs, int Np, const int* l_mu, const double* vp,
             const double* inv_r2, const unsigned char* live, int count, hipStream_t s) {
    if (count <= 0) return;
    dim3 g(static_cast<unsigned>((npairs + 255) / 256), count);
    hipLaunchKernelGGL(k_build_V, g, dim3(256), 0, s, V, vs, n_exp, pi, pj, vptr, vsig, vval, npairs, Np, live);
    dim3 gd((Np + 255) / 256, count);
    hipLaunchKernelGGL(k_V_diag, gd, dim3(256), 0, s, V, l_mu, vp, inv_r2, Np, live);
}

void build_B(double* B, const double* chi, int n_occ, int n_lambda, const int64_t* xptr, const int* xlam,
             const double* xval, int Np, int n_sigma, const double* fac, double h2_12, int count, hipStream_t s) {
    if (count <= 0) return;
    dim3 g((Np * n_sigma + 255) / 256, count);
    hipLaunchKernelGGL(k_build_B, g, dim3(256), 0, s, B, chi, n_occ, n_lambda, xptr, xlam, xval, Np, n_sigma, fac,
                       h2_12);
}

void scaled_diag(double* out, const double* V, double c, int Np, int count, hipStream_t s) {
    if (count <= 0) return;
    dim3 g((Np + 255) / 256, count);
    hipLaunchKernelGGL(k_scaled_diag, g, dim3(256), 0, s, out, V, c, Np);
}

}  // namespace sehf
