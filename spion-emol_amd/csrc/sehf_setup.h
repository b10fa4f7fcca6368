// sehf_setup.h -- energy-independent inputs of the scan, built on the host.
//
// Mirrors what the reference builds once per molecule before (or, for the
// wavefunction setup, inside) its energy loop:
//   * Potentials::build (src/scatt/Potentials.cpp:144-453): the coupling
//     matrix V(r_n) = sum_s V_s(r_n) G(mu, s, nu) + V_pol + l(l+1)/2r^2,
//     r_n = n dr;
//   * WavefunctionSetup::prepare (src/scatt/WavefunctionSetup.cpp:219-376):
//     l_exch = min(l_cont, 10), n_transition from the 1e-15 orbital cut-off,
//     the Gaunt triplets G(mu, lambda, sigma);
//   * ExchangeCoupling::compute_into (src/scatt/ExchangeCoupling.cpp:66-136):
//     B(n) = h^2/12 Q(n), Q = 2 sqrt(2 pi)/r sum_lambda G chi_lambda;
//   * the dipole angular tables and Simpson-weighted integrand rows of
//     DipoleMatrixElement (src/scatt/DipoleMatrixElement.cpp:204-282).
// The hot path only streams the results.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/sehf.h"

namespace sehf {

class Setup {
public:
    explicit Setup(const sehf_problem& p);

    const sehf_problem P;   // copied; its arrays are read only during construction
    int l_cont, l_exch, l_orb, Np, n_sigma, n_lambda, l_exp, n_exp, n_occ, Nr, Nf, Nt;
    int n_tr = 0, n_start = 1;
    double dr;
    std::vector<int> l_mu, m_mu, l_sig;

    // ---- per radial index (solver radius r = (n+1) dr; potential radius n dr)
    void potential(int ir, double* V) const;   // Np x Np, col-major, r = ir*dr
    // V(n) and B(n) for n in [n0, n0+count), in parallel over n.
    void build_chunk(int n0, int count, double* V, double* B) const;
    void exchange_B(int ir, double* B) const;  // Np x Nf, col-major, (h^2/12) Q
    void dinv(int n, double* d) const;         // Nf: 1/max(1 - h^2/12 l(l+1)/r^2, W_min)
    double polarization(double r) const;
    // d V_pol / d alpha at potential index ir (r = ir dr): the weight with
    // which the polarisation strength enters the diagonal of V.
    double polarization_per_alpha(int ir) const;

    // ---- dipole: per-radial-point weighted rows and the energy-independent
    // orthogonalisation integrals.
    //   rows (n_pts x (6 + n_occ) x Np, point-major, each block col-major
    //   (6+n_occ) x Np): slot s < 6:  w_ir rfac_s(ir) Xi_s(ir, mu)
    //                     slot 6+a:   w_ir phi_a(ir, mu)
    //   dcorr (6 x n_occ): int rfac phi_a . Xi_s
    void dipole_tables(std::vector<double>& rows, std::vector<double>& dcorr, int& n_pts) const;

    // ---- inputs of the device-side table build (sehf_setup_kernels.hip):
    // the same numbers potential() and exchange_B() combine on the host.
    // V_s(r_n) for n in [n0, n0+count): out[(n-n0) * n_exp + s].
    void vsigma_table(int n0, int count, double* out) const;
    // chi_j,lambda(r_n) = r_n F_j,lambda(r_n): out[((n-n0) * n_occ + j) * n_lambda + lambda].
    void chi_table(int n0, int count, double* out) const;
    const std::vector<int>& pot_pair_i() const { return pair_i_; }
    const std::vector<int>& pot_pair_j() const { return pair_j_; }
    const std::vector<int64_t>& pot_ptr() const { return vptr_; }
    const std::vector<int>& pot_sigma() const { return vsig_; }
    const std::vector<double>& pot_val() const { return vval_; }
    const std::vector<int64_t>& exch_ptr() const { return xptr_; }
    const std::vector<int>& exch_lambda() const { return xlam_; }
    const std::vector<double>& exch_val() const { return xval_; }

    static constexpr double W_min = 0.001;        // SchurInverter.hpp:74
    static constexpr int stab_n_max = 60;          // SchurInverter.hpp:75
    static constexpr double shift_margin = 0.01;   // SchurInverter.hpp:76

private:
    // potential Gaunt, CSR over upper-triangle pairs (GauntSparse.hpp:81-200)
    std::vector<int> pair_i_, pair_j_;
    std::vector<int64_t> vptr_;
    std::vector<int> vsig_;
    std::vector<double> vval_;
    // exchange Gaunt, CSR over rows mu*n_sigma + sigma (ExchangeCoupling.cpp:13-54)
    std::vector<int64_t> xptr_;
    std::vector<int> xlam_;
    std::vector<double> xval_;

    void v_sigma(double r, int ir, std::vector<double>& out) const;
    void build_potential_gaunt();
    void build_exchange_gaunt();
    void find_n_transition();
    double chi(int j, int lam, int ir) const;   // r F_{j,lam}(r), r = ir dr
};

}  // namespace sehf
