// sehf_host_math.h -- host numerics used while preparing device inputs.
//
// Energy-independent angular algebra (Wigner 3j, real Gaunt coefficients,
// real spherical harmonics) and the per-energy scalar tables (Riccati-Bessel
// values for the matching and fit radii, smallest eigenvalue of the few
// small-r Schur matrices that need the stability shift).  None of this is
// on the per-radial-step path.
#pragma once

#include <vector>

namespace sehf {

// Channel index l*l + l + m (src/angular/Gaunt.hpp:46-56).
inline int lm_index(int l, int m) { return l * l + l + m; }
void index_lm(int idx, int& l, int& m);

// Wigner 3j symbol (Racah's closed form; long double with a __float128
// re-evaluation when the alternating sum cancels).
double wigner3j(int j1, int j2, int j3, int m1, int m2, int m3);

// Real Gaunt coefficient  int Y^R_{l1 m1} Y^R_{l2 m2} Y^R_{l3 m3} dOmega,
// real harmonics without the Condon-Shortley phase (Gaunt.hpp:143-181).
double real_gaunt(int l1, int m1, int l2, int m2, int l3, int m3);

// Table-driven real Gaunt coefficients for l <= lmax: the theta integral
// int S_l1^|m1| S_l2^|m2| S_l3^|m3| dx by Gauss-Legendre quadrature (exact:
// the integrand is a polynomial of degree l1+l2+l3 whenever the phi integral
// is non-zero), times the phi integral of the three real azimuthal factors
// in closed form.  About 100 multiply-adds per coefficient, no cancellation
// beyond the quadrature sum; used to build the sparse coupling tables.
class GauntTable {
public:
    explicit GauntTable(int lmax);
    double operator()(int l1, int m1, int l2, int m2, int l3, int m3) const;
    int lmax() const { return lmax_; }

private:
    int lmax_, nq_;
    std::vector<double> w_;     // weights
    std::vector<double> S_;     // S_l^m(x_q): index ((l*(l+1))/2 + m) * nq + q
};

// Real spherical harmonic Y^R_{lm}(theta, phi) in the same convention.
double real_ylm(int l, int m, double theta, double phi);

// Riccati-Bessel  x j_l(x)  and Riccati-Neumann  x y_l(x)  (y_0 = -cos x).
double ric_j(int l, double x);
double ric_y(int l, double x);
// Their x-derivatives from the recurrence u_l' = u_{l-1} - (l/x) u_l
// (u_0' = cos x for j, sin x for y).
double ric_jp(int l, double x);
double ric_yp(int l, double x);

// Smallest eigenvalue of a symmetric n x n matrix given column-major
// (lower triangle read).
double sym_min_eig(int n, const double* a);

// Composite Simpson weights on n uniformly spaced samples, reproducing the
// 1/3 rule for odd n and 1/3 + 3/8 stitching for even n
// (src/scatt/DipoleMatrixElement.cpp:142-192).
std::vector<double> simpson_weights(int n, double h);

// Five-point first derivative with 4th-order one-sided ends
// (src/scatt/DipoleMatrixElement.cpp:49-74).
void deriv5(const double* f, int n, double h, double* df);

}  // namespace sehf
