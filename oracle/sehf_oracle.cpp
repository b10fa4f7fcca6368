// sehf_oracle.cpp -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
//
// CPU FP64 restatement of the reference's energy-dependent SE-HF
// photoionization path.  Every function cites the reference file:line whose
// algorithm it restates.  Third-party numerics the reference takes from
// libraries absent here are restated from their published algorithms:
//   * GSL gsl_sf_coupling_3j (no pinned version, CMakeLists.txt:139): Racah's
//     closed-form sum (Edmonds, "Angular Momentum in QM", eq. 3.7.3), evaluated
//     with log-factorials as GSL does, here in long double;
//   * GSL gsl_sf_bessel_jl_e / yl_e: spherical Bessel j_l, y_l (Abramowitz &
//     Stegun 10.1) via series / Miller downward recurrence / upward recurrence;
//   * Eigen 3.4 PartialPivLU::inverse, FullPivLU::inverse, complex LU and
//     SelfAdjointEigenSolver (smallest eigenvalue only): textbook partial-pivot
//     LU (Golub & Van Loan 3.4), full-pivot LU, and Householder
//     tridiagonalisation + Sturm bisection.
// The reference's MKL branches (dsytrf/dsytri) are not restated: the oracle
// follows the no-MKL build, where inverse_symmetric_indefinite is
// PartialPivLU + 1/2(M+M^T) (LapackInverse.hpp:174-176).
//
// Parity pins live in tests/ (sympy 3j, scipy Bessel, the reference tests'
// golden Gaunt table, closed forms and invariants).

#include "sehf_oracle.h"

#include <quadmath.h>

#include <algorithm>
#include <array>
#include <limits>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

using cd = std::complex<double>;
static thread_local std::string g_err;

// ---------------------------------------------------------------------------
// Column-major dense matrix.
// ---------------------------------------------------------------------------
template <class T>
struct MatT {
    int r = 0, c = 0;
    std::vector<T> a;
    MatT() = default;
    MatT(int r_, int c_) : r(r_), c(c_), a(static_cast<size_t>(r_) * c_, T(0)) {}
    T& operator()(int i, int j) { return a[static_cast<size_t>(j) * r + i]; }
    const T& operator()(int i, int j) const { return a[static_cast<size_t>(j) * r + i]; }
    T* col(int j) { return a.data() + static_cast<size_t>(j) * r; }
    const T* col(int j) const { return a.data() + static_cast<size_t>(j) * r; }
    void zero() { std::fill(a.begin(), a.end(), T(0)); }
    void resize(int r_, int c_) { r = r_; c = c_; a.assign(static_cast<size_t>(r_) * c_, T(0)); }
};
using Mat = MatT<double>;
using CMat = MatT<cd>;

static Mat identity(int n) { Mat I(n, n); for (int i = 0; i < n; ++i) I(i, i) = 1.0; return I; }
static Mat transpose(const Mat& A) {
    Mat T(A.c, A.r);
    for (int j = 0; j < A.c; ++j) for (int i = 0; i < A.r; ++i) T(j, i) = A(i, j);
    return T;
}
// C = alpha*(A*B) + beta*C  (column-major).  BLAS/Eigen semantics: the
// product is formed first and then combined with C, so an update such as
// S = A - Bs*B^T groups as A_ij - (sum_p ...), the way the reference's
// GEMM evaluates it (SchurInverter.cpp:347-351).
// Rounding-order probe (noise-floor measurements only): with g_gemm_rev set
// on this thread the product sums run over p in reverse order.
static thread_local bool g_gemm_rev = false;
// Threads for the loops over independent output columns (GEMM, LU update,
// LU solve): each column's operations are unchanged, so the results are
// bitwise those of one thread.
static int g_threads = 1;
static void gemm(double alpha, const Mat& A, const Mat& B, double beta, Mat& C) {
    const int m = A.r, k = A.c, n = B.c;
    if (C.r != m || C.c != n) { C.resize(m, n); beta = 0.0; }
    // output columns in groups of 8 that share each pass over A's columns;
    // every element keeps its own sum in p order
    constexpr int SG = 8;
    const int ng = (n + SG - 1) / SG;
    const bool par = g_threads > 1 && static_cast<double>(m) * k * n > 4e6;
#pragma omp parallel num_threads(g_threads) if (par)
    {
    std::vector<double> t(static_cast<size_t>(m) * SG);
#pragma omp for schedule(static)
    for (int g = 0; g < ng; ++g) {
        const int j0 = g * SG, gw = std::min(SG, n - j0);
        std::fill(t.begin(), t.end(), 0.0);
        for (int pp = 0; pp < k; ++pp) {
            const int p = g_gemm_rev ? k - 1 - pp : pp;
            const double* ap = A.col(p);
            for (int q = 0; q < gw; ++q) {
                const double s = B(p, j0 + q);
                if (s == 0.0) continue;
                double* tq = t.data() + static_cast<size_t>(q) * m;
                for (int i = 0; i < m; ++i) tq[i] += s * ap[i];
            }
        }
        for (int q = 0; q < gw; ++q) {
            double* cj = C.col(j0 + q);
            const double* tq = t.data() + static_cast<size_t>(q) * m;
            if (beta == 0.0) for (int i = 0; i < m; ++i) cj[i] = alpha * tq[i];
            else if (beta == 1.0) for (int i = 0; i < m; ++i) cj[i] = cj[i] + alpha * tq[i];
            else for (int i = 0; i < m; ++i) cj[i] = beta * cj[i] + alpha * tq[i];
        }
    }
    }
}
static Mat mul(const Mat& A, const Mat& B) { Mat C(A.r, B.c); gemm(1.0, A, B, 0.0, C); return C; }
static void symmetrize(Mat& M) {  // M <- 0.5 (M + M^T)
    const int n = M.r;
    for (int j = 0; j < n; ++j)
        for (int i = j + 1; i < n; ++i) {
            const double v = 0.5 * (M(i, j) + M(j, i));
            M(i, j) = v; M(j, i) = v;
        }
    for (int i = 0; i < n; ++i) M(i, i) = 0.5 * (M(i, i) + M(i, i));
}
static double max_abs(const Mat& M) { double m = 0; for (double v : M.a) m = std::max(m, std::abs(v)); return m; }

// ---------------------------------------------------------------------------
// Dense inverses (LapackInverse.hpp:53-92 no-MKL branch = Eigen PartialPivLU).
// ---------------------------------------------------------------------------
template <class T>
static double absval(const T& v) { return std::abs(v); }

// In-place partial-pivot LU: P A = L U; perm[k] = row swapped into k.
// Right-looking elimination in the order of the unblocked algorithm: at
// step k the pivot is the first row of largest |a(i,k)|, i >= k; rows k
// and p swap in every column; l = a(i,k) / a(k,k); then a(i,j) -= l * a(k,j)
// for i, j > k, skipped where a(k,j) == 0 (a zero pivot skips the step).
// For cache reuse the columns right of a 32-column panel receive the
// panel's steps together, one column at a time and step by step, after the
// panel is factored: each element still sees the same operations in the
// same order (the row swaps commute with the deferred updates because they
// move the panel's multipliers with the rows), so the factors are bitwise
// those of the unblocked loop.
template <class T>
static void lu_partial(MatT<T>& a, std::vector<int>& perm) {
    const int n = a.r;
    perm.resize(n);
    constexpr int NB = 32;
    std::vector<char> live(n, 0);
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int k1 = std::min(n, k0 + NB);
        for (int k = k0; k < k1; ++k) {
            int p = k; double best = absval(a(k, k));
            for (int i = k + 1; i < n; ++i) { const double v = absval(a(i, k)); if (v > best) { best = v; p = i; } }
            perm[k] = p;
            if (p != k) for (int j = 0; j < n; ++j) std::swap(a(k, j), a(p, j));
            const T piv = a(k, k);
            if (piv == T(0)) continue;
            live[k] = 1;
            T* ck = a.col(k);
            for (int i = k + 1; i < n; ++i) ck[i] /= piv;
            for (int j = k + 1; j < k1; ++j) {
                T* cj = a.col(j);
                const T f = cj[k];
                if (f == T(0)) continue;
                for (int i = k + 1; i < n; ++i) cj[i] -= ck[i] * f;
            }
        }
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1 && n - k1 > 64)
        for (int j = k1; j < n; ++j) {
            T* cj = a.col(j);
            for (int k = k0; k < k1; ++k) {
                if (!live[k]) continue;
                const T f = cj[k];
                if (f == T(0)) continue;
                const T* ck = a.col(k);
                for (int i = k + 1; i < n; ++i) cj[i] -= ck[i] * f;
            }
        }
    }
}
// Solve (L U) X = P B in place on X (B given in X), using LU from lu_partial:
// per column, unit-lower forward substitution then upper backward
// substitution, skipping zero entries.  Columns are taken in groups of 8
// that share each pass over L and U; every column sees the same operations
// in the same order as alone.
template <class T>
static void lu_solve(const MatT<T>& lu, const std::vector<int>& perm, MatT<T>& X) {
    const int n = lu.r, m = X.c;
    for (int k = 0; k < n; ++k)
        if (perm[k] != k) for (int j = 0; j < m; ++j) std::swap(X(k, j), X(perm[k], j));
    constexpr int SG = 8;
    const int ng = (m + SG - 1) / SG;
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1 && static_cast<double>(n) * n * m > 4e6)
    for (int g = 0; g < ng; ++g) {
        const int j0 = g * SG, gw = std::min(SG, m - j0);
        T* xs[SG];
        for (int q = 0; q < gw; ++q) xs[q] = X.col(j0 + q);
        for (int k = 0; k < n; ++k) {            // unit-lower forward
            const T* lk = lu.col(k);
            for (int q = 0; q < gw; ++q) {
                T* x = xs[q];
                const T xk = x[k];
                if (xk == T(0)) continue;
                for (int i = k + 1; i < n; ++i) x[i] -= lk[i] * xk;
            }
        }
        for (int k = n - 1; k >= 0; --k) {       // upper backward
            const T* uk = lu.col(k);
            const T d = lu(k, k);
            for (int q = 0; q < gw; ++q) {
                T* x = xs[q];
                x[k] /= d;
                const T xk = x[k];
                if (xk == T(0)) continue;
                for (int i = 0; i < k; ++i) x[i] -= uk[i] * xk;
            }
        }
    }
}
template <class T>
static MatT<T> inverse_pp(const MatT<T>& A) {
    MatT<T> lu = A; std::vector<int> perm;
    lu_partial(lu, perm);
    MatT<T> X(A.r, A.r);
    for (int i = 0; i < A.r; ++i) X(i, i) = T(1);
    lu_solve(lu, perm, X);
    return X;
}
// Gauss-Jordan elimination with partial pivoting: an equally valid
// elimination order, used only to measure how far two correct rounding
// orders of the reference's algorithm drift apart (the parity noise floor).
static Mat inverse_gj(const Mat& A) {
    const int n = A.r;
    Mat a = A;
    std::vector<int> piv(n);
    for (int k = 0; k < n; ++k) {
        int p = k; double best = std::abs(a(k, k));
        for (int i = k + 1; i < n; ++i) if (std::abs(a(i, k)) > best) { best = std::abs(a(i, k)); p = i; }
        piv[k] = p;
        if (p != k) for (int j = 0; j < n; ++j) std::swap(a(k, j), a(p, j));
        const double pinv = 1.0 / a(k, k);
        for (int j = 0; j < n; ++j) if (j != k) a(k, j) *= pinv;
        a(k, k) = pinv;
        for (int j = 0; j < n; ++j) {
            if (j == k) continue;
            const double f = a(k, j);
            double* cj = a.col(j);
            const double* ck = a.col(k);
            for (int i = 0; i < n; ++i) if (i != k) cj[i] -= ck[i] * f;
        }
        for (int i = 0; i < n; ++i) if (i != k) a(i, k) = -a(i, k) * pinv;
    }
    for (int k = n - 1; k >= 0; --k) if (piv[k] != k) for (int i = 0; i < n; ++i) std::swap(a(i, k), a(i, piv[k]));
    return a;
}
// Rounding-order probes (noise-floor measurements only): bit 0 selects
// Gauss-Jordan in place of LU for the dense inverses, bit 1 accumulates the
// asymptotic-fit sums over the window from its outer end inwards; bits 2, 3
// and 4 select Gauss-Jordan for one stage's inverses only (Schur inverses,
// forward ratio steps, K matrix); bit 5 runs the partial-pivot LU in 80-bit
// extended precision (rounded to FP64 once at the end); bit 6 sums the
// products of the W^{-1} application (forward ratio steps and
// back-propagation) in reverse order.
static int g_inverse_mode = 0;
enum InvStage { INV_ANY = -1, INV_SINV = 0, INV_FRP = 1, INV_KME = 2 };
static Mat inverse_ext(const Mat& A) {
    MatT<long double> a(A.r, A.c);
    for (size_t t = 0; t < A.a.size(); ++t) a.a[t] = A.a[t];
    const MatT<long double> x = inverse_pp(a);
    Mat X(A.r, A.c);
    for (size_t t = 0; t < X.a.size(); ++t) X.a[t] = static_cast<double>(x.a[t]);
    return X;
}
static Mat inverse_general(const Mat& A, int stage = INV_ANY) {
    const bool gj = (g_inverse_mode & 1) || (stage >= 0 && ((g_inverse_mode >> (2 + stage)) & 1));
    if (gj) return inverse_gj(A);
    if (g_inverse_mode & 32) return inverse_ext(A);
    return inverse_pp(A);
}
// inverse_symmetric_indefinite, no-MKL branch (LapackInverse.hpp:174-176).
static Mat inverse_symmetric_indefinite(const Mat& A) { Mat X = inverse_general(A, INV_SINV); symmetrize(X); return X; }

// Full-pivot LU inverse (Eigen FullPivLU; AsymptoticAmplitudes.cpp:317-322).
static bool inverse_fullpiv(const Mat& A, Mat& Xout) {
    const int n = A.r;
    Mat a = A;
    std::vector<int> rp(n), cp(n);
    double maxpiv = 0.0;
    std::vector<double> pivs(n);
    for (int k = 0; k < n; ++k) {
        int pi = k, pj = k; double best = -1.0;
        for (int j = k; j < n; ++j) for (int i = k; i < n; ++i) {
            const double v = std::abs(a(i, j)); if (v > best) { best = v; pi = i; pj = j; }
        }
        rp[k] = pi; cp[k] = pj;
        if (pi != k) for (int j = 0; j < n; ++j) std::swap(a(k, j), a(pi, j));
        if (pj != k) for (int i = 0; i < n; ++i) std::swap(a(i, k), a(i, pj));
        pivs[k] = std::abs(a(k, k)); maxpiv = std::max(maxpiv, pivs[k]);
        const double piv = a(k, k);
        if (piv == 0.0) continue;
        for (int i = k + 1; i < n; ++i) a(i, k) /= piv;
        for (int j = k + 1; j < n; ++j) { const double f = a(k, j); for (int i = k + 1; i < n; ++i) a(i, j) -= a(i, k) * f; }
    }
    const double thr = std::numeric_limits<double>::epsilon() * n * maxpiv;
    for (int k = 0; k < n; ++k) if (pivs[k] <= thr) return false;
    // A = P^T L U Q^T with row swaps rp and col swaps cp -> solve for identity.
    Mat X(n, n); for (int i = 0; i < n; ++i) X(i, i) = 1.0;
    for (int k = 0; k < n; ++k) if (rp[k] != k) for (int j = 0; j < n; ++j) std::swap(X(k, j), X(rp[k], j));
    for (int j = 0; j < n; ++j) {
        double* x = X.col(j);
        for (int k = 0; k < n; ++k) { const double xk = x[k]; for (int i = k + 1; i < n; ++i) x[i] -= a(i, k) * xk; }
        for (int k = n - 1; k >= 0; --k) { x[k] /= a(k, k); const double xk = x[k]; for (int i = 0; i < k; ++i) x[i] -= a(i, k) * xk; }
    }
    for (int k = n - 1; k >= 0; --k) if (cp[k] != k) for (int j = 0; j < n; ++j) std::swap(X(k, j), X(cp[k], j));
    Xout = X;
    return true;
}

// Smallest eigenvalue of a symmetric matrix from its LOWER triangle
// (Eigen SelfAdjointEigenSolver, SchurInverter.cpp:43-48): Householder
// tridiagonalisation, then bisection on the Sturm sequence.
static double min_eig_sym(const Mat& M) {
    const int n = M.r;
    if (n == 1) return M(0, 0);
    std::vector<double> A(static_cast<size_t>(n) * n);
    for (int j = 0; j < n; ++j) for (int i = j; i < n; ++i) { A[(size_t)j * n + i] = M(i, j); A[(size_t)i * n + j] = M(i, j); }
    auto a = [&](int i, int j) -> double& { return A[(size_t)j * n + i]; };
    std::vector<double> d(n), e(n, 0.0), v(n), w(n);
    for (int k = 0; k < n - 2; ++k) {
        double alpha2 = 0.0;
        for (int i = k + 1; i < n; ++i) alpha2 += a(i, k) * a(i, k);
        const double nrm = std::sqrt(alpha2);
        if (nrm == 0.0) { e[k] = 0.0; continue; }
        const double alpha = (a(k + 1, k) > 0) ? -nrm : nrm;
        for (int i = 0; i < n; ++i) v[i] = 0.0;
        for (int i = k + 1; i < n; ++i) v[i] = a(i, k);
        v[k + 1] -= alpha;
        double vn2 = 0.0; for (int i = k + 1; i < n; ++i) vn2 += v[i] * v[i];
        e[k] = alpha;
        if (vn2 == 0.0) continue;
        const double beta = 2.0 / vn2;
        // p = beta * A v ; K = beta/2 * v^T p ; w = p - K v ; A -= v w^T + w v^T
        for (int i = k + 1; i < n; ++i) { double s = 0.0; for (int j = k + 1; j < n; ++j) s += a(i, j) * v[j]; w[i] = beta * s; }
        double K = 0.0; for (int i = k + 1; i < n; ++i) K += v[i] * w[i]; K *= 0.5 * beta;
        for (int i = k + 1; i < n; ++i) w[i] -= K * v[i];
        for (int j = k + 1; j < n; ++j) for (int i = k + 1; i < n; ++i) a(i, j) -= v[i] * w[j] + w[i] * v[j];
    }
    for (int i = 0; i < n; ++i) d[i] = a(i, i);
    e[n - 2] = a(n - 1, n - 2);
    // off-diagonals: e[k] links d[k], d[k+1]
    double lo = 1e300, hi = -1e300;
    for (int i = 0; i < n; ++i) {
        const double r = (i > 0 ? std::abs(e[i - 1]) : 0.0) + (i < n - 1 ? std::abs(e[i]) : 0.0);
        lo = std::min(lo, d[i] - r); hi = std::max(hi, d[i] + r);
    }
    auto count_below = [&](double x) {
        int cnt = 0; double q = d[0] - x;
        if (q < 0) ++cnt;
        for (int i = 1; i < n; ++i) {
            if (q == 0.0) q = 1e-300;
            q = d[i] - x - e[i - 1] * e[i - 1] / q;
            if (q < 0) ++cnt;
        }
        return cnt;
    };
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (count_below(mid) >= 1) hi = mid; else lo = mid;
    }
    return 0.5 * (lo + hi);
}

// ---------------------------------------------------------------------------
// Angular algebra (Gaunt.hpp:46-181).
// ---------------------------------------------------------------------------
inline void idx_to_lm(int idx, int& l, int& m) {   // Gaunt.hpp:52-56
    l = static_cast<int>(std::sqrt(static_cast<double>(idx)));
    if ((l + 1) * (l + 1) <= idx) ++l;
    m = idx - l * l - l;
}
inline int lm_to_idx(int l, int m) { return l * l + l + m; }

static const std::vector<long double>& log_fact() {
    static const std::vector<long double> t = [] {
        std::vector<long double> v(2048);
        v[0] = 0.0L;
        for (size_t i = 1; i < v.size(); ++i) v[i] = v[i - 1] + logl(static_cast<long double>(i));
        return v;
    }();
    return t;
}

// Wigner 3j by Racah's sum (the formula gsl_sf_coupling_3j evaluates), in
// long double; when the alternating sum cancels by more than two digits it is
// re-evaluated in __float128 so the result stays exact to ~1e-16 up to l~100.
static const std::vector<__float128>& log_fact_q() {
    static const std::vector<__float128> t = [] {
        std::vector<__float128> v(2048);
        v[0] = 0;
        for (size_t i = 1; i < v.size(); ++i) v[i] = v[i - 1] + logq(static_cast<__float128>(i));
        return v;
    }();
    return t;
}
template <class R, class LF, class EXP>
static R racah_sum(int j1, int j2, int j3, int m1, int m2, int m3, const LF& lf, EXP ex, R& maxterm) {
    const R lnpre = R(0.5) * (lf[j1 + j2 - j3] + lf[j1 - j2 + j3] + lf[-j1 + j2 + j3] - lf[j1 + j2 + j3 + 1]
                              + lf[j1 + m1] + lf[j1 - m1] + lf[j2 + m2] + lf[j2 - m2] + lf[j3 + m3] + lf[j3 - m3]);
    const int kmin = std::max(0, std::max(j2 - j3 - m1, j1 - j3 + m2));
    const int kmax = std::min(j1 + j2 - j3, std::min(j1 - m1, j2 + m2));
    R pos = 0, neg = 0;
    maxterm = 0;
    for (int k = kmin; k <= kmax; ++k) {
        const R t = ex(lnpre - (lf[k] + lf[j3 - j2 + k + m1] + lf[j3 - j1 + k - m2] + lf[j1 + j2 - j3 - k]
                                + lf[j1 - k - m1] + lf[j2 - k + m2]));
        if (t > maxterm) maxterm = t;
        if (k & 1) neg += t; else pos += t;
    }
    R s = pos - neg;
    if ((j1 - j2 - m3) & 1) s = -s;
    return s;
}
static double w3j(int j1, int j2, int j3, int m1, int m2, int m3) {
    if (m1 + m2 + m3 != 0) return 0.0;
    if (j3 < std::abs(j1 - j2) || j3 > j1 + j2) return 0.0;
    if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(m3) > j3) return 0.0;
    if (m1 == 0 && m2 == 0) {
        // closed form of the m = 0 sum (no cancellation): J = j1+j2+j3 even,
        // g = J/2: (-1)^g sqrt[(J-2j1)!(J-2j2)!(J-2j3)!/(J+1)!] g!/[(g-j1)!(g-j2)!(g-j3)!]
        const int J = j1 + j2 + j3;
        if (J & 1) return 0.0;
        const int g = J / 2;
        const auto& lf = log_fact();
        const long double l = 0.5L * (lf[J - 2 * j1] + lf[J - 2 * j2] + lf[J - 2 * j3] - lf[J + 1]) + lf[g] - lf[g - j1] -
                              lf[g - j2] - lf[g - j3];
        const long double v = expl(l);
        return static_cast<double>((g & 1) ? -v : v);
    }
    long double mt = 0;
    const long double s = racah_sum<long double>(j1, j2, j3, m1, m2, m3, log_fact(),
                                                 [](long double v) { return expl(v); }, mt);
    if (mt <= 100.0L * fabsl(s)) return static_cast<double>(s);
    __float128 mq = 0;
    const __float128 q = racah_sum<__float128>(j1, j2, j3, m1, m2, m3, log_fact_q(),
                                               [](__float128 v) { return expq(v); }, mq);
    return static_cast<double>(q);
}

inline double phase(int m) { return (m & 1) ? -1.0 : 1.0; }

static double gaunt_complex_no_conj(int l1, int m1, int l2, int m2, int l3, int m3) {  // Gaunt.hpp:65-90
    if (m1 + m2 + m3 != 0) return 0.0;
    if (std::abs(l1 - l2) > l3) return 0.0;
    if (l1 + l2 < l3) return 0.0;
    if ((l1 + l2 + l3) & 1) return 0.0;
    const double a = w3j(l1, l2, l3, 0, 0, 0);
    const double b = w3j(l1, l2, l3, m1, m2, m3);
    const double pre = std::sqrt((2.0 * l1 + 1.0) * (2.0 * l2 + 1.0) * (2.0 * l3 + 1.0) / (4.0 * M_PI));
    return pre * a * b;
}

static cd U_real_to_complex(int mR, int mC) {  // Gaunt.hpp:96-113
    const double s = M_SQRT1_2;
    if (mR == 0) return (mC == 0) ? cd(1.0, 0.0) : cd(0.0, 0.0);
    if (mR > 0) {
        if (mC == -mR) return {s, 0.0};
        if (mC == mR) return {s * phase(mR), 0.0};
        return {0.0, 0.0};
    }
    const int p = -mR;
    if (mC == -p) return {0.0, s};
    if (mC == p) return {0.0, -s * phase(p)};
    return {0.0, 0.0};
}

static double gaunt_real(int l1, int mR1, int l2, int mR2, int l3, int mR3) {  // Gaunt.hpp:118-140
    if (std::abs(l1 - l2) > l3 || l1 + l2 < l3) return 0.0;
    if ((l1 + l2 + l3) & 1) return 0.0;
    cd sum(0.0, 0.0);
    const int c1[2] = {-std::abs(mR1), std::abs(mR1)};
    const int c2[2] = {-std::abs(mR2), std::abs(mR2)};
    // Only mC = +-|mR| carry a non-zero U (Gaunt.hpp:96-113); visit them in
    // the reference's ascending-m order, once each.
    for (int a = 0; a < 2; ++a) {
        const int m1 = c1[a];
        if (a == 1 && m1 == c1[0]) continue;
        const cd U1 = U_real_to_complex(mR1, m1);
        if (U1.real() == 0.0 && U1.imag() == 0.0) continue;
        for (int b = 0; b < 2; ++b) {
            const int m2 = c2[b];
            if (b == 1 && m2 == c2[0]) continue;
            const cd U2 = U_real_to_complex(mR2, m2);
            if (U2.real() == 0.0 && U2.imag() == 0.0) continue;
            const int m3 = -(m1 + m2);
            if (m3 < -l3 || m3 > l3) continue;
            const cd U3 = U_real_to_complex(mR3, m3);
            if (U3.real() == 0.0 && U3.imag() == 0.0) continue;
            const double Gc = gaunt_complex_no_conj(l1, m1, l2, m2, l3, m3);
            if (Gc == 0.0) continue;
            sum += U1 * U2 * U3 * Gc;
        }
    }
    return sum.real();
}

static double sph_legendre_noCS(int l, int m, double theta) {  // Gaunt.hpp:152-171
    if (l < 0 || m < 0 || m > l) return 0.0;
    const double s = std::sin(theta), c = std::cos(theta);
    double S_mm = 1.0 / std::sqrt(4.0 * M_PI);
    for (int k = 1; k <= m; ++k) S_mm *= std::sqrt((2.0 * k + 1.0) / (2.0 * k)) * s;
    if (l == m) return S_mm;
    double S_prev = S_mm;
    double S_cur = std::sqrt(2.0 * m + 3.0) * c * S_mm;
    if (l == m + 1) return S_cur;
    double S_pp = S_prev;
    S_prev = S_cur;
    for (int k = m + 2; k <= l; ++k) {
        const double a = std::sqrt((4.0 * k * k - 1.0) / (static_cast<double>(k) * k - m * m));
        const double b = std::sqrt((static_cast<double>(k - 1) * (k - 1) - m * m) / (4.0 * (k - 1.0) * (k - 1.0) - 1.0));
        const double S_k = a * (c * S_prev - b * S_pp);
        S_pp = S_prev;
        S_prev = S_k;
    }
    return S_prev;
}
static double real_Ylm(int l, int m, double theta, double phi) {  // Gaunt.hpp:173-181
    if (l < 0 || std::abs(m) > l) return 0.0;
    const int am = std::abs(m);
    const double S = sph_legendre_noCS(l, am, theta);
    if (m == 0) return S;
    if (m > 0) return std::sqrt(2.0) * S * std::cos(m * phi);
    return std::sqrt(2.0) * S * std::sin(am * phi);
}

// ---------------------------------------------------------------------------
// Spherical Bessel (GSL gsl_sf_bessel_jl_e / yl_e) and Riccati forms
// (ForwardRPropagator.cpp:64-76, KMatrixExtractor.cpp:50-73).
// ---------------------------------------------------------------------------
static double sph_j(int l, double x) {
    if (x == 0.0) return (l == 0) ? 1.0 : 0.0;
    const double x2 = x * x;
    if (x2 < 0.1 * (2.0 * l + 3.0) || x < 1e-3) {
        // power series: j_l(x) = x^l/(2l+1)!! * sum_k (-x^2/2)^k / (k! prod_{i=1..k}(2l+2i+1))
        double pre = 1.0;
        for (int i = 1; i <= l; ++i) pre *= x / (2.0 * i + 1.0);
        double term = 1.0, sum = 1.0;
        for (int k = 1; k < 200; ++k) {
            term *= -0.5 * x2 / (k * (2.0 * l + 2.0 * k + 1.0));
            sum += term;
            if (std::abs(term) < 1e-18 * std::abs(sum)) break;
        }
        return pre * sum;
    }
    const double s = std::sin(x), c = std::cos(x);
    const double j0 = s / x;
    if (l == 0) return j0;
    const double j1 = s / x2 - c / x;
    if (l == 1) return j1;
    if (x > l) {  // upward recurrence is stable for l < x
        double jm = j0, jc = j1;
        for (int n = 1; n < l; ++n) { const double jn = (2.0 * n + 1.0) / x * jc - jm; jm = jc; jc = jn; }
        return jc;
    }
    // Miller downward recurrence, normalised to whichever of j0, j1 is larger.
    const int nstart = l + 40 + static_cast<int>(std::sqrt(40.0 * (l + x)));
    double fnext = 0.0, f = 1e-280, fl = 0.0;   // f holds f_n, fnext f_{n+1}
    for (int n = nstart; n >= 1; --n) {
        const double fm = (2.0 * n + 1.0) / x * f - fnext;
        fnext = f; f = fm;                       // now f = f_{n-1}
        if (n - 1 == l) fl = f;
        if (std::abs(f) > 1e250) { f *= 1e-250; fnext *= 1e-250; fl *= 1e-250; }
    }
    // f = f_0, fnext = f_1
    const double scale = (std::abs(j0) >= std::abs(j1)) ? j0 / f : j1 / fnext;
    return fl * scale;
}
static double sph_y(int l, double x) {
    const double s = std::sin(x), c = std::cos(x);
    const double y0 = -c / x;
    if (l == 0) return y0;
    const double y1 = -c / (x * x) - s / x;
    double ym = y0, yc = y1;
    for (int n = 1; n < l; ++n) { const double yn = (2.0 * n + 1.0) / x * yc - ym; ym = yc; yc = yn; }
    return yc;
}
static double riccati_j(int l, double x) { if (x < 1e-300) return 0.0; return x * sph_j(l, x); }
static double riccati_y(int l, double x) {
    if (x < 1e-300) throw std::runtime_error("riccati_y called at x ~ 0");
    return x * sph_y(l, x);
}
// d/dx of the Riccati forms u_l = x f_l:  u_l' = u_{l-1} - (l/x) u_l, with
// u_{-1} = cos x (j) and sin x (y, where u_0 = -cos x).
static double riccati_jp(int l, double x) {
    if (l == 0) return std::cos(x);
    if (x < 1e-300) return 0.0;
    return riccati_j(l - 1, x) - static_cast<double>(l) / x * riccati_j(l, x);
}
static double riccati_yp(int l, double x) {
    if (l == 0) return std::sin(x);
    return riccati_y(l - 1, x) - static_cast<double>(l) / x * riccati_y(l, x);
}

// ---------------------------------------------------------------------------
// Quadrature helpers (DipoleMatrixElement.cpp:49-74, :142-192).
// ---------------------------------------------------------------------------
static void five_point_derivative(const double* f, int n, double dr, double* df) {
    if (n < 5) {
        if (n >= 2) {
            df[0] = (f[1] - f[0]) / dr;
            for (int i = 1; i < n - 1; ++i) df[i] = (f[i + 1] - f[i - 1]) / (2.0 * dr);
            df[n - 1] = (f[n - 1] - f[n - 2]) / dr;
        } else if (n == 1) df[0] = 0.0;
        return;
    }
    const double inv_12h = 1.0 / (12.0 * dr);
    for (int i = 2; i < n - 2; ++i) df[i] = (-f[i + 2] + 8.0 * f[i + 1] - 8.0 * f[i - 1] + f[i - 2]) * inv_12h;
    df[0] = (-25.0 * f[0] + 48.0 * f[1] - 36.0 * f[2] + 16.0 * f[3] - 3.0 * f[4]) * inv_12h;
    df[1] = (-3.0 * f[0] - 10.0 * f[1] + 18.0 * f[2] - 6.0 * f[3] + 1.0 * f[4]) * inv_12h;
    df[n - 2] = (-1.0 * f[n - 5] + 6.0 * f[n - 4] - 18.0 * f[n - 3] + 10.0 * f[n - 2] + 3.0 * f[n - 1]) * inv_12h;
    df[n - 1] = (3.0 * f[n - 5] - 16.0 * f[n - 4] + 36.0 * f[n - 3] - 48.0 * f[n - 2] + 25.0 * f[n - 1]) * inv_12h;
}
static double simpson(const double* f, int n_pts, double h) {
    if (n_pts < 2) return 0.0;
    auto one_third = [&](int n) {
        double s = f[0] + f[n - 1];
        for (int i = 1; i < n - 1; i += 2) s += 4.0 * f[i];
        for (int i = 2; i < n - 2; i += 2) s += 2.0 * f[i];
        return s * h / 3.0;
    };
    if (n_pts == 2) return 0.5 * h * (f[0] + f[1]);
    if (n_pts % 2 == 1) return one_third(n_pts);
    if (n_pts < 4) {
        double s = 0.5 * (f[0] + f[n_pts - 1]);
        for (int i = 1; i < n_pts - 1; ++i) s += f[i];
        return s * h;
    }
    const double a = one_third(n_pts - 3);
    const double b = (3.0 * h / 8.0) * (f[n_pts - 4] + 3.0 * f[n_pts - 3] + 3.0 * f[n_pts - 2] + f[n_pts - 1]);
    return a + b;
}

// ---------------------------------------------------------------------------
// The solver.
// ---------------------------------------------------------------------------
struct GauntRow { std::vector<int> col; std::vector<double> val; };

static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

class Solver {
public:
    const sehf_oracle_problem& P;
    int l_cont, l_exch, l_orb, n_mu, n_sigma, n_lambda, l_exp, n_exp, n_occ, Nr, N_f, N_tot;
    double dr;
    int n_tr = 0;                       // n_transition
    std::vector<int> l_mu, m_mu, l_sig;
    // potential sparse Gaunt (GauntSparse.hpp:81-200): rows = upper-triangle pairs
    std::vector<std::pair<int, int>> pair_to_mu;
    std::vector<GauntRow> gsp;
    // exchange G_combined (ExchangeCoupling.cpp:13-54): row mu*n_sigma+sigma, col lambda
    std::vector<GauntRow> gcomb;
    int n_lambda_max = 0;
    // chi[ir] (n_occ x n_lambda) (WavefunctionSetup.cpp:155-214)
    std::vector<Mat> chi;

    static constexpr double W_min = 0.001;       // SchurInverter.hpp:74
    static constexpr int stab_n_max = 60;         // SchurInverter.hpp:75
    static constexpr double shift_margin = 0.01;  // SchurInverter.hpp:76

    explicit Solver(const sehf_oracle_problem& p) : P(p) {
        l_cont = p.l_cont;
        l_exch = std::min(l_cont, 10);                        // WavefunctionSetup.cpp:238-240
        l_orb = l_cont + l_exch;
        n_mu = (l_cont + 1) * (l_cont + 1);
        n_sigma = (l_exch + 1) * (l_exch + 1);
        n_lambda = (l_orb + 1) * (l_orb + 1);
        l_exp = 2 * l_cont;                                   // Parameters.hpp:69
        n_exp = (l_exp + 1) * (l_exp + 1);
        n_occ = p.n_occ; Nr = p.Nr; dr = p.dr;
        N_f = n_occ * n_sigma; N_tot = n_mu + N_f;
        if (n_lambda > p.nlm_orb) throw std::runtime_error("l_max_orbitals > stored orbital channels");
        if (p.nlm_vh < n_exp) throw std::runtime_error("V_H has fewer rows than (2 l_cont + 1)^2");
        l_mu.resize(n_mu); m_mu.resize(n_mu);
        for (int mu = 0; mu < n_mu; ++mu) idx_to_lm(mu, l_mu[mu], m_mu[mu]);
        l_sig.resize(n_sigma);
        for (int s = 0; s < n_sigma; ++s) { int l, m; idx_to_lm(s, l, m); l_sig[s] = l; }
        const double t0 = now_s();
        build_potential_gaunt();
        const double t1 = now_s();
        prepare_setup();
        if (std::getenv("SEHF_ORACLE_VERBOSE"))
            std::fprintf(stderr, "[oracle] setup: potential Gaunt %.2f s, exchange tables %.2f s\n", t1 - t0, now_s() - t1);
    }

    // ---- GauntSparseMatrix::build (GauntSparse.hpp:81-200) ----
    void build_potential_gaunt() {
        const int ch = n_mu;
        const long n_pairs = static_cast<long>(ch) * (ch + 1) / 2;
        pair_to_mu.resize(n_pairs);
        auto pair_index = [&](int mu_p, int mu) -> long {
            return static_cast<long>(mu_p) * ch - (static_cast<long>(mu_p) * (mu_p - 1)) / 2 + (mu - mu_p);
        };
        for (int mp = 0; mp < ch; ++mp) for (int mu = mp; mu < ch; ++mu) pair_to_mu[pair_index(mp, mu)] = {mp, mu};
        gsp.assign(n_pairs, GauntRow{});
        std::vector<std::vector<std::pair<long, std::pair<int, double>>>> loc;
        int nthr = 1;
#ifdef _OPENMP
        nthr = omp_get_max_threads();
#endif
        loc.resize(nthr);
#pragma omp parallel for schedule(dynamic, 1)
        for (int sigma = 0; sigma < n_exp; ++sigma) {
            int tid = 0;
#ifdef _OPENMP
            tid = omp_get_thread_num();
#endif
            int l_s, m_s; idx_to_lm(sigma, l_s, m_s);
            for (int l2 = 0; l2 <= l_cont; ++l2) {
                const int l1_min = std::abs(l2 - l_s), l1_max = std::min(l_cont, l2 + l_s);
                for (int m2 = -l2; m2 <= l2; ++m2) {
                    const int mu_p = lm_to_idx(l2, m2);
                    for (int l1 = l1_min; l1 <= l1_max; ++l1) {
                        if (((l1 + l_s + l2) & 1) != 0) continue;
                        // real-harmonic selection rule: G = 0 unless
                        // |m1| in {|m2|+|m_s|, ||m2|-|m_s||} (the skipped calls are exact zeros)
                        for (int m1 = -l1; m1 <= l1; ++m1) {
                            const int am = std::abs(m1);
                            if (am != std::abs(m2) + std::abs(m_s) && am != std::abs(std::abs(m2) - std::abs(m_s))) continue;
                            const int mu = lm_to_idx(l1, m1);
                            if (mu < mu_p) continue;
                            const double G = gaunt_real(l2, m2, l_s, m_s, l1, m1);
                            if (std::abs(G) > 1e-15) loc[tid].push_back({pair_index(mu_p, mu), {sigma, G}});
                        }
                    }
                }
            }
        }
        for (auto& v : loc) for (auto& t : v) { gsp[t.first].col.push_back(t.second.first); gsp[t.first].val.push_back(t.second.second); }
        // Eigen's setFromTriplets leaves each row sorted by column.
        for (auto& r : gsp) sort_row(r);
    }
    static void sort_row(GauntRow& r) {
        std::vector<int> idx(r.col.size());
        for (size_t i = 0; i < idx.size(); ++i) idx[i] = static_cast<int>(i);
        std::sort(idx.begin(), idx.end(), [&](int a, int b) { return r.col[a] < r.col[b]; });
        GauntRow s; s.col.reserve(idx.size()); s.val.reserve(idx.size());
        for (int i : idx) { s.col.push_back(r.col[i]); s.val.push_back(r.val[i]); }
        r = std::move(s);
    }

    // ---- Potentials::compute_V_sigma_total (Potentials.cpp:40-88) ----
    std::vector<double> V_sigma_total(double r, int ir) const {
        std::vector<double> Vs(n_exp, 0.0);
        const int nmax_ee = std::min(n_exp, P.nlm_vh);
        for (int s = 0; s < nmax_ee; ++s) Vs[s] += P.V_H[static_cast<size_t>(s) * Nr + ir];
        for (int a = 0; a < P.n_atoms; ++a) {
            const double Z = static_cast<double>(P.Z[a]);
            const double Rx = P.xyz[3 * a], Ry = P.xyz[3 * a + 1], Rz = P.xyz[3 * a + 2];
            const double R = std::sqrt(Rx * Rx + Ry * Ry + Rz * Rz);
            double th = 0.0, ph = 0.0;
            if (R > 1e-14) { th = std::acos(std::clamp(Rz / R, -1.0, 1.0)); ph = std::atan2(Ry, Rx); }
            const double rl = std::min(r, R), rg = std::max(r, R);
            for (int l = 0; l <= l_exp; ++l) {
                const double radial = (4.0 * M_PI / (2.0 * l + 1.0)) * std::pow(rl, l) / std::pow(rg, l + 1);
                for (int m = -l; m <= l; ++m) {
                    double Y;
                    if (R < 1e-14) Y = (l == 0) ? (1.0 / (2.0 * std::sqrt(M_PI))) : 0.0;
                    else Y = real_Ylm(l, m, th, ph);
                    if (std::abs(Y) > 1e-15) Vs[lm_to_idx(l, m)] += -Z * radial * Y;
                }
            }
        }
        return Vs;
    }
    // ---- Potentials::build assemble_V_at (Potentials.cpp:324-365) ----
    Mat potential(int ir) const {
        const double r = 0.0 + ir * dr;   // Parameters::r_min = 0 (preproc rmin)
        Mat V(n_mu, n_mu);
        if (!(r > 1e-14)) return V;
        const std::vector<double> Vs = V_sigma_total(r, ir);
        for (size_t p = 0; p < gsp.size(); ++p) {     // GauntSparse.hpp:212-222
            double acc = 0.0;
            const GauntRow& g = gsp[p];
            for (size_t t = 0; t < g.col.size(); ++t) acc += g.val[t] * Vs[g.col[t]];
            const auto pr = pair_to_mu[p];
            V(pr.first, pr.second) = acc;
            if (pr.first != pr.second) V(pr.second, pr.first) = acc;
        }
        const double vpol = polarization(r);        // Potentials.cpp:111-129 (diag only)
        if (vpol != 0.0) for (int mu = 0; mu < n_mu; ++mu) V(mu, mu) += vpol;
        const double inv_r2 = 1.0 / (r * r);
        for (int mu = 0; mu < n_mu; ++mu) V(mu, mu) += l_mu[mu] * (l_mu[mu] + 1.0) * 0.5 * inv_r2;
        return V;
    }
    double polarization(double r) const {
        if (P.pol_kind == 1) {                     // Potentials.cpp:111-129
            if (P.pol_alpha == 0.0 || r < 1.0e-6) return 0.0;
            const double u = r / 1.5, u6 = u * u * u * u * u * u;
            const double damp = 1.0 - std::exp(-u6);
            return -0.5 * P.pol_alpha * damp * damp / (r * r * r * r);
        }
        if (P.pol_kind == 2) {                     // soft form, same diagonal slot
            if (r < 1.0e-6) return 0.0;
            const double q = r * r + P.pol_rc * P.pol_rc;
            return -P.pol_alpha / (2.0 * q * q);
        }
        return 0.0;
    }
    // dV_pol/d alpha at potential index n (both forms are linear in alpha;
    // V is zero at r = 0)
    double pol_unit_at(int n) const {
        const double r = n * dr;
        if (!(r > 1e-14) || r < 1.0e-6) return 0.0;
        if (P.pol_kind == 1) {
            const double u = r / 1.5, u6 = u * u * u * u * u * u;
            const double damp = 1.0 - std::exp(-u6);
            return -0.5 * damp * damp / (r * r * r * r);
        }
        if (P.pol_kind == 2) {
            const double q = r * r + P.pol_rc * P.pol_rc;
            return -1.0 / (2.0 * q * q);
        }
        return 0.0;
    }

    // ---- WavefunctionSetup::prepare + build_G_vector (WavefunctionSetup.cpp:21-81, :219-376) ----
    void prepare_setup() {
        const int Nlm_store = P.nlm_orb, Nr_store = Nr;
        const int n_lscan = std::min(n_lambda, Nlm_store);
        int last_ir = 0;
        for (int ir = Nr_store - 1; ir >= 0; --ir) {   // chi_cutoff = 1e-15 scan
            double mx = 0.0;
            for (int j = 0; j < n_occ; ++j)
                for (int lam = 0; lam < n_lscan; ++lam)
                    mx = std::max(mx, std::abs(P.psi_lm[(static_cast<size_t>(j) * Nlm_store + lam) * Nr + ir]));
            if (mx > 1e-15) { last_ir = ir; break; }
        }
        const int buffer = std::max(100, static_cast<int>(5.0 / dr));
        n_tr = std::min({Nr, Nr_store, last_ir + 1 + buffer});
        if (n_tr < 1) n_tr = 1;
        chi.assign(n_tr, Mat(n_occ, n_lambda));
        for (int ir = 0; ir < n_tr; ++ir) {
            const double r = 0.0 + ir * dr;
            for (int j = 0; j < n_occ; ++j)
                for (int lam = 0; lam < n_lambda; ++lam)
                    chi[ir](j, lam) = r * P.psi_lm[(static_cast<size_t>(j) * Nlm_store + lam) * Nr + ir];
        }
        // G triplets (mu, lambda, sigma) with |G| > 1e-14, stacked into
        // G_combined[mu*n_sigma + sigma, lambda] (ExchangeCoupling.cpp:13-54).
        gcomb.assign(static_cast<size_t>(n_mu) * n_sigma, GauntRow{});
        n_lambda_max = 0;
#pragma omp parallel for schedule(dynamic, 1)
        for (int mu = 0; mu < n_mu; ++mu) {
            const int lm = l_mu[mu], mm = m_mu[mu];
            for (int lam = 0; lam < n_lambda; ++lam) {
                int ll, ml; idx_to_lm(lam, ll, ml);
                for (int s = 0; s < n_sigma; ++s) {
                    int ls, ms; idx_to_lm(s, ls, ms);
                    // triangle, parity and real-harmonic m rules: the skipped
                    // calls are exact zeros, so the table is unchanged
                    if (ll < std::abs(lm - ls) || ll > lm + ls || ((lm + ll + ls) & 1)) continue;
                    const int am = std::abs(ml);
                    if (am != std::abs(mm) + std::abs(ms) && am != std::abs(std::abs(mm) - std::abs(ms))) continue;
                    const double val = gaunt_real(lm, mm, ll, ml, ls, ms);
                    if (std::abs(val) > 1e-14) {
                        GauntRow& row = gcomb[static_cast<size_t>(mu) * n_sigma + s];
                        row.col.push_back(lam); row.val.push_back(val);
                    }
                }
            }
        }
        for (auto& r : gcomb) for (int c : r.col) n_lambda_max = std::max(n_lambda_max, c + 1);
    }

    // ---- ExchangeCoupling::compute_into (ExchangeCoupling.cpp:66-136) ----
    Mat exchange_Q(int ir) const {
        Mat Q(n_mu, N_f);
        const double r = 0.0 + ir * dr;   // EC gets data.rmin (main.cpp:1131-1133)
        if (r < 1e-10 || ir >= n_tr) return Q;
        const double factor = 2.0 * std::sqrt(2.0 * M_PI) / r;
        const Mat& c = chi[ir];
        for (int mu = 0; mu < n_mu; ++mu)
            for (int s = 0; s < n_sigma; ++s) {
                const GauntRow& g = gcomb[static_cast<size_t>(mu) * n_sigma + s];
                for (int i = 0; i < n_occ; ++i) {
                    double acc = 0.0;
                    for (size_t t = 0; t < g.col.size(); ++t) acc += g.val[t] * c(i, g.col[t]);
                    Q(mu, i * n_sigma + s) = factor * acc;
                }
            }
        return Q;
    }

    // ---- D^{-1} with Johnson clamp at solver radius r=(n+1)dr (SchurInverter.cpp:332-339) ----
    std::vector<double> Dinv(int n) const {
        const double r = dr + n * dr, r2 = r * r, h2_12 = dr * dr / 12.0;
        std::vector<double> d(N_f);
        for (int f = 0; f < N_f; ++f) {
            const int l = l_sig[f % n_sigma];
            const double centrif = (r2 > 1e-30) ? static_cast<double>(l) * (l + 1) / r2 : 0.0;
            double Df = 1.0 - h2_12 * centrif;
            if (Df < W_min) Df = W_min;
            d[f] = 1.0 / Df;
        }
        return d;
    }
    Mat Bmat(int n) const {   // B = (h^2/12) Q(n), zero past n_transition (WInverseOperator.cpp:39-62)
        Mat B(n_mu, N_f);
        if (n >= n_tr) return B;
        B = exchange_Q(n);
        const double h2_12 = dr * dr / 12.0;
        for (double& v : B.a) v *= h2_12;
        return B;
    }

    // ---- SchurInverter stability helpers (SchurInverter.cpp:30-59) ----
    static double gershgorin_min(const Mat& M) {
        const int n = M.r; double lb = 1e300;
        for (int i = 0; i < n; ++i) {
            double off = 0.0;
            for (int j = 0; j < n; ++j) if (i != j) off += std::abs(M(i, j));
            lb = std::min(lb, M(i, i) - off);
        }
        return lb;
    }
    static double min_eig_fast(const Mat& M, double thr) {
        const double g = gershgorin_min(M);
        if (g > thr + 0.05) return g;
        return min_eig_sym(M);
    }
    static bool stability_shift(Mat& M) {
        const double me = min_eig_fast(M, W_min);
        if (me >= W_min) return false;
        const double delta = W_min - me + shift_margin;
        for (int i = 0; i < M.r; ++i) M(i, i) += delta;
        return true;
    }

    // ---- SchurInverter::compute_sinv_at (SchurInverter.cpp:247-415) ----
    Mat sinv_at(int n, double E, int& shiftA, int& shiftS) const {
        const double h = dr, h2_6 = h * h / 6.0;
        Mat A(n_mu, n_mu);
        for (int i = 0; i < n_mu; ++i) A(i, i) = 1.0;
        for (int i = 0; i < n_mu; ++i) A(i, i) += h2_6 * E;
        {
            const Mat V = potential(n);
            for (size_t t = 0; t < A.a.size(); ++t) A.a[t] -= h2_6 * V.a[t];
        }
        if (n < stab_n_max) {
            symmetrize(A);
            if (stability_shift(A)) ++shiftA;
        }
        const bool exchange_on = (n < n_tr);
        Mat S;
        if (!exchange_on) {
            S = A;
        } else {
            const std::vector<double> Di = Dinv(n);
            const Mat B = Bmat(n);
            Mat Bs = B;
            for (int f = 0; f < N_f; ++f) for (int i = 0; i < n_mu; ++i) Bs(i, f) *= Di[f];
            S = A;
            const Mat Bt = transpose(B);
            gemm(-1.0, Bs, Bt, 1.0, S);
            if (n < stab_n_max) { if (stability_shift(S)) ++shiftS; }
        }
        return inverse_symmetric_indefinite(S);
    }

    // ---- WInverseOperator::apply (WInverseOperator.cpp:75-128) ----
    // Y = W_n^{-1} Z, with Sinv, B(n), D^{-1}(n).
    void winv_apply(int n, const Mat& Sinv, const Mat& Z, Mat& Y) const {
        struct RevScope {   // bit 6 of the rounding-order probes: reversed sums here
            bool old = g_gemm_rev;
            RevScope() { if (g_inverse_mode & 64) g_gemm_rev = true; }
            ~RevScope() { g_gemm_rev = old; }
        } rev_scope;
        const int nc = Z.c;
        Y.resize(N_tot, nc);
        const std::vector<double> Di = Dinv(n);
        Mat tmp_f(N_f, nc), rhs(n_mu, nc);
        for (int j = 0; j < nc; ++j) {
            for (int f = 0; f < N_f; ++f) tmp_f(f, j) = Di[f] * Z(n_mu + f, j);
            for (int i = 0; i < n_mu; ++i) rhs(i, j) = Z(i, j);
        }
        const bool ex = (n < n_tr);
        Mat B;
        if (ex) { B = Bmat(n); gemm(-1.0, B, tmp_f, 1.0, rhs); }
        Mat Ypsi = mul(Sinv, rhs);
        Mat Yf = tmp_f;
        if (ex) {
            Mat BtY = mul(transpose(B), Ypsi);
            for (int j = 0; j < nc; ++j) for (int f = 0; f < N_f; ++f) Yf(f, j) = tmp_f(f, j) - Di[f] * BtY(f, j);
        }
        for (int j = 0; j < nc; ++j) {
            for (int i = 0; i < n_mu; ++i) Y(i, j) = Ypsi(i, j);
            for (int f = 0; f < N_f; ++f) Y(n_mu + f, j) = Yf(f, j);
        }
    }

    // ---- ForwardRPropagator (ForwardRPropagator.cpp:49-134, :299-386) ----
    int n_start() const {
        int lmf = 0; for (int s = 0; s < n_sigma; ++s) lmf = std::max(lmf, l_sig[s]);
        const double r_crit = dr * std::sqrt(static_cast<double>(lmf * (lmf + 1)) / 12.0);
        int ns = std::max(1, static_cast<int>(std::ceil(2.0 * r_crit / dr)));
        ns = std::min(ns, Nr - 1);
        return ns;
    }
    Mat analytic_init(int n, double E) const {
        Mat R(N_tot, N_tot);
        if (n <= 0) return R;
        const double h = dr, h2_12 = h * h / 12.0, h2_6 = h * h / 6.0;
        const double k = std::sqrt(2.0 * E);
        const double r_n = dr + n * h, r_np1 = dr + (n + 1) * h;
        const Mat Vn = potential(n), Vnp1 = potential(n + 1);
        for (int mu = 0; mu < n_mu; ++mu) {
            const double Mn = 1.0 + h2_6 * (E - Vn(mu, mu));
            const double Mnp1 = 1.0 + h2_6 * (E - Vnp1(mu, mu));
            const double yn = riccati_j(l_mu[mu], k * r_n), ynp1 = riccati_j(l_mu[mu], k * r_np1);
            const double den = Mnp1 * ynp1;
            if (std::abs(den) > 1e-300) R(mu, mu) = (Mn * yn) / den;
        }
        const double rn2 = r_n * r_n, rnp12 = r_np1 * r_np1;
        for (int f = 0; f < N_f; ++f) {
            const int l = l_sig[f % n_sigma];
            const double cn = (rn2 > 1e-30) ? double(l * (l + 1)) / rn2 : 0.0;
            const double cnp1 = (rnp12 > 1e-30) ? double(l * (l + 1)) / rnp12 : 0.0;
            double Mn = 1.0 - h2_12 * cn, Mnp1 = 1.0 - h2_12 * cnp1;
            if (Mn < W_min) Mn = W_min;
            if (Mnp1 < W_min) Mnp1 = W_min;
            const double yn = std::pow(r_n, static_cast<double>(l + 1)), ynp1 = std::pow(r_np1, static_cast<double>(l + 1));
            const double den = Mnp1 * ynp1;
            if (std::abs(den) > 1e-300) R(n_mu + f, n_mu + f) = (Mn * yn) / den;
        }
        return R;
    }
    // Tangent of analytic_init along (E, alpha): dk = tE / k, and the
    // potential-index weights w_n = tE - tA * dV_pol/dalpha enter M_n.
    Mat analytic_init_tangent(int n, double E, double tE, double tA) const {
        Mat dR(N_tot, N_tot);
        if (n <= 0) return dR;
        const double h = dr, h2_6 = h * h / 6.0;
        const double k = std::sqrt(2.0 * E), dk = tE / k;
        const double r_n = dr + n * h, r_np1 = dr + (n + 1) * h;
        const double wn = tE - tA * pol_unit_at(n), wn1 = tE - tA * pol_unit_at(n + 1);
        const Mat Vn = potential(n), Vnp1 = potential(n + 1);
        for (int mu = 0; mu < n_mu; ++mu) {
            const int l = l_mu[mu];
            const double Mn = 1.0 + h2_6 * (E - Vn(mu, mu));
            const double Mnp1 = 1.0 + h2_6 * (E - Vnp1(mu, mu));
            const double yn = riccati_j(l, k * r_n), ynp1 = riccati_j(l, k * r_np1);
            const double dyn = riccati_jp(l, k * r_n) * r_n * dk, dynp1 = riccati_jp(l, k * r_np1) * r_np1 * dk;
            const double a = Mn * yn, b = Mnp1 * ynp1;
            const double da = h2_6 * wn * yn + Mn * dyn, db = h2_6 * wn1 * ynp1 + Mnp1 * dynp1;
            if (std::abs(b) > 1e-300) dR(mu, mu) = (da * b - a * db) / (b * b);
        }
        return dR;   // the f block does not depend on E or alpha
    }
    // One FRP step: Rinv_n = sym(inv(12 W^{-1} - 10 I - Rinv_prev)).
    Mat frp_step(int n, const Mat& Sinv, const Mat& Rprev) const {
        Mat Y;
        winv_apply(n, Sinv, identity(N_tot), Y);
        Mat R(N_tot, N_tot);
        for (int j = 0; j < N_tot; ++j)
            for (int i = 0; i < N_tot; ++i)
                R(i, j) = (12.0 * Y(i, j) - 10.0 * (i == j ? 1.0 : 0.0)) - Rprev(i, j);
        Mat Ri = inverse_general(R, INV_FRP);
        symmetrize(Ri);
        return Ri;
    }
};

struct Scan {
    // full per-energy run (main.cpp:1081-1623 wiring)
    static void run(const Solver& S, double k, sehf_oracle_out* out);
};


void Scan::run(const Solver& S, double k, sehf_oracle_out* out) {
    const double t0 = now_s();
    const double E = 0.5 * k * k;
    const int Np = S.n_mu, Nf = S.N_f, Nt = S.N_tot, Nr = S.Nr, n_occ = S.n_occ;
    const double h = S.dr, h2_6 = h * h / 6.0;
    double ts[8] = {0};
    // Tangent along (E, alpha) = (tE, tA) when requested.  Every E- or
    // alpha-dependence of A_n = I + h^2/6 (E - V_n) is a multiple of I,
    // dA_n = h^2/6 w_n I with w_n = tE - tA dV_pol/dalpha(n); a stability
    // shift rebuilds A (or S) as M - lambda_min(M) I + const, which removes
    // that dependence (sig[n] = 0).  The Riccati functions move with
    // dk = tE / k.
    const bool tan = out->tangent != 0;
    const double tE = out->tan_E, tA = out->tan_alpha, dk = tE / k;
    auto wn = [&](int n) { return tE - tA * S.pol_unit_at(n); };
    // ---- Sinv (SchurInverter::build, SchurInverter.cpp:465-473) and FRP
    // (ForwardRPropagator::run, ForwardRPropagator.cpp:136-401) in one sweep.
    // The reference keeps every Sinv_n and Rinv_n (disk or RAM); here only
    // Rinv at every L-th point is kept (L = ceil(sqrt(Nr-1)) > n_start) and
    // each segment is recomputed by the same operations during the
    // back-propagation below, so every value is bitwise the same as with
    // the full store.
    const int ns = S.n_start();
    const int N = Nr - 1;
    int L = static_cast<int>(std::ceil(std::sqrt(static_cast<double>(N))));
    L = std::max(L, ns + 1);
    const int n_seg = (N + L - 1) / L;
    std::vector<Mat> ck(n_seg), dck(tan ? n_seg : 0);   // ck[s] = Rinv_{sL-1} (s >= 1)
    std::vector<char> sig(Nr, 1);
    int shA = 0, shS = 0;
    Mat Epsi(Nt, Np);
    for (int i = 0; i < Np; ++i) Epsi(i, i) = 1.0;
    auto dump_at = [](int nd, const int* idx, int n) {
        for (int i = 0; i < nd; ++i) if (idx[i] == n) return i;
        return -1;
    };
    auto put_at = [](double* dst, size_t stride, int i, const Mat& M) {
        if (dst && i >= 0) std::memcpy(dst + static_cast<size_t>(i) * stride, M.a.data(), M.a.size() * sizeof(double));
    };
    // one ratio step's tangent: dRinv_n = Rinv_n (dRinv_{n-1} + 12 sig h^2/6 w_n G G^T) Rinv_n,
    // G = W_n^{-1}[:, psi] (the derivative of 12 W^{-1} - 10 I - Rinv_{n-1})
    auto frp_tan = [&](int n, const Mat& sv, const Mat& Rn, const Mat* dprev) {
        Mat G;
        S.winv_apply(n, sv, Epsi, G);
        Mat H = dprev ? *dprev : Mat(Nt, Nt);
        gemm(12.0 * (sig[n] ? 1.0 : 0.0) * h2_6 * wn(n), G, transpose(G), 1.0, H);
        const Mat T = mul(H, Rn);
        Mat dR = mul(Rn, T);
        symmetrize(dR);
        return dR;
    };
    Mat Rprev(Nt, Nt), dRprev;
    for (int n = 0; n < Nr; ++n) {
        const int a0 = shA, s0 = shS;
        const Mat sv = S.sinv_at(n, E, shA, shS);
        sig[n] = (shA == a0 && shS == s0) ? 1 : 0;
        put_at(out->dump_sinv, static_cast<size_t>(Np) * Np, dump_at(out->n_dump_sinv, out->dump_sinv_n, n), sv);
        if (n > N - 1) continue;   // Rinv_N is not used
        Mat Rn, dRn;
        if (n < ns) {
            Rn = S.analytic_init(n, E);
            if (tan) dRn = S.analytic_init_tangent(n, E, tE, tA);
        } else {
            Rn = S.frp_step(n, sv, Rprev);
            if (tan) dRn = frp_tan(n, sv, Rn, n >= 1 ? &dRprev : nullptr);
        }
        const int di = dump_at(out->n_dump_rinv, out->dump_rinv_n, n);
        put_at(out->dump_rinv, static_cast<size_t>(Nt) * Nt, di, Rn);
        if (tan) put_at(out->dump_drinv, static_cast<size_t>(Nt) * Nt, di, dRn);
        Rprev = std::move(Rn);
        if (tan) dRprev = std::move(dRn);
        if ((n + 1) % L == 0 && (n + 1) / L < n_seg) {
            ck[(n + 1) / L] = Rprev;
            if (tan) dck[(n + 1) / L] = dRprev;
        }
    }
    const Mat& R_last = Rprev;    // Rinv_{N-1}
    const Mat& dR_last = dRprev;
    const double t1 = now_s(); ts[1] = t1 - t0;
    const double t2 = t1; ts[2] = 0.0;
    // ---- KMatrixExtractor::extract (KMatrixExtractor.cpp:75-198) ----
    const int nM = Nr - 2;
    Mat Kk, dKk;
    double unit_err_kme = 0.0;
    {
        const Mat R = inverse_general(R_last, INV_KME);
        Mat Ssch(Np, Np);
        double npf = 0.0, nfp = 0.0;
        for (int j = Np; j < Nt; ++j) for (int i = 0; i < Np; ++i) npf += R(i, j) * R(i, j);
        for (int j = 0; j < Np; ++j) for (int i = Np; i < Nt; ++i) nfp += R(i, j) * R(i, j);
        npf = std::sqrt(npf); nfp = std::sqrt(nfp);
        for (int j = 0; j < Np; ++j) for (int i = 0; i < Np; ++i) Ssch(i, j) = R(i, j);
        if (Nf > 0 && !(npf < 1e-14 && nfp < 1e-14)) {
            Mat Rff(Nf, Nf), Rpf(Np, Nf), Rfp(Nf, Np);
            for (int j = 0; j < Nf; ++j) for (int i = 0; i < Nf; ++i) Rff(i, j) = R(Np + i, Np + j);
            for (int j = 0; j < Nf; ++j) for (int i = 0; i < Np; ++i) Rpf(i, j) = R(i, Np + j);
            for (int j = 0; j < Np; ++j) for (int i = 0; i < Nf; ++i) Rfp(i, j) = R(Np + i, j);
            const Mat Rffi = inverse_general(Rff, INV_KME);
            const Mat T = mul(Rpf, Rffi);
            gemm(-1.0, T, Rfp, 1.0, Ssch);
        }
        symmetrize(Ssch);
        const double h2_12 = h * h / 12.0, k2 = k * k;
        const double r_in = h + nM * h, r_out = h + (nM + 1) * h;
        std::vector<double> Jin(Np), Jout(Np), Yin(Np), Yout(Np), Win(Np), Wout(Np);
        for (int mu = 0; mu < Np; ++mu) {
            const int l = S.l_mu[mu];
            Jin[mu] = riccati_j(l, k * r_in); Jout[mu] = riccati_j(l, k * r_out);
            Yin[mu] = riccati_y(l, k * r_in); Yout[mu] = riccati_y(l, k * r_out);
            Win[mu] = 1.0 + h2_12 * (k2 - static_cast<double>(l * (l + 1)) / (r_in * r_in));
            Wout[mu] = 1.0 + h2_12 * (k2 - static_cast<double>(l * (l + 1)) / (r_out * r_out));
        }
        Mat Am(Np, Np), Bm(Np, Np);
        for (int j = 0; j < Np; ++j)
            for (int i = 0; i < Np; ++i) {
                Am(i, j) = ((i == j) ? Wout[j] * Jout[j] : 0.0) - Ssch(i, j) * (Win[j] * Jin[j]);
                Bm(i, j) = Ssch(i, j) * (Win[j] * Yin[j]) - ((i == j) ? Wout[j] * Yout[j] : 0.0);
            }
        Mat lu = Bm; std::vector<int> perm; lu_partial(lu, perm);
        Kk = Am; lu_solve(lu, perm, Kk);
        symmetrize(Kk);
        if (tan) {
            // S_schur = ((Rinv_{nM})_psi,psi)^{-1}, so dS = -S (dRinv)_psi,psi S;
            // dK = sym(Bm^{-1} (dAm - dBm K)).
            Mat dRpp(Np, Np);
            for (int j = 0; j < Np; ++j) for (int i = 0; i < Np; ++i) dRpp(i, j) = dR_last(i, j);
            Mat dSs(Np, Np);
            gemm(-1.0, mul(Ssch, dRpp), Ssch, 0.0, dSs);
            Mat dAm(Np, Np), dBm(Np, Np);
            const double dW = h2_12 * 2.0 * k * dk;
            for (int j = 0; j < Np; ++j) {
                const int l = S.l_mu[j];
                const double dJin = riccati_jp(l, k * r_in) * r_in * dk, dJout = riccati_jp(l, k * r_out) * r_out * dk;
                const double dYin = riccati_yp(l, k * r_in) * r_in * dk, dYout = riccati_yp(l, k * r_out) * r_out * dk;
                const double WJin = Win[j] * Jin[j], dWJin = dW * Jin[j] + Win[j] * dJin;
                const double WYin = Win[j] * Yin[j], dWYin = dW * Yin[j] + Win[j] * dYin;
                for (int i = 0; i < Np; ++i) {
                    dAm(i, j) = ((i == j) ? dW * Jout[j] + Wout[j] * dJout : 0.0) - dSs(i, j) * WJin - Ssch(i, j) * dWJin;
                    dBm(i, j) = dSs(i, j) * WYin + Ssch(i, j) * dWYin - ((i == j) ? dW * Yout[j] + Wout[j] * dYout : 0.0);
                }
            }
            Mat rhs = dAm;
            gemm(-1.0, dBm, Kk, 1.0, rhs);
            lu_solve(lu, perm, rhs);
            symmetrize(rhs);
            dKk = rhs;
        }
        // S = (I + iK)(I - iK)^{-1}; unitarity diagnostic
        CMat lhs(Np, Np), rhs(Np, Np);
        for (int j = 0; j < Np; ++j) for (int i = 0; i < Np; ++i) {
            lhs(i, j) = cd(i == j ? 1.0 : 0.0, Kk(i, j)); rhs(i, j) = cd(i == j ? 1.0 : 0.0, -Kk(i, j));
        }
        const CMat ri = inverse_pp(rhs);
        CMat Sm(Np, Np);
        for (int j = 0; j < Np; ++j) for (int p = 0; p < Np; ++p) { const cd b = ri(p, j); for (int i = 0; i < Np; ++i) Sm(i, j) += lhs(i, p) * b; }
        for (int j = 0; j < Np; ++j) for (int i = 0; i < Np; ++i) {
            cd s(0, 0); for (int p = 0; p < Np; ++p) s += std::conj(Sm(p, i)) * Sm(p, j);
            unit_err_kme = std::max(unit_err_kme, std::abs(s - cd(i == j ? 1.0 : 0.0, 0.0)));
        }
    }
    // make_psi_boundary (KMatrixExtractor.cpp:200-224)
    Mat psiN(Np, Np);
    {
        const double rN = h + (Nr - 1) * h;
        for (int j = 0; j < Np; ++j) for (int i = 0; i < Np; ++i) {
            const int l = S.l_mu[i];
            psiN(i, j) = ((i == j) ? riccati_j(l, k * rN) : 0.0) + riccati_y(l, k * rN) * Kk(i, j);
        }
    }
    Mat dpsiN;
    if (tan) {
        const double rN = h + (Nr - 1) * h;
        dpsiN.resize(Np, Np);
        for (int j = 0; j < Np; ++j) for (int i = 0; i < Np; ++i) {
            const int l = S.l_mu[i];
            dpsiN(i, j) = ((i == j) ? riccati_jp(l, k * rN) * rN * dk : 0.0) + riccati_yp(l, k * rN) * rN * dk * Kk(i, j) +
                          riccati_y(l, k * rN) * dKk(i, j);
        }
    }
    const double t3 = now_s(); ts[3] = t3 - t2;
    // ---- energy-independent dipole tables (DipoleMatrixElement.cpp:584-965),
    // so the integrands can be formed point by point during the sweep ----
    const int n_main_hi = std::min(Nr - 1, std::max(1, S.n_tr - 1));   // main.cpp:1267-1270
    // initial state and bystander orbitals (main.cpp:746-822)
    const int Nlm_init = S.n_lambda;
    const int n_chi = S.n_tr;
    Mat chi_init(Nr, Nlm_init);
    std::vector<Mat> phi(n_occ, Mat(Nr, Nlm_init));
    for (int j = 0; j < n_occ; ++j)
        for (int ir = 0; ir < n_chi; ++ir)
            for (int lam = 0; lam < Nlm_init; ++lam) phi[j](ir, lam) = S.chi[ir](j, lam);
    if (S.P.has_init) {
        // /initial_state/psi_lm read raw (no r factor), main.cpp:781-802
        const int Nr_take = std::min(n_chi, Nr);
        for (int ir = 0; ir < Nr_take; ++ir)
            for (int lam = 0; lam < Nlm_init; ++lam) chi_init(ir, lam) = S.P.init_psi_lm[static_cast<size_t>(lam) * Nr + ir];
    } else {
        for (int ir = 0; ir < n_chi; ++ir)
            for (int lam = 0; lam < Nlm_init; ++lam) chi_init(ir, lam) = S.chi[ir](n_occ - 1, lam);
    }
    Mat dchi(Nr, Nlm_init);
    {
        std::vector<double> f(Nr), df(Nr);
        for (int nu = 0; nu < Nlm_init; ++nu) {
            for (int ir = 0; ir < Nr; ++ir) f[ir] = chi_init(ir, nu);
            five_point_derivative(f.data(), Nr, h, df.data());
            for (int ir = 0; ir < Nr; ++ir) dchi(ir, nu) = df[ir];
        }
    }
    std::vector<int> l_nu(Nlm_init), m_nu(Nlm_init);
    for (int nu = 0; nu < Nlm_init; ++nu) idx_to_lm(nu, l_nu[nu], m_nu[nu]);
    auto velocity_coef = [](int lm, int ln) -> double {    // DipoleMatrixElement.cpp:90-95
        if (lm == ln + 1) return -static_cast<double>(ln + 1);
        if (lm == ln - 1) return static_cast<double>(ln);
        return 0.0;
    };
    const int qs[3] = {+1, -1, 0};
    std::vector<Mat> Ang(3, Mat(Np, Nlm_init));
    for (int qi = 0; qi < 3; ++qi)
        for (int mu = 0; mu < Np; ++mu)
            for (int nu = 0; nu < Nlm_init; ++nu) {
                if (std::abs(S.l_mu[mu] - l_nu[nu]) != 1) continue;
                const double a = std::sqrt(4.0 * M_PI / 3.0) * gaunt_real(S.l_mu[mu], S.m_mu[mu], 1, qs[qi], l_nu[nu], m_nu[nu]);
                if (a != 0.0) Ang[qi](mu, nu) = a;
            }
    const int n_hi = std::min(Nr, n_main_hi + 1), n_lo = 0, n_pts = n_hi - n_lo;
    if (n_pts < 5) throw std::runtime_error("DipoleMatrixElement: too few radial points");
    std::vector<Mat> Xi(6, Mat(n_pts, Np));
    for (int slot = 0; slot < 6; ++slot) {
        const Mat& Ag = Ang[slot % 3];
        const bool L = slot < 3;
        for (int ir = n_lo; ir < n_hi; ++ir) {
            const double r = h + ir * h, inv_r = (r > 1e-30) ? 1.0 / r : 0.0;
            for (int mu = 0; mu < Np; ++mu) {
                double s = 0.0;
                for (int nu = 0; nu < Nlm_init; ++nu) {
                    const double a = Ag(mu, nu);
                    if (a == 0.0) continue;
                    const double c = chi_init(ir, nu);
                    if (L) s += a * c;
                    else s += a * (dchi(ir, nu) + velocity_coef(S.l_mu[mu], l_nu[nu]) * inv_r * c);
                }
                Xi[slot](ir - n_lo, mu) = s;
            }
        }
    }
    std::vector<Mat> integ(6, Mat(n_pts, Np));           // [slot](row, beta)
    std::vector<Mat> integ_b(n_occ, Mat(n_pts, Np));     // [alpha](row, beta)
    std::vector<Mat> dinteg(tan ? 6 : 0, Mat(n_pts, Np)), dinteg_b(tan ? n_occ : 0, Mat(n_pts, Np));
    auto integrands = [&](int ir, const Mat& ps, const Mat* dps) {
        if (ir < n_lo || ir >= n_hi) return;
        const int row = ir - n_lo;
        const double r = h + ir * h;
        for (int slot = 0; slot < 6; ++slot) {
            const double rf = (slot < 3) ? r : 1.0;
            for (int b = 0; b < Np; ++b) {
                double s = 0.0;
                for (int mu = 0; mu < Np; ++mu) s += Xi[slot](row, mu) * ps(mu, b);
                integ[slot](row, b) = rf * s;
            }
        }
        for (int a = 0; a < n_occ; ++a)
            for (int b = 0; b < Np; ++b) {
                double s = 0.0;
                for (int mu = 0; mu < Np; ++mu) s += phi[a](ir, mu) * ps(mu, b);
                integ_b[a](row, b) = s;
            }
        if (!dps) return;
        for (int slot = 0; slot < 6; ++slot) {
            const double rf = (slot < 3) ? r : 1.0;
            for (int b = 0; b < Np; ++b) {
                double s = 0.0;
                for (int mu = 0; mu < Np; ++mu) s += Xi[slot](row, mu) * (*dps)(mu, b);
                dinteg[slot](row, b) = rf * s;
            }
        }
        for (int a = 0; a < n_occ; ++a)
            for (int b = 0; b < Np; ++b) {
                double s = 0.0;
                for (int mu = 0; mu < Np; ++mu) s += phi[a](ir, mu) * (*dps)(mu, b);
                dinteg_b[a](row, b) = s;
            }
    };
    // ---- BackPropagator::run (BackPropagator.cpp:199-758) ----
    // The driver keeps psi on [0, n_tr-1] for the dipole (main.cpp:1267-1270),
    // whose integrands are formed here point by point, and on the 300-point
    // tail for the asymptotic fit (main.cpp:1335), which is stored.
    const int n_asym = 300, asym_off = N - n_asym + 1;
    std::vector<Mat> psi_asym(n_asym), dpsi_asym(tan ? n_asym : 0);
    auto store = [&](int n, const Mat& psi) {
        if (n >= asym_off && n <= N) psi_asym[n - asym_off] = psi;
    };
    auto dstore = [&](int n, const Mat& dpsi) {
        if (n >= asym_off && n <= N) dpsi_asym[n - asym_off] = dpsi;
    };
    store(N, psiN);
    if (tan) dstore(N, dpsiN);
    // Z at the outer boundary: Z_psi = psi + h^2/6 (E psi - V_N psi) + B_N f_N
    // with f_N = -D_N^{-1} B_N^T psi, Z_f = 0.  Its tangent is the same map on
    // dpsi_N plus h^2/6 w_N psi_N.
    Mat Z(Nt, Np), dZ;
    {
        const Mat BN = S.Bmat(N);
        const std::vector<double> DiN = S.Dinv(N);
        const Mat VN = S.potential(N);
        auto zmap = [&](const Mat& ps, Mat& Zo) {
            Mat fN = mul(transpose(BN), ps);
            for (int j = 0; j < Np; ++j) for (int f = 0; f < Nf; ++f) fN(f, j) = -(DiN[f] * fN(f, j));
            Mat Vpsi(Np, Np); gemm(h2_6, VN, ps, 0.0, Vpsi);
            Mat Bf = mul(BN, fN);
            Zo.resize(Nt, Np);
            for (int j = 0; j < Np; ++j) for (int i = 0; i < Np; ++i)
                Zo(i, j) = ((ps(i, j) + h2_6 * E * ps(i, j)) - Vpsi(i, j)) + Bf(i, j);
        };
        zmap(psiN, Z);
        if (tan) {
            zmap(dpsiN, dZ);
            const double c = h2_6 * wn(N);
            for (int j = 0; j < Np; ++j) for (int i = 0; i < Np; ++i) dZ(i, j) += c * psiN(i, j);
        }
    }
    integrands(N, psiN, tan ? &dpsiN : nullptr);
    {
        const int di = dump_at(out->n_dump_psi, out->dump_psi_n, N);
        put_at(out->dump_psi, static_cast<size_t>(Np) * Np, di, psiN);
        if (tan) put_at(out->dump_dpsi, static_cast<size_t>(Np) * Np, di, dpsiN);
    }
    Mat Y, Yd;
    std::vector<Mat> seg_s(L), seg_r(L), seg_dr(tan ? L : 0);
    for (int sg = n_seg - 1; sg >= 0; --sg) {
        const int n0 = sg * L, n1 = std::min(n0 + L, N);
        // recompute the segment's Sinv, Rinv (and dRinv) from its checkpoint
        for (int n = n0; n < n1; ++n) {
            int dA = 0, dS = 0;
            seg_s[n - n0] = S.sinv_at(n, E, dA, dS);
            if (n < ns) {
                seg_r[n - n0] = S.analytic_init(n, E);
                if (tan) seg_dr[n - n0] = S.analytic_init_tangent(n, E, tE, tA);
            } else {
                const Mat& prev = (n == n0) ? ck[sg] : seg_r[n - n0 - 1];
                seg_r[n - n0] = S.frp_step(n, seg_s[n - n0], prev);
                if (tan) {
                    const Mat* dprev = (n == n0) ? &dck[sg] : &seg_dr[n - n0 - 1];
                    seg_dr[n - n0] = frp_tan(n, seg_s[n - n0], seg_r[n - n0], n >= 1 ? dprev : nullptr);
                }
            }
        }
        for (int n = n1 - 1; n >= n0; --n) {
            const Mat& rn = seg_r[n - n0];
            const Mat& sn = seg_s[n - n0];
            const Mat Zold = tan ? Z : Mat();
            Z = mul(rn, Z);
            S.winv_apply(n, sn, Z, Y);
            Mat psi(Np, Np);
            for (int j = 0; j < Np; ++j) for (int i = 0; i < Np; ++i) psi(i, j) = Y(i, j);
            Mat dpsi;
            if (tan) {
                // dZ_n = dRinv_n Z_{n+1} + Rinv_n dZ_{n+1}.  For n >= n_start,
                // dRinv_n = Rinv_n H_n Rinv_n with H_n = dRinv_{n-1} + s_n G G^T,
                // so dRinv_n Z_{n+1} = Rinv_n H_n Z_n and G^T Z_n = psi_n:
                //   dZ_n = Rinv_n (dZ_{n+1} + dRinv_{n-1} Z_n + s_n G psi_n).
                // This grouping never multiplies by dRinv_n itself, which grows
                // like 1/delta^2 where Rinv_n passes near a pole (a near-singular
                // ratio); its rounding error then lies along the pole direction,
                // which Rinv_{n+1} maps to O(delta).
                Mat br = dZ;
                if (n >= ns) {
                    const Mat& dRm1 = (n == n0) ? dck[sg] : seg_dr[n - n0 - 1];
                    gemm(1.0, dRm1, Z, 1.0, br);
                    Mat G;
                    S.winv_apply(n, sn, Epsi, G);
                    gemm(12.0 * (sig[n] ? 1.0 : 0.0) * h2_6 * wn(n), G, psi, 1.0, br);
                    dZ = mul(rn, br);
                } else {
                    dZ = mul(rn, br);
                    gemm(1.0, seg_dr[n - n0], Zold, 1.0, dZ);
                }
                // dpsi = (W^{-1} dZ')_psi + (dW^{-1} Z')_psi, dW^{-1} = -sig h^2/6 w G G^T
                // and G^T Z' = psi, G_psi = Sinv
                S.winv_apply(n, sn, dZ, Yd);
                dpsi.resize(Np, Np);
                for (int j = 0; j < Np; ++j) for (int i = 0; i < Np; ++i) dpsi(i, j) = Yd(i, j);
                gemm(-(sig[n] ? 1.0 : 0.0) * h2_6 * wn(n), sn, psi, 1.0, dpsi);
            }
            integrands(n, psi, tan ? &dpsi : nullptr);
            const int di = dump_at(out->n_dump_psi, out->dump_psi_n, n);
            put_at(out->dump_psi, static_cast<size_t>(Np) * Np, di, psi);
            if (tan) put_at(out->dump_dpsi, static_cast<size_t>(Np) * Np, di, dpsi);
            if (tan) dstore(n, dpsi);
            store(n, std::move(psi));
        }
    }
    auto get_psi = [&](int n) -> const Mat& {
        if (n >= asym_off && n <= N) return psi_asym[n - asym_off];
        throw std::runtime_error("get_psi outside the kept tail");
    };
    auto get_dpsi = [&](int n) -> const Mat& {
        if (n >= asym_off && n <= N) return dpsi_asym[n - asym_off];
        throw std::runtime_error("get_dpsi outside the kept tail");
    };
    const double t4 = now_s(); ts[4] = t4 - t3;
    // ---- AsymptoticAmplitudes::extract (AsymptoticAmplitudes.cpp:64-359) ----
    int n_fs = std::max(Nr - 200, static_cast<int>(0.9 * Nr));
    int n_fe = Nr - 1 - 3;
    n_fs = std::max(n_fs, asym_off);
    n_fe = std::min(n_fe, asym_off + n_asym - 1);
    if (n_fe - n_fs + 1 < 10) throw std::runtime_error("fit window too small");
    const int n_fit = n_fe - n_fs + 1;
    std::vector<int> uls(S.l_mu.begin(), S.l_mu.end());
    std::sort(uls.begin(), uls.end()); uls.erase(std::unique(uls.begin(), uls.end()), uls.end());
    const int nl = static_cast<int>(uls.size());
    std::vector<std::vector<double>> jt(nl, std::vector<double>(n_fit)), yt(nl, std::vector<double>(n_fit));
    for (int t = 0; t < nl; ++t)
        for (int i = 0; i < n_fit; ++i) {
            const double r = h + (n_fs + i) * h;
            jt[t][i] = riccati_j(uls[t], k * r);
            yt[t][i] = riccati_y(uls[t], k * r);
        }
    std::vector<int> lidx(Np);
    for (int mu = 0; mu < Np; ++mu) lidx[mu] = static_cast<int>(std::lower_bound(uls.begin(), uls.end(), S.l_mu[mu]) - uls.begin());
    std::vector<std::array<double, 4>> Minv(nl);
    std::vector<char> closed(nl, 0);
    int n_closed = 0;
    for (int t = 0; t < nl; ++t) {
        double m00 = 0, m01 = 0, m11 = 0;
        for (int i = 0; i < n_fit; ++i) { m00 += jt[t][i] * jt[t][i]; m01 += jt[t][i] * yt[t][i]; m11 += yt[t][i] * yt[t][i]; }
        const double m10 = m01;
        const double det = m00 * m11 - m01 * m10;
        const double mx = std::max(std::max(std::abs(m00), std::abs(m01)), std::max(std::abs(m10), std::abs(m11)));
        const bool singular = !(std::abs(det) > 1e-20 * mx * mx);
        if (singular) {
            if (!S.P.allow_closed) throw std::runtime_error("AsymptoticAmplitudes: singular normal matrix (closed channel)");
            closed[t] = 1; ++n_closed;
            Minv[t] = {1, 0, 0, 1};
        } else {
            const double id = 1.0 / det;
            Minv[t] = {m11 * id, -m10 * id, -m01 * id, m00 * id};   // (00,10,01,11)
        }
    }
    double max_resid = 0.0, max_psi = 0.0;
    Mat YJ(Np, Np), YY(Np, Np);
    for (int ii = 0; ii < n_fit; ++ii) {
        const int i = (g_inverse_mode & 2) ? n_fit - 1 - ii : ii;
        const Mat& ps = get_psi(n_fs + i);
        max_psi = std::max(max_psi, max_abs(ps));
        for (int j = 0; j < Np; ++j) for (int mu = 0; mu < Np; ++mu) {
            YJ(mu, j) += jt[lidx[mu]][i] * ps(mu, j);
            YY(mu, j) += yt[lidx[mu]][i] * ps(mu, j);
        }
    }
    Mat Af(Np, Np), Bf(Np, Np);
    for (int mu = 0; mu < Np; ++mu) {
        const int t = lidx[mu];
        if (closed[t]) { Af(mu, mu) = 1.0; continue; }
        const auto& M = Minv[t];
        for (int j = 0; j < Np; ++j) {
            Af(mu, j) = M[0] * YJ(mu, j) + M[2] * YY(mu, j);
            Bf(mu, j) = M[1] * YJ(mu, j) + M[3] * YY(mu, j);
        }
    }
    for (int i = 0; i < n_fit; ++i) {
        const Mat& ps = get_psi(n_fs + i);
        for (int mu = 0; mu < Np; ++mu) {
            if (closed[lidx[mu]]) continue;
            const double jc = jt[lidx[mu]][i], yc = yt[lidx[mu]][i];
            for (int j = 0; j < Np; ++j) max_resid = std::max(max_resid, std::abs(jc * Af(mu, j) + yc * Bf(mu, j) - ps(mu, j)));
        }
    }
    // tangent of the fit: the normal matrices and the tables move with dk,
    // YJ / YY with dpsi and the tables
    Mat dAf, dBf;
    if (tan) {
        std::vector<std::vector<double>> djt(nl, std::vector<double>(n_fit)), dyt(nl, std::vector<double>(n_fit));
        for (int t = 0; t < nl; ++t)
            for (int i = 0; i < n_fit; ++i) {
                const double r = h + (n_fs + i) * h;
                djt[t][i] = riccati_jp(uls[t], k * r) * r * dk;
                dyt[t][i] = riccati_yp(uls[t], k * r) * r * dk;
            }
        std::vector<std::array<double, 4>> dMinv(nl);
        for (int t = 0; t < nl; ++t) {
            if (closed[t]) { dMinv[t] = {0, 0, 0, 0}; continue; }
            double d00 = 0, d01 = 0, d11 = 0;
            for (int i = 0; i < n_fit; ++i) {
                d00 += 2.0 * jt[t][i] * djt[t][i];
                d01 += djt[t][i] * yt[t][i] + jt[t][i] * dyt[t][i];
                d11 += 2.0 * yt[t][i] * dyt[t][i];
            }
            // dMinv = -Minv dM Minv (symmetric 2x2; storage (00,10,01,11))
            const auto& M = Minv[t];
            const double a = M[0], b = M[1], c = M[3];
            const double x00 = -(a * (d00 * a + d01 * b) + b * (d01 * a + d11 * b));
            const double x01 = -(a * (d00 * b + d01 * c) + b * (d01 * b + d11 * c));
            const double x11 = -(b * (d00 * b + d01 * c) + c * (d01 * b + d11 * c));
            dMinv[t] = {x00, x01, x01, x11};
        }
        Mat dYJ(Np, Np), dYY(Np, Np);
        for (int ii = 0; ii < n_fit; ++ii) {
            const int i = (g_inverse_mode & 2) ? n_fit - 1 - ii : ii;
            const Mat& ps = get_psi(n_fs + i);
            const Mat& dps = get_dpsi(n_fs + i);
            for (int j = 0; j < Np; ++j) for (int mu = 0; mu < Np; ++mu) {
                const int t = lidx[mu];
                dYJ(mu, j) += djt[t][i] * ps(mu, j) + jt[t][i] * dps(mu, j);
                dYY(mu, j) += dyt[t][i] * ps(mu, j) + yt[t][i] * dps(mu, j);
            }
        }
        dAf.resize(Np, Np); dBf.resize(Np, Np);
        for (int mu = 0; mu < Np; ++mu) {
            const int t = lidx[mu];
            if (closed[t]) continue;
            const auto& M = Minv[t];
            const auto& dM = dMinv[t];
            for (int j = 0; j < Np; ++j) {
                dAf(mu, j) = dM[0] * YJ(mu, j) + M[0] * dYJ(mu, j) + dM[2] * YY(mu, j) + M[2] * dYY(mu, j);
                dBf(mu, j) = dM[1] * YJ(mu, j) + M[1] * dYJ(mu, j) + dM[3] * YY(mu, j) + M[3] * dYY(mu, j);
            }
        }
    }
    double AmI = 0.0;
    for (int j = 0; j < Np; ++j) for (int i = 0; i < Np; ++i) AmI = std::max(AmI, std::abs(Af(i, j) - (i == j ? 1.0 : 0.0)));
    Mat Ainv;
    if (!inverse_fullpiv(Af, Ainv)) throw std::runtime_error("AsymptoticAmplitudes: fitted A matrix is singular.");
    const Mat Kfit = mul(Bf, Ainv);
    double Ksym = 0.0;
    for (int j = 0; j < Np; ++j) for (int i = 0; i < Np; ++i) Ksym = std::max(Ksym, std::abs(Kfit(i, j) - Kfit(j, i)));
    double S_unit = 0.0;
    {
        CMat apb(Np, Np), amb(Np, Np);
        for (size_t t = 0; t < Af.a.size(); ++t) { apb.a[t] = cd(Af.a[t], Bf.a[t]); amb.a[t] = cd(Af.a[t], -Bf.a[t]); }
        const CMat ai = inverse_pp(amb);
        CMat Sm(Np, Np);
        for (int j = 0; j < Np; ++j) for (int p = 0; p < Np; ++p) { const cd b = ai(p, j); for (int i = 0; i < Np; ++i) Sm(i, j) += apb(i, p) * b; }
        for (int j = 0; j < Np; ++j) for (int i = 0; i < Np; ++i) {
            cd s(0, 0); for (int p = 0; p < Np; ++p) s += std::conj(Sm(p, i)) * Sm(p, j);
            S_unit = std::max(S_unit, std::abs(s - cd(i == j ? 1.0 : 0.0, 0.0)));
        }
    }
    const double t5 = now_s(); ts[5] = t5 - t4;
    // ---- DipoleMatrixElement::compute_six (DipoleMatrixElement.cpp:584-965): Simpson sums ----
    Mat bov(Np, n_occ), dbov(tan ? Np : 0, tan ? n_occ : 0);
    for (int a = 0; a < n_occ; ++a) for (int b = 0; b < Np; ++b) bov(b, a) = simpson(integ_b[a].col(b), n_pts, h);
    if (tan)
        for (int a = 0; a < n_occ; ++a) for (int b = 0; b < Np; ++b) dbov(b, a) = simpson(dinteg_b[a].col(b), n_pts, h);
    CMat M(Np, Np);
    {
        CMat AmiB(Np, Np);
        for (size_t t = 0; t < Af.a.size(); ++t) AmiB.a[t] = cd(Af.a[t], -Bf.a[t]);
        const CMat inv = inverse_pp(AmiB);
        for (int j = 0; j < Np; ++j) for (int i = 0; i < Np; ++i) M(i, j) = std::conj(inv(j, i));
    }
    CMat o(Np, n_occ);
    for (int a = 0; a < n_occ; ++a) for (int p = 0; p < Np; ++p) { const cd b = bov(p, a); for (int i = 0; i < Np; ++i) o(i, a) += M(i, p) * b; }
    // tangent: M = (C^{-1})^dagger with C = A - iB, dM = -M dC^dagger M, so
    // d(M x) = M (dx - dC^dagger (M x)) for any E-dependent x
    auto mtan = [&](const std::vector<cd>& Mx, const std::vector<double>& dx) {
        std::vector<cd> v(Np);
        for (int i = 0; i < Np; ++i) {
            cd s(dx[i], 0.0);
            for (int p = 0; p < Np; ++p) s -= cd(dAf(p, i), dBf(p, i)) * Mx[p];   // (dC^dagger)(i,p) = conj(dC(p,i))
            v[i] = s;
        }
        std::vector<cd> r(Np, cd(0, 0));
        for (int p = 0; p < Np; ++p) for (int i = 0; i < Np; ++i) r[i] += M(i, p) * v[p];
        return r;
    };
    CMat dO(tan ? Np : 0, tan ? n_occ : 0);
    if (tan)
        for (int a = 0; a < n_occ; ++a) {
            std::vector<cd> oa(Np); std::vector<double> dba(Np);
            for (int i = 0; i < Np; ++i) { oa[i] = o(i, a); dba[i] = dbov(i, a); }
            const std::vector<cd> r = mtan(oa, dba);
            for (int i = 0; i < Np; ++i) dO(i, a) = r[i];
        }
    for (int slot = 0; slot < 6; ++slot) {
        std::vector<double> draw(Np);
        for (int b = 0; b < Np; ++b) draw[b] = simpson(integ[slot].col(b), n_pts, h);
        std::vector<cd> Draw(Np, cd(0, 0));
        for (int p = 0; p < Np; ++p) for (int i = 0; i < Np; ++i) Draw[i] += M(i, p) * draw[p];
        std::vector<double> dcorr(n_occ);
        for (int a = 0; a < n_occ; ++a) {
            std::vector<double> ic(n_pts);
            for (int ir = n_lo; ir < n_hi; ++ir) {
                const int row = ir - n_lo;
                const double r = h + ir * h, rf = (slot < 3) ? r : 1.0;
                double s = 0.0;
                for (int mu = 0; mu < Np; ++mu) s += phi[a](ir, mu) * Xi[slot](row, mu);
                ic[row] = rf * s;
            }
            dcorr[a] = simpson(ic.data(), n_pts, h);
        }
        std::vector<cd> D(Np);
        for (int mu = 0; mu < Np; ++mu) {
            cd sub = 0.0;
            for (int a = 0; a < n_occ; ++a) sub += 2.0 * o(mu, a) * dcorr[a];
            D[mu] = Draw[mu] - sub;
        }
        if (tan) {
            std::vector<double> ddraw(Np);
            for (int b = 0; b < Np; ++b) ddraw[b] = simpson(dinteg[slot].col(b), n_pts, h);
            const std::vector<cd> dDraw = mtan(Draw, ddraw);
            for (int mu = 0; mu < Np; ++mu) {
                cd sub = 0.0;
                for (int a = 0; a < n_occ; ++a) sub += 2.0 * dO(mu, a) * dcorr[a];
                const cd dDm = dDraw[mu] - sub;
                if (out->dD) { out->dD[2 * (slot * Np + mu)] = dDm.real(); out->dD[2 * (slot * Np + mu) + 1] = dDm.imag(); }
            }
            if (out->dd_raw) for (int b = 0; b < Np; ++b) out->dd_raw[slot * Np + b] = ddraw[b];
        }
        if (out->d_raw) for (int b = 0; b < Np; ++b) out->d_raw[slot * Np + b] = draw[b];
        if (out->D_raw) for (int b = 0; b < Np; ++b) { out->D_raw[2 * (slot * Np + b)] = Draw[b].real(); out->D_raw[2 * (slot * Np + b) + 1] = Draw[b].imag(); }
        if (out->D) for (int b = 0; b < Np; ++b) { out->D[2 * (slot * Np + b)] = D[b].real(); out->D[2 * (slot * Np + b) + 1] = D[b].imag(); }
        if (out->d_corr) for (int a = 0; a < n_occ; ++a) out->d_corr[slot * n_occ + a] = dcorr[a];
    }
    const double t6 = now_s(); ts[6] = t6 - t5; ts[7] = t6 - t0;
    // ---- outputs ----
    auto put = [](double* dst, const Mat& M) { if (dst) std::memcpy(dst, M.a.data(), M.a.size() * sizeof(double)); };
    put(out->K_kme, Kk); put(out->A, Af); put(out->B, Bf); put(out->K_fit, Kfit); put(out->b_overlap, bov);
    if (tan) { put(out->dK, dKk); put(out->dA, dAf); put(out->dB, dBf); put(out->db_overlap, dbov); }
    if (out->diag) {
        double* d = out->diag;
        d[0] = max_resid / std::max(max_psi, 1e-30); d[1] = max_resid; d[2] = Ksym; d[3] = AmI;
        d[4] = S.n_tr; d[5] = ns; d[6] = n_fs; d[7] = n_fe; d[8] = shA; d[9] = shS; d[10] = n_closed;
        d[11] = unit_err_kme; d[12] = S_unit;
    }
    if (out->t_stage) for (int i = 0; i < 8; ++i) out->t_stage[i] = ts[i];
}

}  // namespace orc

// ---------------------------------------------------------------------------
// C entry points.
// ---------------------------------------------------------------------------
extern "C" {

const char* sehf_oracle_last_error(void) { return orc::g_err.c_str(); }

void sehf_oracle_set_inverse_mode(int mode) { orc::g_inverse_mode = mode; }

void sehf_oracle_set_threads(int threads) { orc::g_threads = std::max(1, threads); }

int sehf_oracle_solve(const sehf_oracle_problem* p, double k, sehf_oracle_out* out) {
    try {
        const double t0 = orc::now_s();
        orc::Solver S(*p);
        const double t1 = orc::now_s();
        orc::Scan::run(S, k, out);
        if (out->t_stage) out->t_stage[0] = t1 - t0;
        return 0;
    } catch (const std::exception& e) {
        orc::g_err = e.what();
        return 1;
    }
}

int sehf_oracle_time_steps(const sehf_oracle_problem* p, double k, int n_steps, double* t) {
    try {
        orc::Solver S(*p);
        const double E = 0.5 * k * k;
        const int Nt = S.N_tot, Np = S.n_mu;
        int shA = 0, shS = 0;
        const int n0 = std::max(S.n_start(), 60);
        const int n1 = std::min(S.Nr - 1, n0 + n_steps);
        double ts = 0, tf = 0, tb = 0;
        orc::Mat Rprev = orc::identity(Nt);
        orc::Mat Z(Nt, Np);
        for (int i = 0; i < Nt; ++i) Z(i, i % Np) = 1.0;
        orc::Mat Y;
        for (int n = n0; n < n1; ++n) {
            const double a = orc::now_s();
            orc::Mat sinv = S.sinv_at(n, E, shA, shS);
            const double b = orc::now_s();
            orc::Mat r = S.frp_step(n, sinv, Rprev);
            const double c = orc::now_s();
            Z = orc::mul(r, Z);
            S.winv_apply(n, sinv, Z, Y);
            const double d = orc::now_s();
            ts += b - a; tf += c - b; tb += d - c;
            Rprev = r;
        }
        const int ns = std::max(1, n1 - n0);
        t[0] = ts / ns; t[1] = tf / ns; t[2] = tb / ns;
        return 0;
    } catch (const std::exception& e) {
        orc::g_err = e.what();
        return 1;
    }
}

int sehf_oracle_time_steps_multi(const sehf_oracle_problem* p, const double* ks, int nk, int n_steps,
                                 double* t, double* wall) {
    try {
        const double t0 = orc::now_s();
        orc::Solver S(*p);
        const double t1 = orc::now_s();
        const int Nt = S.N_tot, Np = S.n_mu;
        const int n0 = std::max(S.n_start(), 60);
        const int n1 = std::min(S.Nr - 1, n0 + n_steps);
        std::vector<std::string> errs(nk);
#pragma omp parallel for num_threads(nk) schedule(static, 1)
        for (int e = 0; e < nk; ++e) {
            try {
                const double E = 0.5 * ks[e] * ks[e];
                int shA = 0, shS = 0;
                double ts = 0, tf = 0, tb = 0;
                orc::Mat Rprev = orc::identity(Nt);
                orc::Mat Z(Nt, Np);
                for (int i = 0; i < Nt; ++i) Z(i, i % Np) = 1.0;
                orc::Mat Y;
                for (int n = n0; n < n1; ++n) {
                    const double a = orc::now_s();
                    orc::Mat sinv = S.sinv_at(n, E, shA, shS);
                    const double b = orc::now_s();
                    orc::Mat r = S.frp_step(n, sinv, Rprev);
                    const double c = orc::now_s();
                    Z = orc::mul(r, Z);
                    S.winv_apply(n, sinv, Z, Y);
                    const double d = orc::now_s();
                    ts += b - a; tf += c - b; tb += d - c;
                    Rprev = r;
                }
                const int ns = std::max(1, n1 - n0);
                t[3 * e + 0] = ts / ns; t[3 * e + 1] = tf / ns; t[3 * e + 2] = tb / ns;
            } catch (const std::exception& ex) {
                errs[e] = ex.what();
            }
        }
        for (auto& m : errs) if (!m.empty()) throw std::runtime_error(m);
        wall[0] = t1 - t0;              // setup (energy independent)
        wall[1] = orc::now_s() - t1;    // concurrent steps
        return 0;
    } catch (const std::exception& e) {
        orc::g_err = e.what();
        return 1;
    }
}

double sehf_oracle_w3j(int j1, int j2, int j3, int m1, int m2, int m3) { return orc::w3j(j1, j2, j3, m1, m2, m3); }
double sehf_oracle_gaunt_real(int l1, int m1, int l2, int m2, int l3, int m3) { return orc::gaunt_real(l1, m1, l2, m2, l3, m3); }
double sehf_oracle_real_ylm(int l, int m, double theta, double phi) { return orc::real_Ylm(l, m, theta, phi); }
double sehf_oracle_riccati_j(int l, double x) { return orc::riccati_j(l, x); }
double sehf_oracle_riccati_y(int l, double x) {
    try { return orc::riccati_y(l, x); } catch (...) { return NAN; }
}
int sehf_oracle_inverse_general(int n, const double* A, double* Ainv) {
    orc::Mat M(n, n);
    std::memcpy(M.a.data(), A, sizeof(double) * n * n);
    const orc::Mat X = orc::inverse_general(M);
    std::memcpy(Ainv, X.a.data(), sizeof(double) * n * n);
    return 0;
}
double sehf_oracle_min_eig_sym(int n, const double* A) {
    orc::Mat M(n, n);
    std::memcpy(M.a.data(), A, sizeof(double) * n * n);
    return orc::min_eig_sym(M);
}
void sehf_oracle_five_point_derivative(const double* f, int n, double dr, double* df) { orc::five_point_derivative(f, n, dr, df); }
double sehf_oracle_simpson(const double* f, int n_pts, double h) { return orc::simpson(f, n_pts, h); }
int sehf_oracle_potential(const sehf_oracle_problem* p, int ir, double* V) {
    try {
        orc::Solver S(*p);
        const orc::Mat M = S.potential(ir);
        std::memcpy(V, M.a.data(), sizeof(double) * M.a.size());
        return 0;
    } catch (const std::exception& e) { orc::g_err = e.what(); return 1; }
}
int sehf_oracle_exchange_Q(const sehf_oracle_problem* p, int ir, double* Q) {
    try {
        orc::Solver S(*p);
        const orc::Mat M = S.exchange_Q(ir);
        std::memcpy(Q, M.a.data(), sizeof(double) * M.a.size());
        return 0;
    } catch (const std::exception& e) { orc::g_err = e.what(); return 1; }
}
int sehf_oracle_frp_step(int n, const double* Winv, const double* Rprev, double* Rinv) {
    orc::Mat R(n, n);
    for (int j = 0; j < n; ++j) for (int i = 0; i < n; ++i)
        R(i, j) = (12.0 * Winv[static_cast<size_t>(j) * n + i] - 10.0 * (i == j ? 1.0 : 0.0)) - Rprev[static_cast<size_t>(j) * n + i];
    orc::Mat X = orc::inverse_general(R);
    orc::symmetrize(X);
    std::memcpy(Rinv, X.a.data(), sizeof(double) * n * n);
    return 0;
}

}  // extern "C"
