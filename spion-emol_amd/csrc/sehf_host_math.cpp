// sehf_host_math.cpp -- host numerics for input preparation (see header).
#include "sehf_host_math.h"

#include <quadmath.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <stdexcept>
#include <vector>

namespace sehf {

void index_lm(int idx, int& l, int& m) {
    l = static_cast<int>(std::sqrt(static_cast<double>(idx)));
    while ((l + 1) * (l + 1) <= idx) ++l;
    while (l > 0 && l * l > idx) --l;
    m = idx - l * l - l;
}

// ---------------------------------------------------------------------------
// Wigner 3j: (-1)^{j1-j2-m3} sqrt(Delta * prod (j+-m)!) sum_k (-1)^k / [...]
// ---------------------------------------------------------------------------
namespace {
template <class R>
struct LogFact {
    std::vector<R> v;
    explicit LogFact(R (*lg)(R)) : v(4096) {
        v[0] = 0;
        for (size_t i = 1; i < v.size(); ++i) v[i] = v[i - 1] + lg(static_cast<R>(i));
    }
};
long double logl_w(long double x) { return logl(x); }
__float128 logq_w(__float128 x) { return logq(x); }
const LogFact<long double>& lf_ld() { static const LogFact<long double> t(&logl_w); return t; }
const LogFact<__float128>& lf_q() { static const LogFact<__float128> t(&logq_w); return t; }

template <class R, class E>
R racah(int j1, int j2, int j3, int m1, int m2, int m3, const std::vector<R>& f, E ex, R& big) {
    const R pre = R(0.5) * (f[j1 + j2 - j3] + f[j1 - j2 + j3] + f[j2 + j3 - j1] - f[j1 + j2 + j3 + 1] +
                            f[j1 + m1] + f[j1 - m1] + f[j2 + m2] + f[j2 - m2] + f[j3 + m3] + f[j3 - m3]);
    const int k0 = std::max({0, j2 - j3 - m1, j1 - j3 + m2});
    const int k1 = std::min({j1 + j2 - j3, j1 - m1, j2 + m2});
    R sum = 0;
    big = 0;
    for (int k = k0; k <= k1; ++k) {
        const R term = ex(pre - (f[k] + f[j3 - j2 + k + m1] + f[j3 - j1 + k - m2] + f[j1 + j2 - j3 - k] +
                                 f[j1 - k - m1] + f[j2 - k + m2]));
        big = std::max(big, term);
        sum += (k & 1) ? -term : term;
    }
    return ((j1 - j2 - m3) & 1) ? -sum : sum;
}
}  // namespace

double wigner3j(int j1, int j2, int j3, int m1, int m2, int m3) {
    if (m1 + m2 + m3 != 0 || j3 < std::abs(j1 - j2) || j3 > j1 + j2) return 0.0;
    if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(m3) > j3) return 0.0;
    long double big = 0;
    const long double s = racah<long double>(j1, j2, j3, m1, m2, m3, lf_ld().v,
                                             [](long double x) { return expl(x); }, big);
    if (big <= 100.0L * fabsl(s)) return static_cast<double>(s);
    __float128 bq = 0;
    return static_cast<double>(racah<__float128>(j1, j2, j3, m1, m2, m3, lf_q().v,
                                                 [](__float128 x) { return expq(x); }, bq));
}

// ---------------------------------------------------------------------------
// Real Gaunt from the complex (no-conjugation) Gaunt integral
//   int Y_{l1m1} Y_{l2m2} Y_{l3m3} = sqrt((2l1+1)(2l2+1)(2l3+1)/4pi) (l1 l2 l3;000)(l1 l2 l3;m1 m2 m3)
// and the real<-complex leg coefficients of the convention
//   Y^R_{l,+m} = [(-1)^m Y_{l,m} + Y_{l,-m}]/sqrt2,  Y^R_{l,-m} = i[Y_{l,-m} - (-1)^m Y_{l,m}]/sqrt2.
// ---------------------------------------------------------------------------
namespace {
using cd = std::complex<double>;
// coefficient of complex Y_{l,mc} in real Y^R_{l,mr}
cd leg(int mr, int mc) {
    const double s = std::sqrt(0.5);
    const double ph = ((std::abs(mr) & 1) ? -1.0 : 1.0);
    if (mr == 0) return mc == 0 ? cd(1, 0) : cd(0, 0);
    if (mr > 0) {
        if (mc == mr) return cd(s * ph, 0);
        if (mc == -mr) return cd(s, 0);
        return cd(0, 0);
    }
    const int p = -mr;
    if (mc == -p) return cd(0, s);
    if (mc == p) return cd(0, -s * ph);
    return cd(0, 0);
}
}  // namespace

double real_gaunt(int l1, int m1, int l2, int m2, int l3, int m3) {
    if (l3 < std::abs(l1 - l2) || l3 > l1 + l2 || ((l1 + l2 + l3) & 1)) return 0.0;
    const double w000 = wigner3j(l1, l2, l3, 0, 0, 0);
    if (w000 == 0.0) return 0.0;
    const double pre = std::sqrt((2.0 * l1 + 1.0) * (2.0 * l2 + 1.0) * (2.0 * l3 + 1.0) / (4.0 * M_PI)) * w000;
    cd acc(0, 0);
    const int a1[2] = {-std::abs(m1), std::abs(m1)};
    const int a2[2] = {-std::abs(m2), std::abs(m2)};
    for (int x = 0; x < (m1 == 0 ? 1 : 2); ++x) {
        const int c1 = a1[x];
        const cd u1 = leg(m1, c1);
        for (int y = 0; y < (m2 == 0 ? 1 : 2); ++y) {
            const int c2 = a2[y];
            const int c3 = -(c1 + c2);
            if (std::abs(c3) != std::abs(m3)) continue;
            const cd u3 = leg(m3, c3);
            if (u3 == cd(0, 0)) continue;
            acc += u1 * leg(m2, c2) * u3 * wigner3j(l1, l2, l3, c1, c2, c3);
        }
    }
    return pre * acc.real();
}

GauntTable::GauntTable(int lmax) : lmax_(lmax) {
    nq_ = (3 * lmax) / 2 + 2;
    // Gauss-Legendre nodes by Newton iteration on P_n
    std::vector<double> x(nq_);
    w_.assign(nq_, 0.0);
    const int n = nq_;
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; ++j) { const double p3 = p2; p2 = p1; p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j; }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            const double z1 = z;
            z = z1 - p1 / pp;
            if (std::fabs(z - z1) < 1e-16) {
                double q1 = 1.0, q2 = 0.0;
                for (int j = 1; j <= n; ++j) { const double q3 = q2; q2 = q1; q1 = ((2.0 * j - 1.0) * z * q2 - (j - 1.0) * q3) / j; }
                pp = n * (z * q1 - q2) / (z * z - 1.0);
                break;
            }
        }
        x[i] = -z; x[n - 1 - i] = z;
        w_[i] = w_[n - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
    }
    // normalised associated Legendre without Condon-Shortley phase (the theta
    // part of Y^R, including the 1/sqrt(4 pi) of the m = 0 harmonic)
    S_.assign(static_cast<size_t>((lmax + 1) * (lmax + 2) / 2) * nq_, 0.0);
    for (int q = 0; q < nq_; ++q) {
        const double ct = x[q], st = std::sqrt(std::max(0.0, 1.0 - ct * ct));
        double pmm = 1.0 / std::sqrt(4.0 * M_PI);
        for (int m = 0; m <= lmax; ++m) {
            if (m > 0) pmm *= std::sqrt((2.0 * m + 1.0) / (2.0 * m)) * st;
            double p0 = 0.0, p1 = pmm;
            S_[(static_cast<size_t>(m * (m + 1) / 2 + m)) * nq_ + q] = pmm;
            for (int l = m + 1; l <= lmax; ++l) {
                double v;
                if (l == m + 1) v = std::sqrt(2.0 * m + 3.0) * ct * pmm;
                else {
                    const double a = std::sqrt((4.0 * l * l - 1.0) / (static_cast<double>(l) * l - m * m));
                    const double b = std::sqrt(((l - 1.0) * (l - 1.0) - m * m) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
                    v = a * (ct * p1 - b * p0);
                }
                p0 = p1; p1 = v;
                S_[(static_cast<size_t>(l * (l + 1) / 2 + m)) * nq_ + q] = v;
            }
        }
    }
}

double GauntTable::operator()(int l1, int m1, int l2, int m2, int l3, int m3) const {
    if (l3 < std::abs(l1 - l2) || l3 > l1 + l2 || ((l1 + l2 + l3) & 1)) return 0.0;
    if (std::abs(m1) > l1 || std::abs(m2) > l2 || std::abs(m3) > l3) return 0.0;
    // phi integral of f1 f2 f3, f_0 = 1, f_{m>0} = sqrt2 cos(m phi),
    // f_{m<0} = sqrt2 sin(|m| phi), via exponentials: sum over sign choices
    // with s1 a + s2 b + s3 c = 0 of the product of coefficients, times 2 pi.
    const int ms[3] = {m1, m2, m3};
    std::complex<double> acc(0.0, 0.0);
    for (int s1 = -1; s1 <= 1; s1 += 2)
        for (int s2 = -1; s2 <= 1; s2 += 2)
            for (int s3 = -1; s3 <= 1; s3 += 2) {
                const int sg[3] = {s1, s2, s3};
                int tot = 0;
                std::complex<double> c(1.0, 0.0);
                bool skip = false;
                for (int t = 0; t < 3; ++t) {
                    const int m = ms[t], a = std::abs(m);
                    if (m == 0) { if (sg[t] < 0) { skip = true; break; } continue; }
                    tot += sg[t] * a;
                    c *= (m > 0) ? std::complex<double>(0.5 * std::sqrt(2.0), 0.0)
                                 : std::complex<double>(0.0, -0.5 * std::sqrt(2.0) * sg[t]);
                }
                if (!skip && tot == 0) acc += c;
            }
    const double phi = 2.0 * M_PI * acc.real();
    if (phi == 0.0) return 0.0;
    const double* a = S_.data() + static_cast<size_t>(l1 * (l1 + 1) / 2 + std::abs(m1)) * nq_;
    const double* b = S_.data() + static_cast<size_t>(l2 * (l2 + 1) / 2 + std::abs(m2)) * nq_;
    const double* c = S_.data() + static_cast<size_t>(l3 * (l3 + 1) / 2 + std::abs(m3)) * nq_;
    double th = 0.0;
    for (int q = 0; q < nq_; ++q) th += w_[q] * a[q] * b[q] * c[q];
    return th * phi;
}

double real_ylm(int l, int m, double theta, double phi) {
    if (l < 0 || std::abs(m) > l) return 0.0;
    const int am = std::abs(m);
    const double st = std::sin(theta), ct = std::cos(theta);
    double pmm = 1.0 / std::sqrt(4.0 * M_PI);
    for (int k = 1; k <= am; ++k) pmm *= std::sqrt((2.0 * k + 1.0) / (2.0 * k)) * st;
    double val = pmm;
    if (l > am) {
        double p0 = pmm, p1 = std::sqrt(2.0 * am + 3.0) * ct * pmm;
        for (int k = am + 2; k <= l; ++k) {
            const double a = std::sqrt((4.0 * k * k - 1.0) / (static_cast<double>(k) * k - am * am));
            const double b = std::sqrt(((k - 1.0) * (k - 1.0) - am * am) / (4.0 * (k - 1.0) * (k - 1.0) - 1.0));
            const double p2 = a * (ct * p1 - b * p0);
            p0 = p1; p1 = p2;
        }
        val = p1;
    }
    if (m == 0) return val;
    return (m > 0) ? std::sqrt(2.0) * val * std::cos(am * phi) : std::sqrt(2.0) * val * std::sin(am * phi);
}

// ---------------------------------------------------------------------------
// Spherical Bessel functions.
// j_l: series for small x, upward recurrence for x > l, otherwise the
// continued fraction for j_l/j_{l-1} (modified Lentz) + downward recurrence,
// normalised to j_0 or j_1.  y_l: upward recurrence (stable).
// ---------------------------------------------------------------------------
static double sph_j_series(int l, double x) {
    double pre = 1.0;
    for (int i = 1; i <= l; ++i) pre *= x / (2.0 * i + 1.0);
    double term = 1.0, sum = 1.0;
    const double q = -0.5 * x * x;
    for (int k = 1; k < 300; ++k) {
        term *= q / (k * (2.0 * l + 2.0 * k + 1.0));
        sum += term;
        if (std::fabs(term) <= 1e-18 * std::fabs(sum)) break;
    }
    return pre * sum;
}

static double sph_j(int l, double x) {
    if (x == 0.0) return l == 0 ? 1.0 : 0.0;
    if (x < 1e-3 || x * x < 0.1 * (2.0 * l + 3.0)) return sph_j_series(l, x);
    const double s = std::sin(x), c = std::cos(x);
    const double j0 = s / x, j1 = s / (x * x) - c / x;
    if (l == 0) return j0;
    if (l == 1) return j1;
    if (x > l) {
        double a = j0, b = j1;
        for (int n = 1; n < l; ++n) { const double t = (2.0 * n + 1.0) / x * b - a; a = b; b = t; }
        return b;
    }
    // ratio r = j_l / j_{l-1} = 1 / ((2l+1)/x - 1/((2l+3)/x - ...))  (modified Lentz)
    const double tiny = 1e-300;
    double f = (2.0 * l + 1.0) / x, C = f, D = 0.0;
    for (int i = 1; i < 100000; ++i) {
        const double bi = (2.0 * (l + i) + 1.0) / x;
        D = bi - D;  D = (std::fabs(D) < tiny) ? tiny : D; D = 1.0 / D;
        C = bi - 1.0 / C; C = (std::fabs(C) < tiny) ? tiny : C;
        // continued fraction with a_i = -1: f = b0 - 1/(b1 - 1/(b2 - ...))
        const double delta = C * D;
        f *= delta;
        if (std::fabs(delta - 1.0) < 1e-16) break;
    }
    // downward recurrence from (f_l, f_{l-1}) = (1, f) with f = j_{l-1}/j_l
    double fl = 1.0, flm1 = f;
    double top = fl;          // value of j_l in these units
    double a = fl, b = flm1;  // a = f_{n}, b = f_{n-1}
    for (int n = l - 1; n >= 1; --n) {
        const double t = (2.0 * n + 1.0) / x * b - a;   // f_{n-1}
        a = b; b = t;
        if (std::fabs(b) > 1e250) { a *= 1e-250; b *= 1e-250; top *= 1e-250; }
    }
    // now a = f_1, b = f_0
    const double scale = (std::fabs(j0) >= std::fabs(j1)) ? j0 / b : j1 / a;
    return top * scale;
}

static double sph_y(int l, double x) {
    const double s = std::sin(x), c = std::cos(x);
    double a = -c / x;
    if (l == 0) return a;
    double b = -c / (x * x) - s / x;
    for (int n = 1; n < l; ++n) { const double t = (2.0 * n + 1.0) / x * b - a; a = b; b = t; }
    return b;
}

double ric_j(int l, double x) { return (x < 1e-300) ? 0.0 : x * sph_j(l, x); }
double ric_y(int l, double x) {
    if (x < 1e-300) throw std::runtime_error("Riccati-Neumann evaluated at x ~ 0");
    return x * sph_y(l, x);
}
double ric_jp(int l, double x) {
    if (l == 0) return std::cos(x);
    if (x < 1e-300) return 0.0;
    return ric_j(l - 1, x) - static_cast<double>(l) / x * ric_j(l, x);
}
double ric_yp(int l, double x) {
    if (l == 0) return std::sin(x);
    return ric_y(l - 1, x) - static_cast<double>(l) / x * ric_y(l, x);
}

// ---------------------------------------------------------------------------
// Smallest eigenvalue: Householder reduction to tridiagonal form, then
// bisection on the Sturm count.
// ---------------------------------------------------------------------------
double sym_min_eig(int n, const double* a_in) {
    if (n == 1) return a_in[0];
    std::vector<double> a(static_cast<size_t>(n) * n);
    for (int j = 0; j < n; ++j)
        for (int i = j; i < n; ++i) a[static_cast<size_t>(j) * n + i] = a[static_cast<size_t>(i) * n + j] = a_in[static_cast<size_t>(j) * n + i];
    auto A = [&](int i, int j) -> double& { return a[static_cast<size_t>(j) * n + i]; };
    std::vector<double> d(n), e(n, 0.0), v(n), p(n);
    for (int k = 0; k + 2 < n; ++k) {
        double nrm2 = 0.0;
        for (int i = k + 1; i < n; ++i) nrm2 += A(i, k) * A(i, k);
        if (nrm2 == 0.0) { e[k] = 0.0; continue; }
        const double nrm = std::sqrt(nrm2);
        const double alpha = (A(k + 1, k) > 0.0) ? -nrm : nrm;
        for (int i = k + 1; i < n; ++i) v[i] = A(i, k);
        v[k + 1] -= alpha;
        double vv = 0.0;
        for (int i = k + 1; i < n; ++i) vv += v[i] * v[i];
        e[k] = alpha;
        if (vv == 0.0) continue;
        const double beta = 2.0 / vv;
#pragma omp parallel for schedule(static) if (n > 256)
        for (int i = k + 1; i < n; ++i) {
            double s = 0.0;
            for (int j = k + 1; j < n; ++j) s += A(j, i) * v[j];
            p[i] = beta * s;
        }
        double K = 0.0;
        for (int i = k + 1; i < n; ++i) K += v[i] * p[i];
        K *= 0.5 * beta;
        for (int i = k + 1; i < n; ++i) p[i] -= K * v[i];
#pragma omp parallel for schedule(static) if (n > 256)
        for (int j = k + 1; j < n; ++j)
            for (int i = k + 1; i < n; ++i) A(i, j) -= v[i] * p[j] + p[i] * v[j];
    }
    for (int i = 0; i < n; ++i) d[i] = A(i, i);
    e[n - 2] = A(n - 1, n - 2);
    double lo = 1e300, hi = -1e300;
    for (int i = 0; i < n; ++i) {
        const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(e[i]) : 0.0);
        lo = std::min(lo, d[i] - r);
        hi = std::max(hi, d[i] + r);
    }
    auto below = [&](double x) {
        int c = 0;
        double q = d[0] - x;
        if (q < 0) ++c;
        for (int i = 1; i < n; ++i) {
            if (q == 0.0) q = 1e-300;
            q = d[i] - x - e[i - 1] * e[i - 1] / q;
            if (q < 0) ++c;
        }
        return c;
    };
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (below(mid) >= 1) hi = mid; else lo = mid;
    }
    return 0.5 * (lo + hi);
}

std::vector<double> simpson_weights(int n, double h) {
    std::vector<double> w(std::max(n, 0), 0.0);
    if (n < 2) return w;
    if (n == 2) { w[0] = w[1] = 0.5 * h; return w; }
    auto third = [&](int m) {   // Simpson 1/3 on samples [0, m), m odd
        w[0] += h / 3.0; w[m - 1] += h / 3.0;
        for (int i = 1; i < m - 1; i += 2) w[i] += 4.0 * h / 3.0;
        for (int i = 2; i < m - 2; i += 2) w[i] += 2.0 * h / 3.0;
    };
    if (n % 2 == 1) { third(n); return w; }
    if (n < 4) {   // trapezoid
        w[0] = w[n - 1] = 0.5 * h;
        for (int i = 1; i < n - 1; ++i) w[i] = h;
        return w;
    }
    third(n - 3);
    const double c = 3.0 * h / 8.0;
    w[n - 4] += c; w[n - 3] += 3.0 * c; w[n - 2] += 3.0 * c; w[n - 1] += c;
    return w;
}

void deriv5(const double* f, int n, double h, double* df) {
    if (n < 5) {
        if (n >= 2) {
            df[0] = (f[1] - f[0]) / h;
            for (int i = 1; i < n - 1; ++i) df[i] = (f[i + 1] - f[i - 1]) / (2.0 * h);
            df[n - 1] = (f[n - 1] - f[n - 2]) / h;
        } else if (n == 1) {
            df[0] = 0.0;
        }
        return;
    }
    const double s = 1.0 / (12.0 * h);
    for (int i = 2; i < n - 2; ++i) df[i] = (f[i - 2] - 8.0 * f[i - 1] + 8.0 * f[i + 1] - f[i + 2]) * s;
    df[0] = (-25.0 * f[0] + 48.0 * f[1] - 36.0 * f[2] + 16.0 * f[3] - 3.0 * f[4]) * s;
    df[1] = (-3.0 * f[0] - 10.0 * f[1] + 18.0 * f[2] - 6.0 * f[3] + f[4]) * s;
    df[n - 2] = (-f[n - 5] + 6.0 * f[n - 4] - 18.0 * f[n - 3] + 10.0 * f[n - 2] + 3.0 * f[n - 1]) * s;
    df[n - 1] = (3.0 * f[n - 5] - 16.0 * f[n - 4] + 36.0 * f[n - 3] - 48.0 * f[n - 2] + 25.0 * f[n - 1]) * s;
}

}  // namespace sehf
