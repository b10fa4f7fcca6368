// sehf_setup.cpp -- energy-independent inputs (see sehf_setup.h).
#include "sehf_setup.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

#include "sehf_host_math.h"

namespace sehf {

Setup::Setup(const sehf_problem& p) : P(p) {
    l_cont = p.l_cont;
    if (l_cont < 0) throw std::runtime_error("sehf_problem: l_cont < 0");
    l_exch = std::min(l_cont, 10);
    l_orb = l_cont + l_exch;
    Np = (l_cont + 1) * (l_cont + 1);
    n_sigma = (l_exch + 1) * (l_exch + 1);
    n_lambda = (l_orb + 1) * (l_orb + 1);
    l_exp = 2 * l_cont;
    n_exp = (l_exp + 1) * (l_exp + 1);
    n_occ = p.n_occ;
    Nr = p.Nr;
    dr = p.dr;
    Nf = n_occ * n_sigma;
    Nt = Np + Nf;
    if (n_occ <= 0 || n_occ > p.n_orb) throw std::runtime_error("sehf_problem: bad n_occ");
    if (Nr < 310) throw std::runtime_error("sehf_problem: Nr must be >= 310 (300-point asymptotic window + fit)");
    if (!(dr > 0.0)) throw std::runtime_error("sehf_problem: dr must be > 0");
    if (p.nlm_vh < n_exp)
        throw std::runtime_error("sehf_problem: V_H has " + std::to_string(p.nlm_vh) +
                                 " rows, (2 l_cont + 1)^2 = " + std::to_string(n_exp) + " needed");
    if (p.nlm_orb < n_lambda)
        throw std::runtime_error("sehf_problem: orbitals store " + std::to_string(p.nlm_orb) +
                                 " channels, (l_cont + l_exch + 1)^2 = " + std::to_string(n_lambda) + " needed");
    l_mu.resize(Np); m_mu.resize(Np);
    for (int mu = 0; mu < Np; ++mu) index_lm(mu, l_mu[mu], m_mu[mu]);
    l_sig.resize(n_sigma);
    for (int s = 0; s < n_sigma; ++s) { int m; index_lm(s, l_sig[s], m); }
    // ForwardRPropagator::compute_n_start_ (ForwardRPropagator.cpp:49-62)
    {
        const int lf = l_exch;
        const double r_crit = dr * std::sqrt(static_cast<double>(lf * (lf + 1)) / 12.0);
        n_start = std::max(1, static_cast<int>(std::ceil(2.0 * r_crit / dr)));
        n_start = std::min(n_start, Nr - 1);
    }
    find_n_transition();
    build_potential_gaunt();
    build_exchange_gaunt();
}

// WavefunctionSetup.cpp:260-320: last radius where any occupied SCE
// coefficient exceeds 1e-15, plus a max(100, 5/dr) point buffer.
void Setup::find_n_transition() {
    const int nl = std::min(n_lambda, P.nlm_orb);
    int last = 0;
    for (int ir = Nr - 1; ir >= 0; --ir) {
        double mx = 0.0;
        for (int j = 0; j < n_occ && mx <= 1e-15; ++j)
            for (int lam = 0; lam < nl; ++lam)
                mx = std::max(mx, std::fabs(P.psi_lm[(static_cast<size_t>(j) * P.nlm_orb + lam) * Nr + ir]));
        if (mx > 1e-15) { last = ir; break; }
    }
    const int buffer = std::max(100, static_cast<int>(5.0 / dr));
    n_tr = std::max(1, std::min(Nr, last + 1 + buffer));
}

double Setup::chi(int j, int lam, int ir) const {
    return (ir * dr) * P.psi_lm[(static_cast<size_t>(j) * P.nlm_orb + lam) * Nr + ir];
}

// Non-zero G(mu', sigma, mu) on the upper triangle mu >= mu', |G| > 1e-15,
// enumerated through the real-harmonic selection rules
// |m_s| in {|m1| + |m2|, ||m1| - |m2||}.
void Setup::build_potential_gaunt() {
    const int64_t npairs = static_cast<int64_t>(Np) * (Np + 1) / 2;
    pair_i_.resize(npairs); pair_j_.resize(npairs);
    {
        int64_t q = 0;
        for (int a = 0; a < Np; ++a)
            for (int b = a; b < Np; ++b) { pair_i_[q] = a; pair_j_[q] = b; ++q; }
    }
    std::vector<std::vector<std::pair<int, double>>> rows(npairs);
    const GauntTable gt(l_exp);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t q = 0; q < npairs; ++q) {
        const int a = pair_i_[q], b = pair_j_[q];
        const int l2 = l_mu[a], m2 = m_mu[a], l1 = l_mu[b], m1 = m_mu[b];
        auto& row = rows[q];
        for (int ls = std::abs(l1 - l2); ls <= std::min(l1 + l2, l_exp); ls += 2) {
            const int cand[2] = {std::abs(m1) + std::abs(m2), std::abs(std::abs(m1) - std::abs(m2))};
            for (int c = 0; c < 2; ++c) {
                if (c == 1 && cand[1] == cand[0]) continue;
                const int am = cand[c];
                if (am > ls) continue;
                for (int sg = (am == 0 ? 1 : -1); sg <= 1; sg += 2) {
                    const int ms = sg * am;
                    const double G = gt(l2, m2, ls, ms, l1, m1);
                    if (std::fabs(G) > 1e-15) row.push_back({lm_index(ls, ms), G});
                }
            }
        }
        std::sort(row.begin(), row.end());
    }
    vptr_.assign(npairs + 1, 0);
    for (int64_t q = 0; q < npairs; ++q) vptr_[q + 1] = vptr_[q] + static_cast<int64_t>(rows[q].size());
    vsig_.resize(vptr_[npairs]); vval_.resize(vptr_[npairs]);
    for (int64_t q = 0; q < npairs; ++q)
        for (size_t t = 0; t < rows[q].size(); ++t) {
            vsig_[vptr_[q] + t] = rows[q][t].first;
            vval_[vptr_[q] + t] = rows[q][t].second;
        }
}

// Non-zero G(mu, lambda, sigma), |G| > 1e-14 (WavefunctionSetup.cpp:21-81),
// stacked as rows mu*n_sigma + sigma with ascending lambda.
void Setup::build_exchange_gaunt() {
    const int nrow = Np * n_sigma;
    std::vector<std::vector<std::pair<int, double>>> rows(nrow);
    const GauntTable gt(l_orb);
#pragma omp parallel for schedule(dynamic, 16)
    for (int rr = 0; rr < nrow; ++rr) {
        const int mu = rr / n_sigma, s = rr % n_sigma;
        const int lm = l_mu[mu], mm = m_mu[mu];
        int ls, ms; index_lm(s, ls, ms);
        auto& row = rows[rr];
        for (int ll = std::abs(lm - ls); ll <= std::min(lm + ls, l_orb); ll += 2) {
            const int cand[2] = {std::abs(mm) + std::abs(ms), std::abs(std::abs(mm) - std::abs(ms))};
            for (int c = 0; c < 2; ++c) {
                if (c == 1 && cand[1] == cand[0]) continue;
                const int am = cand[c];
                if (am > ll) continue;
                for (int sg = (am == 0 ? 1 : -1); sg <= 1; sg += 2) {
                    const int ml = sg * am;
                    const double G = gt(lm, mm, ll, ml, ls, ms);
                    if (std::fabs(G) > 1e-14) row.push_back({lm_index(ll, ml), G});
                }
            }
        }
        std::sort(row.begin(), row.end());
    }
    xptr_.assign(nrow + 1, 0);
    for (int rr = 0; rr < nrow; ++rr) xptr_[rr + 1] = xptr_[rr] + static_cast<int64_t>(rows[rr].size());
    xlam_.resize(xptr_[nrow]); xval_.resize(xptr_[nrow]);
    for (int rr = 0; rr < nrow; ++rr)
        for (size_t t = 0; t < rows[rr].size(); ++t) {
            xlam_[xptr_[rr] + t] = rows[rr][t].first;
            xval_[xptr_[rr] + t] = rows[rr][t].second;
        }
}

// V_s(r) = V_H,s(r) - sum_atoms Z 4pi/(2l+1) r<^l / r>^(l+1) Y^R_s(R_hat)
// (Potentials.cpp:40-88).
void Setup::v_sigma(double r, int ir, std::vector<double>& out) const {
    out.assign(n_exp, 0.0);
    const int nmax = std::min(n_exp, P.nlm_vh);
    for (int s = 0; s < nmax; ++s) out[s] += P.V_H[static_cast<size_t>(s) * Nr + ir];
    for (int a = 0; a < P.n_atoms; ++a) {
        const double Z = P.Z[a];
        const double x = P.xyz[3 * a], y = P.xyz[3 * a + 1], z = P.xyz[3 * a + 2];
        const double R = std::sqrt(x * x + y * y + z * z);
        double th = 0.0, ph = 0.0;
        if (R > 1e-14) { th = std::acos(std::clamp(z / R, -1.0, 1.0)); ph = std::atan2(y, x); }
        const double rl = std::min(r, R), rg = std::max(r, R);
        for (int l = 0; l <= l_exp; ++l) {
            const double radial = (4.0 * M_PI / (2.0 * l + 1.0)) * std::pow(rl, l) / std::pow(rg, l + 1);
            for (int m = -l; m <= l; ++m) {
                const double Y = (R < 1e-14) ? ((l == 0) ? 1.0 / (2.0 * std::sqrt(M_PI)) : 0.0) : real_ylm(l, m, th, ph);
                if (std::fabs(Y) > 1e-15) out[lm_index(l, m)] += -Z * radial * Y;
            }
        }
    }
}

double Setup::polarization(double r) const {
    if (P.pol_kind == 1) {
        if (P.pol_alpha == 0.0 || r < 1.0e-6) return 0.0;
        const double u = r / 1.5, u6 = u * u * u * u * u * u;
        const double d = 1.0 - std::exp(-u6);
        return -0.5 * P.pol_alpha * d * d / (r * r * r * r);
    }
    if (P.pol_kind == 2) {
        if (r < 1.0e-6) return 0.0;
        const double q = r * r + P.pol_rc * P.pol_rc;
        return -P.pol_alpha / (2.0 * q * q);
    }
    return 0.0;
}

void Setup::vsigma_table(int n0, int count, double* out) const {
    // v_sigma(r) for many r: the spherical harmonics of the nuclei do not
    // depend on r, so they are evaluated once; every element is formed with
    // the same operations in the same order as v_sigma.
    const int nmax = std::min(n_exp, P.nlm_vh);
    std::vector<double> Yat(static_cast<size_t>(P.n_atoms) * n_exp, 0.0), Rat(P.n_atoms);
    for (int a = 0; a < P.n_atoms; ++a) {
        const double x = P.xyz[3 * a], y = P.xyz[3 * a + 1], z = P.xyz[3 * a + 2];
        const double R = std::sqrt(x * x + y * y + z * z);
        Rat[a] = R;
        double th = 0.0, ph = 0.0;
        if (R > 1e-14) { th = std::acos(std::clamp(z / R, -1.0, 1.0)); ph = std::atan2(y, x); }
        for (int l = 0; l <= l_exp; ++l)
            for (int m = -l; m <= l; ++m)
                Yat[static_cast<size_t>(a) * n_exp + lm_index(l, m)] =
                    (R < 1e-14) ? ((l == 0) ? 1.0 / (2.0 * std::sqrt(M_PI)) : 0.0) : real_ylm(l, m, th, ph);
    }
#pragma omp parallel for schedule(static)
    for (int q = 0; q < count; ++q) {
        const int ir = n0 + q;
        const double r = ir * dr;
        double* o = out + static_cast<size_t>(q) * n_exp;
        for (int s = 0; s < n_exp; ++s) o[s] = 0.0;
        for (int s = 0; s < nmax; ++s) o[s] += P.V_H[static_cast<size_t>(s) * Nr + ir];
        for (int a = 0; a < P.n_atoms; ++a) {
            const double Z = P.Z[a], R = Rat[a];
            const double rl = std::min(r, R), rg = std::max(r, R);
            for (int l = 0; l <= l_exp; ++l) {
                const double radial = (4.0 * M_PI / (2.0 * l + 1.0)) * std::pow(rl, l) / std::pow(rg, l + 1);
                for (int m = -l; m <= l; ++m) {
                    const double Y = Yat[static_cast<size_t>(a) * n_exp + lm_index(l, m)];
                    if (std::fabs(Y) > 1e-15) o[lm_index(l, m)] += -Z * radial * Y;
                }
            }
        }
    }
}

void Setup::chi_table(int n0, int count, double* out) const {
#pragma omp parallel for schedule(static)
    for (int q = 0; q < count; ++q)
        for (int j = 0; j < n_occ; ++j)
            for (int lam = 0; lam < n_lambda; ++lam)
                out[(static_cast<size_t>(q) * n_occ + j) * n_lambda + lam] = chi(j, lam, n0 + q);
}

double Setup::polarization_per_alpha(int ir) const {
    const double r = ir * dr;
    if (!(r > 1e-14) || r < 1.0e-6) return 0.0;   // V(ir) is zero at r = 0
    if (P.pol_kind == 1) {
        const double u = r / 1.5, u6 = u * u * u * u * u * u;
        const double d = 1.0 - std::exp(-u6);
        return -0.5 * d * d / (r * r * r * r);
    }
    if (P.pol_kind == 2) {
        const double q = r * r + P.pol_rc * P.pol_rc;
        return -1.0 / (2.0 * q * q);
    }
    return 0.0;
}

void Setup::potential(int ir, double* V) const {
    std::fill(V, V + static_cast<size_t>(Np) * Np, 0.0);
    const double r = ir * dr;
    if (!(r > 1e-14)) return;
    std::vector<double> vs;
    v_sigma(r, ir, vs);
    const int64_t npairs = static_cast<int64_t>(pair_i_.size());
    for (int64_t q = 0; q < npairs; ++q) {
        double acc = 0.0;
        for (int64_t t = vptr_[q]; t < vptr_[q + 1]; ++t) acc += vval_[t] * vs[vsig_[t]];
        const int a = pair_i_[q], b = pair_j_[q];
        V[a + static_cast<size_t>(b) * Np] = acc;
        V[b + static_cast<size_t>(a) * Np] = acc;
    }
    const double vp = polarization(r), inv_r2 = 1.0 / (r * r);
    for (int mu = 0; mu < Np; ++mu) {
        double& d = V[mu + static_cast<size_t>(mu) * Np];
        if (vp != 0.0) d += vp;
        d += l_mu[mu] * (l_mu[mu] + 1.0) * 0.5 * inv_r2;
    }
}

void Setup::build_chunk(int n0, int count, double* V, double* B) const {
    const size_t pp = static_cast<size_t>(Np) * Np, pf = static_cast<size_t>(Np) * Nf;
#pragma omp parallel for schedule(dynamic, 1)
    for (int q = 0; q < count; ++q) {
        potential(n0 + q, V + q * pp);
        if (Nf > 0) exchange_B(n0 + q, B + q * pf);
    }
}

void Setup::exchange_B(int ir, double* B) const {
    std::fill(B, B + static_cast<size_t>(Np) * Nf, 0.0);
    const double r = ir * dr;
    if (r < 1e-10 || ir >= n_tr) return;
    const double factor = 2.0 * std::sqrt(2.0 * M_PI) / r;
    const double h2_12 = dr * dr / 12.0;
    std::vector<double> ch(static_cast<size_t>(n_occ) * n_lambda);
    for (int j = 0; j < n_occ; ++j)
        for (int lam = 0; lam < n_lambda; ++lam) ch[static_cast<size_t>(j) * n_lambda + lam] = chi(j, lam, ir);
    for (int mu = 0; mu < Np; ++mu)
        for (int s = 0; s < n_sigma; ++s) {
            const int rr = mu * n_sigma + s;
            for (int j = 0; j < n_occ; ++j) {
                double acc = 0.0;
                const double* c = ch.data() + static_cast<size_t>(j) * n_lambda;
                for (int64_t t = xptr_[rr]; t < xptr_[rr + 1]; ++t) acc += xval_[t] * c[xlam_[t]];
                B[mu + static_cast<size_t>(j * n_sigma + s) * Np] = h2_12 * (factor * acc);
            }
        }
}

void Setup::dinv(int n, double* d) const {
    const double r = dr + n * dr, r2 = r * r, h2_12 = dr * dr / 12.0;
    for (int f = 0; f < Nf; ++f) {
        const int l = l_sig[f % n_sigma];
        const double centrif = (r2 > 1e-30) ? static_cast<double>(l) * (l + 1) / r2 : 0.0;
        double Df = 1.0 - h2_12 * centrif;
        if (Df < W_min) Df = W_min;
        d[f] = 1.0 / Df;
    }
}

// DipoleMatrixElement::compute_six inputs: Ang_q (q = +1, -1, 0 for x, y, z),
// Xi for length / velocity gauge with r = (ir+1) dr, the bystander orbitals
// phi_a = chi_a (r = ir dr), Simpson weights over [0, n_pts),
// n_pts = min(Nr, n_transition).  Initial state: HOMO chi, or the raw
// /initial_state/psi_lm columns (main.cpp:746-822).
void Setup::dipole_tables(std::vector<double>& rows, std::vector<double>& dcorr, int& n_pts_out) const {
    const int n_main_hi = std::min(Nr - 1, std::max(1, n_tr - 1));   // main.cpp:1267-1270
    const int n_pts = std::min(Nr, n_main_hi + 1);
    n_pts_out = n_pts;
    const int nslot = 6 + n_occ;
    const int nl = n_lambda;
    std::vector<double> ci(static_cast<size_t>(Nr) * nl, 0.0), dci(static_cast<size_t>(Nr) * nl, 0.0);
    const int n_chi = n_tr;
    for (int lam = 0; lam < nl; ++lam) {
        double* c = ci.data() + static_cast<size_t>(lam) * Nr;
        if (P.has_init) {
            for (int ir = 0; ir < std::min(n_chi, Nr); ++ir) c[ir] = P.init_psi_lm[static_cast<size_t>(lam) * Nr + ir];
        } else {
            for (int ir = 0; ir < n_chi; ++ir) c[ir] = chi(n_occ - 1, lam, ir);
        }
        deriv5(c, Nr, dr, dci.data() + static_cast<size_t>(lam) * Nr);
    }
    // angular tables, sparse per mu: (nu, value) for |l_mu - l_nu| = 1
    const int qs[3] = {+1, -1, 0};
    std::vector<std::vector<std::pair<int, double>>> ang(3 * Np);
    for (int qi = 0; qi < 3; ++qi)
        for (int mu = 0; mu < Np; ++mu) {
            const int lm = l_mu[mu];
            for (int ln = lm - 1; ln <= lm + 1; ln += 2) {
                if (ln < 0 || ln > l_orb) continue;
                for (int mn = -ln; mn <= ln; ++mn) {
                    const double a = std::sqrt(4.0 * M_PI / 3.0) * real_gaunt(lm, m_mu[mu], 1, qs[qi], ln, mn);
                    if (a != 0.0) ang[qi * Np + mu].push_back({lm_index(ln, mn), a});
                }
            }
            std::sort(ang[qi * Np + mu].begin(), ang[qi * Np + mu].end());
        }
    const std::vector<double> w = simpson_weights(n_pts, dr);
    rows.assign(static_cast<size_t>(n_pts) * nslot * Np, 0.0);
    dcorr.assign(static_cast<size_t>(6) * n_occ, 0.0);
    std::vector<double> dc_acc(static_cast<size_t>(6) * n_occ * n_pts, 0.0);
#pragma omp parallel for schedule(static)
    for (int ir = 0; ir < n_pts; ++ir) {
        const double r = dr + ir * dr, inv_r = (r > 1e-30) ? 1.0 / r : 0.0;
        double* blk = rows.data() + static_cast<size_t>(ir) * nslot * Np;
        std::vector<double> xi(static_cast<size_t>(6) * Np, 0.0);
        for (int slot = 0; slot < 6; ++slot) {
            const int qi = slot % 3;
            const bool L = slot < 3;
            for (int mu = 0; mu < Np; ++mu) {
                double s = 0.0;
                for (const auto& e : ang[qi * Np + mu]) {
                    const double c = ci[static_cast<size_t>(e.first) * Nr + ir];
                    if (L) s += e.second * c;
                    else {
                        const int ln = static_cast<int>(std::sqrt(static_cast<double>(e.first) + 0.5));
                        const int lmu = l_mu[mu];
                        const double vc = (lmu == ln + 1) ? -static_cast<double>(ln + 1) : ((lmu == ln - 1) ? static_cast<double>(ln) : 0.0);
                        s += e.second * (dci[static_cast<size_t>(e.first) * Nr + ir] + vc * inv_r * c);
                    }
                }
                xi[static_cast<size_t>(slot) * Np + mu] = s;
                const double rf = L ? r : 1.0;
                blk[slot + static_cast<size_t>(mu) * nslot] = w[ir] * rf * s;
            }
        }
        for (int a = 0; a < n_occ; ++a) {
            for (int mu = 0; mu < Np; ++mu) {
                const double ph = (ir < n_chi) ? chi(a, mu, ir) : 0.0;
                blk[6 + a + static_cast<size_t>(mu) * nslot] = w[ir] * ph;
            }
            for (int slot = 0; slot < 6; ++slot) {
                double s = 0.0;
                for (int mu = 0; mu < Np; ++mu) s += ((ir < n_chi) ? chi(a, mu, ir) : 0.0) * xi[static_cast<size_t>(slot) * Np + mu];
                const double rf = (slot < 3) ? r : 1.0;
                dc_acc[(static_cast<size_t>(slot) * n_occ + a) * n_pts + ir] = rf * s;
            }
        }
    }
    for (int slot = 0; slot < 6; ++slot)
        for (int a = 0; a < n_occ; ++a) {
            double s = 0.0;
            const double* f = dc_acc.data() + (static_cast<size_t>(slot) * n_occ + a) * n_pts;
            for (int ir = 0; ir < n_pts; ++ir) s += w[ir] * f[ir];
            dcorr[static_cast<size_t>(slot) * n_occ + a] = s;
        }
}

}  // namespace sehf
