// sehf_scan.cpp -- the batched energy scan on one device.
//
// All energies of a batch advance through the radial grid together.  Per
// radial step n and energy b (SchurInverter / ForwardRPropagator,
// src/scatt/SchurInverter.cpp:247-415, ForwardRPropagator.cpp:299-386):
//   Sinv_b = sym((S0_n + shift_b(n) I)^{-1}),   S0_n = I - h^2/6 V_n - B_n D_n^{-1} B_n^T
//   X_b    = Sinv_b (B_n D_n^{-1})
//   U_b    = 12 W_n^{-1} - 10 I - R_{n-1,b}   (W^{-1} from Sinv_b, X_b, D^{-1})
//   R_n,b  = sym(U_b^{-1})
// Only sqrt(Nr)-spaced R checkpoints are kept.  The back-propagation
// (BackPropagator.cpp:603-703) walks the grid downwards one segment at a
// time, recomputing that segment's Sinv and R from its checkpoint, and folds
// the dipole integrands (DipoleMatrixElement.cpp:584-965) and the asymptotic
// fit residual (AsymptoticAmplitudes.cpp:64-359) into the sweep, so psi is
// never stored; the fit sums over the ~200-point window at the top of the
// grid come from a shorter pre-pass over the top segments.
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <condition_variable>
#include <mutex>
#include <string>
#include <thread>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sehf.h"
#include "sehf_common.h"
#include "sehf_host_math.h"
#include "sehf_plan.h"
#include "sehf_psi.h"
#include "sehf_scan_kernels.h"
#include "sehf_setup.h"
#include "sehf_setup_kernels.h"

namespace sehf {
std::string& last_error_ref();
}

using namespace sehf;

namespace {

struct Batch {
    int B = 0;   // capacity
    DevArray<double> Rp, Rc, U, Ms, X, Z, Zt, T, Psi, acc, YJ, YY, Am, Bm, Kd, P, Vpsi, G, Bf;
    DevArray<double> ckpt, segR, segS, W2, W2i, Rff, Rffi, T1, DRre, DRim, ore, oim, Dre, Dim, Ainv, Kfit;
    DevArray<double> dA, dS, zeros, cE, anl, tab, jy, Jc, Yc, coef, red, nrm;
    DevArray<double> SinvR;   // forward sweep: Sinv ring of two groups of points, filled a group ahead on the side stream
    DevArray<double> Ad, Ad2;   // forward-accumulated dipole rows (ns x Nt per energy, ping-pong)
    // streamed tables (energy independent, shared by the batch): S0_n, V_n and
    // B_n scratch of the side stream, B D^{-1} of the two forward ring slots
    // and of every point of the segment being recomputed
    DevArray<double> tS0, tV, tB, ringBD, segBD;
    GjBuffers gj;             // inversions on the main stream
    GjBuffers gj2;            // Schur inversions on the side stream
    // tangent along (E, alpha), allocated by the first tangent run
    bool has_tan = false;
    DevArray<double> dRp, dRc, dckpt, dsegR, Gt, dZ, dZt, P1, dPsi, dacc, dYJ, dYY, dAm, dBm, dKd, dP, Q1, Q2;
    DevArray<double> vre, vim, dDRre, dDRim, dore, doim, dDre, dDim;
    DevArray<double> wF, wP, wN, danl, dtab, djy, dJc, dYc, dcoef;
};

// One energy group: its own stream and batch buffers.  Groups run
// concurrently (one host thread each) and share the energy-independent
// device tables read-only.
struct Group {
    hipStream_t st = nullptr;    // main stream: forward recursion, K, back-propagation, fit, dipole
    hipStream_t st2 = nullptr;   // side stream: the Schur inverses Sinv_n, which depend on n and E only
    hipEvent_t ev_in = nullptr, ev_bwd = nullptr, ev_s[2] = {nullptr, nullptr}, ev_f[2] = {nullptr, nullptr};
    std::vector<hipEvent_t> ev_seg;
    Batch bt;
    double t_stage[8] = {0};
    double t_probe = 0.0;   // seconds of the last probe run (autotuner)
    // psi files: pinned staging ring between the device and the files
    static constexpr int kIoSlots = 4;
    DevArray<double> io_stage;
    hipEvent_t ev_io[kIoSlots] = {nullptr, nullptr, nullptr, nullptr};
    // SEHF_SCHUR_STREAM=main runs the Schur inverses on the main stream (one
    // stream per group: the same work in one queue, for tools that follow a
    // single queue such as per-dispatch counter collection).
    void init() {
        SEHF_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        const char* e = std::getenv("SEHF_SCHUR_STREAM");
        if (e && std::string(e) == "main") st2 = st;
        else SEHF_HIP_CHECK(hipStreamCreateWithFlags(&st2, hipStreamNonBlocking));
        for (hipEvent_t* e : {&ev_in, &ev_bwd, &ev_s[0], &ev_s[1], &ev_f[0], &ev_f[1]})
            SEHF_HIP_CHECK(hipEventCreateWithFlags(e, hipEventDisableTiming));
        for (hipEvent_t& e : ev_io) SEHF_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    void ensure_seg_events(int L) {
        while (static_cast<int>(ev_seg.size()) < L) {
            hipEvent_t e;
            SEHF_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            ev_seg.push_back(e);
        }
    }
    ~Group() {
        for (hipEvent_t e : {ev_in, ev_bwd, ev_s[0], ev_s[1], ev_f[0], ev_f[1]})
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : ev_seg) (void)hipEventDestroy(e);
        for (hipEvent_t e : ev_io)
            if (e) (void)hipEventDestroy(e);
        if (st2 && st2 != st) (void)hipStreamDestroy(st2);
        if (st) (void)hipStreamDestroy(st);
    }
};

}  // namespace

struct sehf_scan {
    sehf_ctx* ctx = nullptr;
    hipStream_t st = nullptr;
    std::unique_ptr<Setup> S;
    int Np = 0, Nf = 0, Nt = 0, Nr = 0, n_tr = 0, n_start = 1, n_pts = 0, nslot = 0, n_bd = 0, nstab = 0;
    double h = 0, c6 = 0;
    DevArray<double> S0, BD, Dinv, VN, BN, negDinvN, Drows, dcorr, cvd, bdb;
    // Streamed tables (SEHF_TABLES_STREAMED): S0_n and B_n D_n^{-1} are not
    // kept for every n but rebuilt per radial step on the device from these
    // inputs (V_s(r_n), chi(r_n), the per-point factors, the Gaunt tables),
    // by the kernels and in the order of the resident build, so bitwise equal.
    int tables = SEHF_TABLES_RESIDENT;
    struct TableInputs {
        DevArray<int> pi, pj, vsig, lmu, xlam;
        DevArray<int64_t> vptr, xptr;
        DevArray<double> vval, xval, vs, vp, ir2, chi, fac;
        DevArray<unsigned char> live;
        int64_t npairs = 0;
    } tin;
    std::vector<double> dcorr_h, vdiag;       // vdiag: (n_start+2) x Np potential diagonal
    std::vector<double> gA, gS, lA, lS;
    std::vector<char> hasLA, hasLS;
    std::vector<std::vector<double>> A0h, S0h;
    std::vector<std::unique_ptr<Group>> groups;
    int n_groups = 1;
    int ctx_device = 0;
    int dipole_path = SEHF_DIPOLES_FORWARD;    // how the dipole integrals are summed (sehf_scan_set_dipole_path)
    bool sinv_spd = true;                       // Sinv by SPD block elimination (else pivoted Gauss-Jordan)
    int ckpt_where = SEHF_CKPT_DEVICE;          // ratio-matrix checkpoints in HBM or pinned host memory
    std::string psi_root;                       // per-energy psi files (sehf_scan_set_psi_store), keep bits
    int psi_keep = 0;
    std::string replay_root;                    // set while sehf_scan_dipole_only replays from psi files
    std::mutex host_mu;   // lazily computed host tables (stability eigenvalues)
    size_t dev_bytes_shared = 0;
    double t_stage[8] = {0};   // setup, forward, kmatrix, back-prop, fit+dipole, host tables
};

// context stream accessor (sehf_ctx is defined in sehf_capi.cpp)
hipStream_t sehf_ctx_stream(const sehf_ctx* c);
int sehf_ctx_device(const sehf_ctx* c);

namespace {

template <class F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        last_error_ref() = e.what();
        return 1;
    } catch (...) {
        last_error_ref() = "unknown error";
        return 1;
    }
}

void build_shared(sehf_scan* s) {
    Setup& S = *s->S;
    hipStream_t st = s->st;
    const int Np = s->Np, Nf = s->Nf, Nr = s->Nr;
    const size_t pp = static_cast<size_t>(Np) * Np, pf = static_cast<size_t>(Np) * Nf;
    const bool streamed = s->tables == SEHF_TABLES_STREAMED;
    s->cvd.alloc(static_cast<size_t>(Np) * Nr);
    s->bdb.alloc(static_cast<size_t>(Np) * Nr);
    fill(s->bdb.p, static_cast<size_t>(Np) * Nr, 0.0, st);
    std::vector<double> cvd_h(static_cast<size_t>(Np) * Nr);
    s->n_bd = std::min(s->n_tr, Nr);
    if (!streamed) {
        s->S0.alloc(pp * Nr);
        if (Nf > 0) s->BD.alloc(pf * s->n_bd);
    }
    s->Dinv.alloc(static_cast<size_t>(std::max(Nf, 1)) * Nr);
    {
        std::vector<double> dh(static_cast<size_t>(std::max(Nf, 1)) * Nr, 0.0);
        for (int n = 0; n < Nr; ++n) S.dinv(n, dh.data() + static_cast<size_t>(n) * std::max(Nf, 1));
        SEHF_HIP_CHECK(hipMemcpyAsync(s->Dinv.p, dh.data(), dh.size() * sizeof(double), hipMemcpyHostToDevice, st));
        SEHF_HIP_CHECK(hipStreamSynchronize(st));
    }
    s->nstab = std::min(Setup::stab_n_max, Nr);
    s->A0h.assign(s->nstab, {});
    s->S0h.assign(s->nstab, {});
    // The dense V(r_n) and B(r_n) are built on the device from V_s(r), chi(r)
    // and the sparse Gaunt tables (sehf_setup_kernels.hip); with
    // SEHF_TABLES_ON_HOST set they are built on the host and uploaded instead
    // (the reference path of the equivalence test: the tables are bitwise
    // equal).  The inputs of the device build are uploaded for every point at
    // once; streamed tables keep them for the per-step rebuild.
    const bool on_host = std::getenv("SEHF_TABLES_ON_HOST") != nullptr && !streamed;
    const int chunk = 32;
    const int n_exp = S.n_exp, n_lam = S.n_lambda, n_occ = S.n_occ;
    std::vector<double> Vh(on_host ? pp * chunk : pp), Bh(on_host ? std::max<size_t>(pf, 1) * chunk : 1);
    DevArray<double> Vd(pp * chunk), Bd(std::max<size_t>(pf, 1) * chunk);
    // chunk-local S0 and B D^{-1} when they are not kept for every point
    DevArray<double> S0k, BDk;
    if (streamed) { S0k.alloc(pp * chunk); BDk.alloc(std::max<size_t>(pf, 1) * chunk); }
    auto& T = s->tin;
    auto upi = [&](auto& d, const auto& h) {
        d.alloc(std::max<size_t>(h.size(), 1));
        if (!h.empty())
            SEHF_HIP_CHECK(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(h[0]), hipMemcpyHostToDevice, st));
        SEHF_HIP_CHECK(hipStreamSynchronize(st));
    };
    if (!on_host) {
        upi(T.pi, S.pot_pair_i()); upi(T.pj, S.pot_pair_j()); upi(T.vptr, S.pot_ptr());
        upi(T.vsig, S.pot_sigma()); upi(T.vval, S.pot_val()); upi(T.lmu, S.l_mu);
        upi(T.xptr, S.exch_ptr()); upi(T.xlam, S.exch_lambda()); upi(T.xval, S.exch_val());
        T.npairs = static_cast<int64_t>(S.pot_pair_i().size());
        std::vector<double> vp_h(Nr), ir2_h(Nr), fac_h(Nr);
        std::vector<unsigned char> live_h(Nr);
        for (int n = 0; n < Nr; ++n) {
            const double r = n * s->h;
            live_h[n] = (r > 1e-14) ? 1 : 0;
            vp_h[n] = S.polarization(r);
            ir2_h[n] = 1.0 / (r * r);
            fac_h[n] = (r < 1e-10) ? 0.0 : 2.0 * std::sqrt(2.0 * M_PI) / r;
        }
        upi(T.vp, vp_h); upi(T.ir2, ir2_h); upi(T.fac, fac_h); upi(T.live, live_h);
        {
            std::vector<double> vs_h(static_cast<size_t>(Nr) * n_exp);
            S.vsigma_table(0, Nr, vs_h.data());
            upi(T.vs, vs_h);
        }
        if (Nf > 0 && s->n_bd > 0) {
            std::vector<double> chi_h(static_cast<size_t>(s->n_bd) * n_occ * n_lam);
            S.chi_table(0, s->n_bd, chi_h.data());
            upi(T.chi, chi_h);
        }
    }
    const double h2_12 = s->h * s->h / 12.0;
    for (int n0 = 0; n0 < Nr; n0 += chunk) {
        const int nc = std::min(chunk, Nr - n0);
        const int nb = std::max(0, std::min(nc, s->n_bd - n0));
        if (on_host) {
            S.build_chunk(n0, nc, Vh.data(), Bh.data());
            for (int q = 0; q < nc; ++q)
                for (int i = 0; i < Np; ++i)
                    cvd_h[static_cast<size_t>(n0 + q) * Np + i] = s->c6 * Vh[q * pp + i + static_cast<size_t>(i) * Np];
            SEHF_HIP_CHECK(hipMemcpyAsync(Vd.p, Vh.data(), nc * pp * sizeof(double), hipMemcpyHostToDevice, st));
            if (Nf > 0 && nb > 0)
                SEHF_HIP_CHECK(hipMemcpyAsync(Bd.p, Bh.data(), nb * pf * sizeof(double), hipMemcpyHostToDevice, st));
        } else {
            build_V(Vd.p, T.vs.p + static_cast<size_t>(n0) * n_exp, n_exp, T.pi.p, T.pj.p, T.vptr.p, T.vsig.p,
                    T.vval.p, T.npairs, Np, T.lmu.p, T.vp.p + n0, T.ir2.p + n0, T.live.p + n0, nc, st);
            scaled_diag(s->cvd.p + static_cast<size_t>(n0) * Np, Vd.p, s->c6, Np, nc, st);
            if (Nf > 0 && nb > 0)
                build_B(Bd.p, T.chi.p + static_cast<size_t>(n0) * n_occ * n_lam, n_occ, n_lam, T.xptr.p, T.xlam.p,
                        T.xval.p, Np, S.n_sigma, T.fac.p + n0, h2_12, nb, st);
        }
        double* S0c = streamed ? S0k.p : s->S0.p + static_cast<size_t>(n0) * pp;
        double* BDc = streamed ? BDk.p : (Nf > 0 ? s->BD.p + static_cast<size_t>(n0) * pf : nullptr);
        // Streamed tables keep S0 only for the stability-shift host copies
        // (n < nstab); past them the chunk needs only B D^{-1} (for bdb).
        const bool need_S0 = !streamed || n0 < s->nstab;
        // S0 = I - h^2/6 V
        if (need_S0) axpbyi(S0c, Np, pp, Vd.p, Np, pp, -s->c6, nullptr, 0, 0, 0.0, 1.0, Np, Np, nc, st);
        if (Nf > 0 && nb > 0) {
            for (int q = 0; q < nb; ++q)
                col_scale(BDc + q * pf, Np, Bd.p + q * pf, Np, s->Dinv.p + static_cast<size_t>(n0 + q) * Nf, Np, Nf,
                          st);
            // (B D^{-1} B^T)_ii as a separate dot product for the per-energy diagonal
            row_dot(s->bdb.p + static_cast<size_t>(n0) * Np, BDc, Bd.p, Np, Nf, nb, st);
            if (need_S0) {
                Gemm g;   // S0 -= (B D^{-1}) B^T
                g.M = Np; g.N = Np; g.K = Nf;
                g.A = BDc; g.lda = Np; g.sA = pf;
                g.B = Bd.p; g.ldb = Np; g.sB = pf; g.tB = 1;
                g.C = S0c; g.ldc = Np; g.sC = pp;
                g.alpha = -1.0; g.beta = 1.0; g.batch = nb;
                gemm(g, st);
            }
        }
        // host copies of the small-r matrices that may need the stability shift
        for (int q = 0; q < nc && n0 + q < s->nstab; ++q) {
            const int n = n0 + q;
            const double* Vq = Vh.data() + (on_host ? q * pp : 0);
            if (!on_host) {
                SEHF_HIP_CHECK(hipMemcpyAsync(Vh.data(), Vd.p + q * pp, pp * sizeof(double), hipMemcpyDeviceToHost, st));
                SEHF_HIP_CHECK(hipStreamSynchronize(st));
            }
            auto& A0 = s->A0h[n];
            A0.resize(pp);
            for (size_t t = 0; t < pp; ++t) A0[t] = -s->c6 * Vq[t];
            for (int i = 0; i < Np; ++i) A0[i + static_cast<size_t>(i) * Np] += 1.0;
            s->S0h[n].resize(pp);
            SEHF_HIP_CHECK(hipMemcpyAsync(s->S0h[n].data(), S0c + q * pp, pp * sizeof(double), hipMemcpyDeviceToHost, st));
        }
        if (n0 + nc >= Nr) {   // boundary n = Nr-1: V_N and unscaled B_N
            const int q = Nr - 1 - n0;
            s->VN.alloc(pp);
            SEHF_HIP_CHECK(hipMemcpyAsync(s->VN.p, Vd.p + q * pp, pp * sizeof(double), hipMemcpyDeviceToDevice, st));
            if (Nf > 0) {
                s->BN.alloc(pf);
                if (q < nb)
                    SEHF_HIP_CHECK(hipMemcpyAsync(s->BN.p, Bd.p + q * pf, pf * sizeof(double), hipMemcpyDeviceToDevice, st));
                else   // beyond the exchange cut-off B is zero (ExchangeCoupling.cpp:66-136)
                    fill(s->BN.p, pf, 0.0, st);
            }
        }
        SEHF_HIP_CHECK(hipStreamSynchronize(st));
    }
    if (on_host)
        SEHF_HIP_CHECK(hipMemcpyAsync(s->cvd.p, cvd_h.data(), cvd_h.size() * sizeof(double), hipMemcpyHostToDevice, st));
    SEHF_HIP_CHECK(hipStreamSynchronize(st));
    if (!streamed) s->tin = sehf_scan::TableInputs();   // only the per-step rebuild needs them
    // Gershgorin bounds of A0_n and S0_n (SchurInverter.cpp:30-49)
    auto gersh = [Np](const std::vector<double>& M) {
        double lb = 1e300;
        for (int i = 0; i < Np; ++i) {
            double off = 0.0;
            for (int j = 0; j < Np; ++j) if (j != i) off += std::fabs(M[i + static_cast<size_t>(j) * Np]);
            lb = std::min(lb, M[i + static_cast<size_t>(i) * Np] - off);
        }
        return lb;
    };
    s->gA.assign(s->nstab, 0.0); s->gS.assign(s->nstab, 0.0);
    s->lA.assign(s->nstab, 0.0); s->lS.assign(s->nstab, 0.0);
    s->hasLA.assign(s->nstab, 0); s->hasLS.assign(s->nstab, 0);
    for (int n = 0; n < s->nstab; ++n) { s->gA[n] = gersh(s->A0h[n]); s->gS[n] = gersh(s->S0h[n]); }
    if (Nf > 0) {
        std::vector<double> nd(Nf);
        S.dinv(Nr - 1, nd.data());
        for (double& v : nd) v = -v;
        s->negDinvN.alloc(Nf);
        SEHF_HIP_CHECK(hipMemcpyAsync(s->negDinvN.p, nd.data(), Nf * sizeof(double), hipMemcpyHostToDevice, st));
    }
    // dipole rows and orthogonalisation integrals
    std::vector<double> rows;
    S.dipole_tables(rows, s->dcorr_h, s->n_pts);
    s->Drows.alloc(rows.size());
    SEHF_HIP_CHECK(hipMemcpyAsync(s->Drows.p, rows.data(), rows.size() * sizeof(double), hipMemcpyHostToDevice, st));
    s->dcorr.alloc(std::max<size_t>(s->dcorr_h.size(), 1));
    SEHF_HIP_CHECK(hipMemcpyAsync(s->dcorr.p, s->dcorr_h.data(), s->dcorr_h.size() * sizeof(double), hipMemcpyHostToDevice, st));
    // potential diagonal for the analytic start (ForwardRPropagator.cpp:78-134)
    s->vdiag.assign(static_cast<size_t>(s->n_start + 2) * Np, 0.0);
    {
        std::vector<double> V(pp);
        for (int n = 0; n <= std::min(s->n_start + 1, Nr - 1); ++n) {
            S.potential(n, V.data());
            for (int i = 0; i < Np; ++i) s->vdiag[static_cast<size_t>(n) * Np + i] = V[i + static_cast<size_t>(i) * Np];
        }
    }
    SEHF_HIP_CHECK(hipStreamSynchronize(st));
}

// Streamed tables: S0_n and, for n < n_bd, B_n D_n^{-1} (into BDout) rebuilt
// on the device for one radial point, with the resident build's kernels in
// its order (bitwise the same values).  S0_n goes to b.tS0.
void tables_at(sehf_scan* s, Batch& b, int n, double* BDout, hipStream_t st) {
    const Setup& S = *s->S;
    auto& T = s->tin;
    const int Np = s->Np, Nf = s->Nf;
    const size_t pp = static_cast<size_t>(Np) * Np, pf = static_cast<size_t>(Np) * Nf;
    build_V(b.tV.p, T.vs.p + static_cast<size_t>(n) * S.n_exp, S.n_exp, T.pi.p, T.pj.p, T.vptr.p, T.vsig.p, T.vval.p,
            T.npairs, Np, T.lmu.p, T.vp.p + n, T.ir2.p + n, T.live.p + n, 1, st);
    axpbyi(b.tS0.p, Np, pp, b.tV.p, Np, pp, -s->c6, nullptr, 0, 0, 0.0, 1.0, Np, Np, 1, st);
    if (Nf > 0 && n < s->n_bd) {
        build_B(b.tB.p, T.chi.p + static_cast<size_t>(n) * S.n_occ * S.n_lambda, S.n_occ, S.n_lambda, T.xptr.p,
                T.xlam.p, T.xval.p, Np, S.n_sigma, T.fac.p + n, s->h * s->h / 12.0, 1, st);
        col_scale(BDout, Np, b.tB.p, Np, s->Dinv.p + static_cast<size_t>(n) * Nf, Np, Nf, st);
        Gemm g;   // S0 -= (B D^{-1}) B^T
        g.M = Np; g.N = Np; g.K = Nf;
        g.A = BDout; g.lda = Np; g.sA = pf;
        g.B = b.tB.p; g.ldb = Np; g.sB = pf; g.tB = 1;
        g.C = b.tS0.p; g.ldc = Np; g.sC = pp;
        g.alpha = -1.0; g.beta = 1.0; g.batch = 1;
        gemm(g, st);
    }
}

double lmin_A(sehf_scan* s, int n) {
    std::lock_guard<std::mutex> lock(s->host_mu);
    if (!s->hasLA[n]) { s->lA[n] = sym_min_eig(s->Np, s->A0h[n].data()); s->hasLA[n] = 1; }
    return s->lA[n];
}
double lmin_S(sehf_scan* s, int n) {
    std::lock_guard<std::mutex> lock(s->host_mu);
    if (!s->hasLS[n]) { s->lS[n] = sym_min_eig(s->Np, s->S0h[n].data()); s->hasLS[n] = 1; }
    return s->lS[n];
}

// Diagonal shift of S0_n for energy E (SchurInverter.cpp:254-372): h^2/6 E,
// plus the Johnson stability shifts of A and S for n < 60.
void sinv_shift(sehf_scan* s, int n, double E, int& shA, int& shS, double& dA_out, double& dS_out) {
    const double cE = s->c6 * E;
    dA_out = dS_out = 0.0;
    if (n >= s->nstab) return;
    const double thr = Setup::W_min;
    double meA = s->gA[n] + cE;
    if (!(meA > thr + 0.05)) meA = lmin_A(s, n) + cE;
    double dA = 0.0;
    if (meA < thr) { dA = thr - meA + Setup::shift_margin; ++shA; }
    dA_out = dA;
    if (n >= s->n_tr || s->Nf == 0) return;
    double meS = s->gS[n] + cE + dA;
    if (!(meS > thr + 0.05)) meS = lmin_S(s, n) + cE + dA;
    if (meS < thr) { dS_out = thr - meS + Setup::shift_margin; ++shS; }
}

}  // namespace

namespace sehf {

// Bytes of the Gauss-Jordan workspace of `batch` matrices of order n.
size_t gj_buffer_bytes(size_t n, size_t batch) {
    return batch * (sizeof(int) * (2 * n + 2 + gj_nom_elems(static_cast<int>(n)) + gj_stgi_elems()) +
                    sizeof(double) * (gj_tbuf_elems(static_cast<int>(n)) + gj_stg_elems()));
}

// The batch buffers of the scan, for `B` energies with checkpoint spacing
// L and n_ck checkpoints per energy: `al(array, count, in_host_memory)` is
// called once per buffer.  alloc_batch allocates through it and the
// storage planner counts through it, so the two cannot disagree.
template <class F>
void batch_arrays(Batch& b, const BatchDims& d, int B, int L, int n_ck, int n_fit, bool tan, F&& al) {
    const size_t Np = d.Np, Nf = d.Nf, Nt = d.Nt, ns = d.ns;
    const size_t tt = Nt * Nt, pp = Np * Np, pf = Np * std::max<size_t>(Nf, 1), ff = std::max<size_t>(Nf, 1) * std::max<size_t>(Nf, 1);
    const bool hc = d.ckpt_host;
    al(b.Rp, tt * B, false); al(b.Rc, tt * B, false); al(b.U, tt * B, false);
    al(b.Ms, pp * B, false); al(b.X, pf * B, false);
    al(b.Z, Nt * Np * B, false); al(b.Zt, Nt * Np * B, false);
    al(b.T, pp * B, false); al(b.Psi, pp * B, false); al(b.acc, ns * Np * B, false);
    al(b.YJ, pp * B, false); al(b.YY, pp * B, false); al(b.Am, pp * B, false); al(b.Bm, pp * B, false);
    al(b.Kd, pp * B, false);
    al(b.P, pp * B, false); al(b.Vpsi, pp * B, false); al(b.G, pf * B, false); al(b.Bf, pp * B, false);
    al(b.ckpt, tt * B * std::max(n_ck, 1), hc);
    al(b.segR, tt * B * L, false); al(b.segS, pp * B * L, false);
    al(b.W2, 4 * pp * B, false); al(b.W2i, 4 * pp * B, false);
    al(b.Rff, ff * B, false); al(b.Rffi, ff * B, false); al(b.T1, pf * B, false);
    al(b.DRre, 6 * Np * B, false); al(b.DRim, 6 * Np * B, false);
    al(b.ore, Np * d.n_occ * B, false); al(b.oim, Np * d.n_occ * B, false);
    al(b.Dre, 6 * Np * B, false); al(b.Dim, 6 * Np * B, false); al(b.Ainv, pp * B, false); al(b.Kfit, pp * B, false);
    al(b.dA, (d.nstab + 1) * B, false); al(b.dS, (d.nstab + 1) * B, false);
    al(b.zeros, static_cast<size_t>(B), false); al(b.cE, static_cast<size_t>(B), false);
    al(b.anl, (d.n_start + 1) * B * Nt, false);
    al(b.tab, 6 * Np * B, false); al(b.jy, 2 * Np * B, false);
    al(b.Jc, std::max<size_t>(n_fit, 1) * Np * B, false); al(b.Yc, std::max<size_t>(n_fit, 1) * Np * B, false);
    al(b.coef, 5 * Np * B, false); al(b.red, 2 * static_cast<size_t>(B), false); al(b.nrm, 2 * static_cast<size_t>(B), false);
    al(b.SinvR, 2 * d.sinv_group * pp * B, false);
    al(b.Ad, ns * Nt * B, false); al(b.Ad2, ns * Nt * B, false);
    if (d.streamed) {
        al(b.tS0, pp, false); al(b.tV, pp, false); al(b.tB, pf, false);
        al(b.ringBD, 2 * d.sinv_group * pf, false); al(b.segBD, pf * L, false);
    }
    if (tan) {
        const size_t no = d.n_occ, nx = std::max<size_t>(6, no);
        al(b.dRp, tt * B, false); al(b.dRc, tt * B, false);
        al(b.dckpt, tt * B * std::max(n_ck, 1), hc); al(b.dsegR, tt * B * L, false);
        al(b.Gt, Np * Nt * B, false); al(b.dZ, Nt * Np * B, false); al(b.dZt, Nt * Np * B, false);
        al(b.P1, pp * B, false); al(b.dPsi, pp * B, false); al(b.dacc, ns * Np * B, false);
        al(b.dYJ, pp * B, false); al(b.dYY, pp * B, false); al(b.dAm, pp * B, false); al(b.dBm, pp * B, false);
        al(b.dKd, pp * B, false);
        al(b.dP, pp * B, false); al(b.Q1, pp * B, false); al(b.Q2, pp * B, false);
        al(b.vre, Np * nx * B, false); al(b.vim, Np * nx * B, false);
        al(b.dDRre, 6 * Np * B, false); al(b.dDRim, 6 * Np * B, false);
        al(b.dore, Np * no * B, false); al(b.doim, Np * no * B, false);
        al(b.dDre, 6 * Np * B, false); al(b.dDim, 6 * Np * B, false);
        al(b.wF, d.Nr * B, false); al(b.wP, d.Nr * B, false); al(b.wN, static_cast<size_t>(B), false);
        al(b.danl, (d.n_start + 1) * B * Nt, false);
        al(b.dtab, 4 * Np * B, false); al(b.djy, 2 * Np * B, false);
        al(b.dJc, std::max<size_t>(n_fit, 1) * Np * B, false); al(b.dYc, std::max<size_t>(n_fit, 1) * Np * B, false);
        al(b.dcoef, 4 * Np * B, false);
    }
}

BatchBytes batch_bytes(const BatchDims& d, int B, int L, int n_ck, int n_fit, bool tan) {
    Batch b;
    BatchBytes out;
    batch_arrays(b, d, B, L, n_ck, n_fit, tan, [&](auto& arr, size_t count, bool host) {
        (host ? out.host : out.device) += count * sizeof(*arr.p);
    });
    out.device += gj_buffer_bytes(std::max(d.Nt, 2 * d.Np), B) + gj_buffer_bytes(d.Np, B * d.sinv_group);
    return out;
}

}  // namespace sehf

namespace {

BatchDims dims_of(const sehf_scan* s) {
    BatchDims d;
    d.Np = s->Np; d.Nf = s->Nf; d.Nt = s->Nt; d.ns = s->nslot; d.n_occ = s->S->n_occ;
    d.nstab = s->nstab; d.n_start = s->n_start; d.Nr = s->Nr;
    d.streamed = s->tables == SEHF_TABLES_STREAMED;
    d.ckpt_host = s->ckpt_where == SEHF_CKPT_HOST;
    d.sinv_group = sinv_group_for(d.Np, s->sinv_spd);
    return d;
}

// n_ck: segment start matrices (checkpoints) kept per energy.
void alloc_batch(sehf_scan* s, Group& g, int B, int L, int n_ck, int n_fit, bool tan) {
    Batch& b = g.bt;
    const BatchDims d = dims_of(s);
    if (b.B >= B && b.segR.n >= static_cast<size_t>(L) * B * s->Nt * s->Nt &&
        b.ckpt.n >= static_cast<size_t>(std::max(n_ck, 1)) * B * s->Nt * s->Nt &&
        b.Jc.n >= static_cast<size_t>(std::max(n_fit, 1)) * B * s->Np && (!tan || b.has_tan) &&
        b.ckpt.host == d.ckpt_host && (!d.streamed || b.segBD.n >= static_cast<size_t>(L) * s->Np * std::max(s->Nf, 1)))
        return;
    b = Batch();
    b.B = B;
    batch_arrays(b, d, B, L, n_ck, n_fit, tan, [](auto& arr, size_t count, bool host) {
        if (host) arr.alloc_host(count);
        else arr.alloc(count);
    });
    fill(b.zeros.p, B, 0.0, g.st);
    b.gj.ensure(static_cast<int>(std::max(d.Nt, 2 * d.Np)), B);
    b.gj2.ensure(static_cast<int>(d.Np), B * static_cast<int>(d.sinv_group));
    b.has_tan = tan;
    g.ensure_seg_events(L);
}

void up(double* d, const std::vector<double>& h, hipStream_t st) {
    if (!h.empty()) SEHF_HIP_CHECK(hipMemcpyAsync(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, st));
}

double wall() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Fit window [n_fs, n_fe] (AsymptoticAmplitudes.cpp:70-91), clamped to the
// 300-point tail that the driver keeps (main.cpp:1335).
void fit_window(int Nr, int& n_fs, int& n_fe) {
    const int N = Nr - 1;
    const int asym_off = std::max(0, N - 300 + 1);
    n_fs = std::max(Nr - 200, static_cast<int>(0.9 * Nr));
    n_fe = Nr - 1 - 3;
    n_fs = std::max(n_fs, asym_off);
    n_fe = std::min(n_fe, asym_off + 300 - 1);
    if (n_fe - n_fs + 1 < 10) throw std::runtime_error("asymptotic fit window is too small");
}

// First segment whose start matrix is kept: every segment when the dipoles
// are back-propagated over the whole grid (and in tangent runs, which always
// back-propagate), else only the segments from the one holding the fit
// window's first point (the only back-propagation left).
// The psi tail, when it is written to file, extends the fit pre-pass down to
// the tail's first point.
int psi_tail_lo(int Nr) { return std::max(0, Nr - kPsiTailPoints); }

bool back_propagated_dipoles(const sehf_scan* s, bool tan) {
    return s->dipole_path == SEHF_DIPOLES_BACKWARD || tan || (s->psi_keep & SEHF_PSI_MAIN);
}

int ckpt_lo(const sehf_scan* s, int L, bool tan) {
    if (back_propagated_dipoles(s, tan)) return 0;
    int n_fs = 0, n_fe = 0;
    fit_window(s->Nr, n_fs, n_fe);
    if (s->psi_keep & SEHF_PSI_TAIL) n_fs = std::min(n_fs, psi_tail_lo(s->Nr));
    return n_fs / L;
}

PsiLayout psi_layout(const sehf_scan* s) {
    PsiLayout l;
    l.Np = s->Np; l.Nr = s->Nr;
    l.tail_lo = psi_tail_lo(s->Nr);
    l.n_tail = s->Nr - l.tail_lo;
    l.n_main = s->n_pts;
    return l;
}

// One batch of energies through the whole path.  With `dir` (per energy the
// direction (t_E, t_alpha)) the forward-mode derivative of every output along
// t_E d/dE + t_alpha d/dalpha is propagated alongside (tout receives it).
// Start line of concurrent probes: every group's host thread prepares its
// tables, then all start their timed steps together.
struct ProbeSync {
    std::mutex mu;
    std::condition_variable cv;
    int waiting = 0, total = 1;
    double t_start = 0.0;
    double arrive(double now) {
        std::unique_lock<std::mutex> lk(mu);
        if (waiting >= total) return t_start;   // already released (a failing thread arriving late)
        if (++waiting == total) {
            t_start = now;
            cv.notify_all();
        } else {
            cv.wait(lk, [&] { return waiting >= total; });
        }
        return t_start;
    }
};

// probe > 0: time only the first `probe` steps of the forward sweep (Schur
// inverses + ratio steps, with the tangent when dir is set) and return;
// nothing is written to out.  g.t_probe = end of this group's steps minus the
// common start of all groups (sync).
// seg >= 0: one segment cycle (see sehf_scan_run_segment): the forward ratio
// steps of radial segment `seg`, its recompute from the segment's start
// matrix and its back-propagation, with nothing written to out; seg_t[4]
// receives the seconds of the host tables, forward, recompute and
// back-propagation parts.
void run_batch(sehf_scan* s, Group& g, const double* ks, const double* dir, int nb, int L, sehf_scan_out* out,
               sehf_scan_tangent_out* tout, int out_off, int probe = 0, ProbeSync* sync = nullptr, int seg = -1,
               double* seg_t = nullptr) {
    const double tw0 = wall();
    Setup& S = *s->S;
    hipStream_t st = g.st;
    Batch& b = g.bt;
    const int Np = s->Np, Nf = s->Nf, Nt = s->Nt, Nr = s->Nr, N = Nr - 1, ns = s->nslot, n_occ = S.n_occ;
    const size_t tt = static_cast<size_t>(Nt) * Nt, pp = static_cast<size_t>(Np) * Np, pf = static_cast<size_t>(Np) * std::max(Nf, 1);
    const int n_seg = (N + L - 1) / L;
    const double h = s->h;
    const bool tan = dir != nullptr;
    // radial points per Schur-inverse call (the SPD inverses of sgrp points
    // share one launch sequence; one point with the Gauss-Jordan inverse)
    const int sgrp = sinv_group_for(static_cast<size_t>(s->Np), s->sinv_spd);
    // -Sinv (B D^{-1}) of the current step: inside Gt = [Sinv | -Sinv B D^{-1}]
    // (= (W^{-1})[:, psi]^T) for tangent runs, else its own buffer
    double* const Xd = tan ? b.Gt.p + pp : b.X.p;
    const int64_t sXd = tan ? static_cast<int64_t>(Np) * Nt : static_cast<int64_t>(pf);
    // psi files: written by a whole-scan run with a psi store set; read back
    // (instead of solving) while sehf_scan_dipole_only replays the fit and
    // the dipoles
    const bool replay = !s->replay_root.empty();
    const PsiLayout plo = psi_layout(s);
    std::unique_ptr<PsiWriter> pw;
    std::unique_ptr<PsiReader> pr;
    if (replay) pr.reset(new PsiReader(s->replay_root, out_off, nb, ks, plo, s->n_pts > 0));
    else if (s->psi_keep && seg < 0 && probe == 0)
        pw.reset(new PsiWriter(s->psi_root, out_off, nb, ks, plo, s->psi_keep & SEHF_PSI_TAIL, s->psi_keep & SEHF_PSI_MAIN));
    if ((pw || pr) && g.io_stage.n < static_cast<size_t>(Group::kIoSlots) * nb * pp)
        g.io_stage.alloc_host(static_cast<size_t>(Group::kIoSlots) * nb * pp);
    // Ring of kIoSlots pinned slots: a slot's file write (or the copy that
    // read it) is completed just before the slot is used again, so the host
    // keeps a few points of device work queued while it does file I/O.
    int io_next = 0;
    int io_pend[Group::kIoSlots][2];
    for (auto& q : io_pend) { q[0] = -1; q[1] = -1; }
    auto io_slot = [&](int i) { return g.io_stage.p + static_cast<size_t>(i) * nb * pp; };
    auto io_retire = [&](int i) {
        if (io_pend[i][0] < 0) return;
        SEHF_HIP_CHECK(hipEventSynchronize(g.ev_io[i]));
        if (pw) pw->put(io_pend[i][0], io_pend[i][1], io_slot(i));
        io_pend[i][0] = -1;
    };
    auto psi_out = [&](int file, int n, const double* dpsi) {
        if (!pw || file < 0 || !pw->wants(file, n)) return;
        const int i = io_next++ % Group::kIoSlots;
        io_retire(i);
        SEHF_HIP_CHECK(hipMemcpyAsync(io_slot(i), dpsi, nb * pp * sizeof(double), hipMemcpyDeviceToHost, g.st));
        SEHF_HIP_CHECK(hipEventRecord(g.ev_io[i], g.st));
        io_pend[i][0] = file; io_pend[i][1] = n;
    };
    auto psi_in = [&](int n, double* dpsi) {
        const int i = io_next++ % Group::kIoSlots;
        io_retire(i);
        pr->get(n, io_slot(i));
        SEHF_HIP_CHECK(hipMemcpyAsync(dpsi, io_slot(i), nb * pp * sizeof(double), hipMemcpyHostToDevice, g.st));
        SEHF_HIP_CHECK(hipEventRecord(g.ev_io[i], g.st));
        io_pend[i][0] = 2; io_pend[i][1] = n;
    };
    auto io_drain = [&] {
        for (int i = 0; i < Group::kIoSlots; ++i) io_retire(i);
    };
    // ---------------- host tables for this batch ----------------
    std::vector<double> E(nb), cE(nb);
    for (int e = 0; e < nb; ++e) { E[e] = 0.5 * ks[e] * ks[e]; cE[e] = s->c6 * E[e]; }
    std::vector<int> shA(nb, 0), shS(nb, 0);
    std::vector<double> dA(static_cast<size_t>(s->nstab) * nb), dS(dA.size());
    for (int n = 0; n < s->nstab; ++n)
        for (int e = 0; e < nb; ++e)
            sinv_shift(s, n, E[e], shA[e], shS[e], dA[static_cast<size_t>(n) * nb + e], dS[static_cast<size_t>(n) * nb + e]);
    // analytic diagonal R for n in [0, n_start)  (ForwardRPropagator.cpp:78-134)
    const int ns0 = s->n_start;
    std::vector<double> anl(static_cast<size_t>(ns0) * nb * Nt, 0.0);
    for (int n = 1; n < ns0; ++n)
        for (int e = 0; e < nb; ++e) {
            double* d = anl.data() + (static_cast<size_t>(n) * nb + e) * Nt;
            const double k = ks[e], h2_12 = h * h / 12.0, h2_6 = h * h / 6.0;
            const double rn = h + n * h, rn1 = h + (n + 1) * h;
            for (int mu = 0; mu < Np; ++mu) {
                const double Mn = 1.0 + h2_6 * (E[e] - s->vdiag[static_cast<size_t>(n) * Np + mu]);
                const double Mn1 = 1.0 + h2_6 * (E[e] - s->vdiag[static_cast<size_t>(n + 1) * Np + mu]);
                const double den = Mn1 * ric_j(S.l_mu[mu], k * rn1);
                if (std::fabs(den) > 1e-300) d[mu] = (Mn * ric_j(S.l_mu[mu], k * rn)) / den;
            }
            for (int f = 0; f < Nf; ++f) {
                const int l = S.l_sig[f % S.n_sigma];
                double Mn = 1.0 - h2_12 * (double(l * (l + 1)) / (rn * rn));
                double Mn1 = 1.0 - h2_12 * (double(l * (l + 1)) / (rn1 * rn1));
                if (Mn < Setup::W_min) Mn = Setup::W_min;
                if (Mn1 < Setup::W_min) Mn1 = Setup::W_min;
                const double den = Mn1 * std::pow(rn1, static_cast<double>(l + 1));
                if (std::fabs(den) > 1e-300) d[Np + f] = (Mn * std::pow(rn, static_cast<double>(l + 1))) / den;
            }
        }
    // K-matrix matching tables (KMatrixExtractor.cpp:126-144), psi_N tables
    std::vector<double> tab(static_cast<size_t>(6) * Np * nb), jy(static_cast<size_t>(2) * Np * nb);
    const double r_in = h + (Nr - 2) * h, r_out = h + (Nr - 1) * h, rN = h + (Nr - 1) * h;
    for (int e = 0; e < nb; ++e) {
        const double k = ks[e], k2 = k * k, h2_12 = h * h / 12.0;
        double* t = tab.data() + static_cast<size_t>(e) * 6 * Np;
        double* q = jy.data() + static_cast<size_t>(e) * 2 * Np;
        for (int mu = 0; mu < Np; ++mu) {
            const int l = S.l_mu[mu];
            t[mu] = ric_j(l, k * r_in); t[Np + mu] = ric_j(l, k * r_out);
            t[2 * Np + mu] = ric_y(l, k * r_in); t[3 * Np + mu] = ric_y(l, k * r_out);
            t[4 * Np + mu] = 1.0 + h2_12 * (k2 - static_cast<double>(l * (l + 1)) / (r_in * r_in));
            t[5 * Np + mu] = 1.0 + h2_12 * (k2 - static_cast<double>(l * (l + 1)) / (r_out * r_out));
            q[mu] = ric_j(l, k * rN); q[Np + mu] = ric_y(l, k * rN);
        }
    }
    int n_fs = 0, n_fe = 0;
    fit_window(Nr, n_fs, n_fe);
    const int n_fit = n_fe - n_fs + 1;
    std::vector<double> Jc(static_cast<size_t>(n_fit) * Np * nb), Yc(Jc.size()), coef(static_cast<size_t>(5) * Np * nb);
    std::vector<double> dJc(tan ? Jc.size() : 0), dYc(dJc.size()), dcoef(tan ? static_cast<size_t>(4) * Np * nb : 0);
    std::vector<int> n_closed(nb, 0);
    {
        int lmax = 0;
        for (int mu = 0; mu < Np; ++mu) lmax = std::max(lmax, S.l_mu[mu]);
        std::vector<double> jl(static_cast<size_t>(n_fit) * (lmax + 1)), yl(jl.size());
        for (int e = 0; e < nb; ++e) {
            const double k = ks[e];
            for (int l = 0; l <= lmax; ++l)
                for (int i = 0; i < n_fit; ++i) {
                    const double r = h + (n_fs + i) * h;
                    jl[static_cast<size_t>(l) * n_fit + i] = ric_j(l, k * r);
                    yl[static_cast<size_t>(l) * n_fit + i] = ric_y(l, k * r);
                }
            std::vector<double> m(5 * (lmax + 1));
            for (int l = 0; l <= lmax; ++l) {
                double a = 0, bb = 0, c = 0;
                for (int i = 0; i < n_fit; ++i) {
                    const double x = jl[static_cast<size_t>(l) * n_fit + i], y = yl[static_cast<size_t>(l) * n_fit + i];
                    a += x * x; bb += x * y; c += y * y;
                }
                const double det = a * c - bb * bb;
                const double mx = std::max(std::fabs(a), std::max(std::fabs(bb), std::fabs(c)));
                if (!(std::fabs(det) > 1e-20 * mx * mx)) {
                    if (!S.P.allow_closed)
                        throw std::runtime_error("asymptotic fit: singular normal matrix for l=" + std::to_string(l) +
                                                 " (closed channel; enable allow_closed)");
                    m[5 * l + 4] = 1.0;
                    ++n_closed[e];
                } else {
                    const double id = 1.0 / det;
                    m[5 * l + 0] = c * id; m[5 * l + 1] = -bb * id; m[5 * l + 2] = -bb * id; m[5 * l + 3] = a * id;
                }
            }
            double* cf = coef.data() + static_cast<size_t>(e) * 5 * Np;
            for (int mu = 0; mu < Np; ++mu) {
                const int l = S.l_mu[mu];
                for (int q = 0; q < 5; ++q) cf[q * Np + mu] = m[5 * l + q];
                for (int i = 0; i < n_fit; ++i) {
                    Jc[(static_cast<size_t>(e) * n_fit + i) * Np + mu] = jl[static_cast<size_t>(l) * n_fit + i];
                    Yc[(static_cast<size_t>(e) * n_fit + i) * Np + mu] = yl[static_cast<size_t>(l) * n_fit + i];
                }
            }
            if (!tan) continue;
            // tangent of the fit tables: dj = j'(kr) r dk, and of the 2x2
            // inverse normal matrices, dMinv = -Minv dM Minv
            const double dk = dir[2 * e] / k;
            std::vector<double> djl(jl.size()), dyl(jl.size()), dm(4 * (lmax + 1), 0.0);
            for (int l = 0; l <= lmax; ++l) {
                double d00 = 0, d01 = 0, d11 = 0;
                for (int i = 0; i < n_fit; ++i) {
                    const size_t q = static_cast<size_t>(l) * n_fit + i;
                    const double r = h + (n_fs + i) * h;
                    djl[q] = ric_jp(l, k * r) * r * dk;
                    dyl[q] = ric_yp(l, k * r) * r * dk;
                    d00 += 2.0 * jl[q] * djl[q];
                    d01 += djl[q] * yl[q] + jl[q] * dyl[q];
                    d11 += 2.0 * yl[q] * dyl[q];
                }
                if (m[5 * l + 4] != 0.0) continue;   // closed: A = I, B = 0 do not move
                const double a = m[5 * l + 0], bq = m[5 * l + 1], c = m[5 * l + 3];
                const double x00 = -(a * (d00 * a + d01 * bq) + bq * (d01 * a + d11 * bq));
                const double x01 = -(a * (d00 * bq + d01 * c) + bq * (d01 * bq + d11 * c));
                const double x11 = -(bq * (d00 * bq + d01 * c) + c * (d01 * bq + d11 * c));
                dm[4 * l + 0] = x00; dm[4 * l + 1] = x01; dm[4 * l + 2] = x01; dm[4 * l + 3] = x11;
            }
            double* dcf = dcoef.data() + static_cast<size_t>(e) * 4 * Np;
            for (int mu = 0; mu < Np; ++mu) {
                const int l = S.l_mu[mu];
                for (int q = 0; q < 4; ++q) dcf[q * Np + mu] = dm[4 * l + q];
                for (int i = 0; i < n_fit; ++i) {
                    dJc[(static_cast<size_t>(e) * n_fit + i) * Np + mu] = djl[static_cast<size_t>(l) * n_fit + i];
                    dYc[(static_cast<size_t>(e) * n_fit + i) * Np + mu] = dyl[static_cast<size_t>(l) * n_fit + i];
                }
            }
        }
    }
    up(b.dA.p, dA, st); up(b.dS.p, dS, st); up(b.cE.p, cE, st); up(b.anl.p, anl, st);
    up(b.tab.p, tab, st); up(b.jy.p, jy, st); up(b.Jc.p, Jc, st); up(b.Yc.p, Yc, st); up(b.coef.p, coef, st);
    // tangent tables (kept alive until the stream is synchronised below)
    std::vector<double> wF, wP, wN, danl, dtab, djy;
    if (tan) {
        // Tangent tables.  E and alpha enter A_n = I + h^2/6 (E - V_n) only as
        // h^2/6 w_n I with w_n = t_E - t_alpha dV_pol/dalpha(n); a stability
        // shift at n rebuilds the matrix with a fixed smallest eigenvalue and
        // removes that dependence (sig = 0).  Per energy and n:
        //   wF = 12 sig h^2/6 w_n  (the derivative of 12 W^{-1}: -12 sig h^2/6 w G G^T
        //                           enters U, and d(U^{-1}) = -U^{-1} dU U^{-1})
        //   wP = -sig h^2/6 w_n    (dW^{-1} applied to Z: -sig h^2/6 w G psi)
        wF.assign(static_cast<size_t>(Nr) * nb, 0.0); wP.assign(wF.size(), 0.0); wN.assign(nb, 0.0);
        for (int n = 0; n < Nr; ++n) {
            const double pu = S.polarization_per_alpha(n);
            for (int e = 0; e < nb; ++e) {
                const size_t q = static_cast<size_t>(n) * nb + e;
                const bool shifted = n < s->nstab && (dA[q] != 0.0 || dS[q] != 0.0);
                const double w = dir[2 * e] - dir[2 * e + 1] * pu, sg = shifted ? 0.0 : 1.0;
                wF[q] = 12.0 * sg * s->c6 * w;
                wP[q] = -sg * s->c6 * w;
                if (n == N) wN[e] = s->c6 * w;
            }
        }
        // analytic start (n < n_start): derivative of the diagonal ratio
        // M_n j(k r_n) / (M_{n+1} j(k r_{n+1})); the f block is constant
        danl.assign(static_cast<size_t>(ns0) * nb * Nt, 0.0);
        for (int n = 1; n < ns0; ++n) {
            const double pn = S.polarization_per_alpha(n), pn1 = S.polarization_per_alpha(n + 1);
            for (int e = 0; e < nb; ++e) {
                double* d = danl.data() + (static_cast<size_t>(n) * nb + e) * Nt;
                const double k = ks[e], dk = dir[2 * e] / k, h2_6 = h * h / 6.0;
                const double wn = dir[2 * e] - dir[2 * e + 1] * pn, wn1 = dir[2 * e] - dir[2 * e + 1] * pn1;
                const double rn = h + n * h, rn1 = h + (n + 1) * h;
                for (int mu = 0; mu < Np; ++mu) {
                    const int l = S.l_mu[mu];
                    const double Mn = 1.0 + h2_6 * (E[e] - s->vdiag[static_cast<size_t>(n) * Np + mu]);
                    const double Mn1 = 1.0 + h2_6 * (E[e] - s->vdiag[static_cast<size_t>(n + 1) * Np + mu]);
                    const double yn = ric_j(l, k * rn), yn1 = ric_j(l, k * rn1);
                    const double dyn = ric_jp(l, k * rn) * rn * dk, dyn1 = ric_jp(l, k * rn1) * rn1 * dk;
                    const double a = Mn * yn, bq = Mn1 * yn1;
                    const double da = h2_6 * wn * yn + Mn * dyn, db = h2_6 * wn1 * yn1 + Mn1 * dyn1;
                    if (std::fabs(bq) > 1e-300) d[mu] = (da * bq - a * db) / (bq * bq);
                }
            }
        }
        // K matching and psi_N tables: d(W J) etc. with dW = h^2/12 2 k dk
        dtab.assign(static_cast<size_t>(4) * Np * nb, 0.0); djy.assign(static_cast<size_t>(2) * Np * nb, 0.0);
        for (int e = 0; e < nb; ++e) {
            const double k = ks[e], dk = dir[2 * e] / k, dW = h * h / 12.0 * 2.0 * k * dk;
            const double* t = tab.data() + static_cast<size_t>(e) * 6 * Np;
            double* d = dtab.data() + static_cast<size_t>(e) * 4 * Np;
            double* q = djy.data() + static_cast<size_t>(e) * 2 * Np;
            for (int mu = 0; mu < Np; ++mu) {
                const int l = S.l_mu[mu];
                const double dJin = ric_jp(l, k * r_in) * r_in * dk, dJout = ric_jp(l, k * r_out) * r_out * dk;
                const double dYin = ric_yp(l, k * r_in) * r_in * dk, dYout = ric_yp(l, k * r_out) * r_out * dk;
                d[mu] = dW * t[Np + mu] + t[5 * Np + mu] * dJout;
                d[Np + mu] = dW * t[mu] + t[4 * Np + mu] * dJin;
                d[2 * Np + mu] = dW * t[2 * Np + mu] + t[4 * Np + mu] * dYin;
                d[3 * Np + mu] = dW * t[3 * Np + mu] + t[5 * Np + mu] * dYout;
                q[mu] = ric_jp(l, k * rN) * rN * dk;
                q[Np + mu] = ric_yp(l, k * rN) * rN * dk;
            }
        }
        up(b.wF.p, wF, st); up(b.wP.p, wP, st); up(b.wN.p, wN, st); up(b.danl.p, danl, st);
        up(b.dtab.p, dtab, st); up(b.djy.p, djy, st); up(b.dJc.p, dJc, st); up(b.dYc.p, dYc, st);
        up(b.dcoef.p, dcoef, st);
    }
    // the side stream reads the per-energy tables (cE, shifts) just uploaded
    hipStream_t st2 = g.st2;
    SEHF_HIP_CHECK(hipEventRecord(g.ev_in, st));
    SEHF_HIP_CHECK(hipStreamWaitEvent(st2, g.ev_in, 0));
    // zero-pivot flags start clear for this batch (a failed earlier run on
    // the same buffers may have left them set)
    fill_int(b.gj.info.p, static_cast<size_t>(nb), 0, st);
    fill_int(b.gj2.info.p, static_cast<size_t>(nb) * sgrp, 0, st2);
    const int64_t sJc = static_cast<int64_t>(n_fit) * Np;

    auto gj = [&](double* mat, double* outp, int n, bool sym) {
        gj_inverse(mat, outp, n, n, static_cast<int64_t>(n) * n, nb, sym, b.gj.work(), st);
    };
    // Sinv_n on the side stream (its own Gauss-Jordan workspace): it needs
    // only the shared tables and this batch's energies, so it runs ahead of
    // the forward recursion that consumes it.
    // S_n is symmetric positive definite (the stability shift below n = 60,
    // A_n within a few percent of I beyond), so by default it is inverted in
    // place by block elimination without pivoting; SEHF_SINV_INVERSE=gj
    // selects the pivoted Gauss-Jordan inverse instead.
    // Streamed tables: the side stream rebuilds S0_n (and B_n D_n^{-1} into
    // bd_out, where the consumers of point n read it) before the Schur inverse.
    // A group of sgrp consecutive points is built point by point (sinv_build)
    // into consecutive slots and then inverted by one call (sinv_invert).
    const bool streamed = s->tables == SEHF_TABLES_STREAMED;
    auto sinv_build = [&](int n, double* Sout, double* bd_out) {
        const double* a = (n < s->nstab) ? b.dA.p + static_cast<size_t>(n) * nb : b.zeros.p;
        const double* d = (n < s->nstab) ? b.dS.p + static_cast<size_t>(n) * nb : b.zeros.p;
        double* Sm = s->sinv_spd ? Sout : b.Ms.p;
        const double* S0n = streamed ? nullptr : s->S0.p + static_cast<size_t>(n) * pp;
        if (streamed) {
            tables_at(s, b, n, bd_out, st2);
            S0n = b.tS0.p;
        }
        sinv_matrix(Sm, S0n, s->cvd.p + static_cast<size_t>(n) * Np,
                    s->bdb.p + static_cast<size_t>(n) * Np, b.cE.p, a, d, Np, nb, st2);
        if (!s->sinv_spd) gj_inverse(b.Ms.p, Sout, Np, Np, static_cast<int64_t>(pp), nb, true, b.gj2.work(), st2);
    };
    // cnt consecutive points' matrices from S (cnt x nb, stride N_psi^2)
    auto sinv_invert = [&](double* S, int cnt) {
        if (s->sinv_spd) spd_inverse(S, Np, Np, static_cast<int64_t>(pp), nb * cnt, true, b.gj2.work(), st2);
    };
    auto frp_step = [&](int n, const double* Sv, const double* Rprev, double* Rout, const double* BDn) {
        init_U(b.U.p, Rprev, Sv, s->Dinv.p + static_cast<size_t>(n) * std::max(Nf, 1), Np, Nt, nb, st);
        if (n < s->n_tr && Nf > 0) {
            Gemm g;   // Xn = -Sinv (B D^{-1})
            g.M = Np; g.N = Nf; g.K = Np;
            g.A = Sv; g.lda = Np; g.sA = pp;
            g.B = BDn; g.ldb = Np; g.sB = 0;
            g.C = Xd; g.ldc = Np; g.sC = sXd; g.alpha = -1.0; g.batch = nb;
            gemm(g, st);
            add_psi_f_blocks(b.U.p, Xd, sXd, 12.0, Np, Nf, Nt, nb, st);
            Gemm h2;  // U_ff -= 12 (B D^{-1})^T Xn
            h2.M = Nf; h2.N = Nf; h2.K = Np;
            h2.A = BDn; h2.lda = Np; h2.sA = 0; h2.tA = 1;
            h2.B = Xd; h2.ldb = Np; h2.sB = sXd;
            h2.C = b.U.p + Np + static_cast<size_t>(Np) * Nt; h2.ldc = Nt; h2.sC = tt;
            h2.alpha = -12.0; h2.beta = 1.0; h2.batch = nb;
            gemm(h2, st);
        }
        gj(b.U.p, Rout, Nt, true);
    };
    // Tangent of one ratio step (after frp_step(n) has left Rinv_n in Rn and
    // -Sinv B D^{-1} in Gt):  with G = W_n^{-1}[:, psi] = [Sinv; -D^{-1} B^T Sinv],
    //   dRinv_n = Rinv_n (dRinv_{n-1} + wF_n G G^T) Rinv_n,  symmetrised.
    // H holds dRinv_{n-1} on entry and is overwritten; U is free scratch.
    auto frp_tangent = [&](int n, const double* Sv, const double* Rn, double* H, double* dRout) {
        const int64_t sG = static_cast<int64_t>(Np) * Nt;
        axpbyi(b.Gt.p, Np, sG, Sv, Np, pp, 1.0, nullptr, 0, 0, 0.0, 0.0, Np, Np, nb, st);   // Gt[:, psi] = Sinv
        const bool ex = n < s->n_tr && Nf > 0;
        Gemm g;   // H += wF_n G G^T  (psi-psi block only when G_f = 0)
        g.M = ex ? Nt : Np; g.N = ex ? Nt : Np; g.K = Np;
        g.A = b.Gt.p; g.lda = Np; g.sA = sG; g.tA = 1;
        g.B = b.Gt.p; g.ldb = Np; g.sB = sG;
        g.C = H; g.ldc = Nt; g.sC = tt; g.beta = 1.0; g.alpha_b = b.wF.p + static_cast<size_t>(n) * nb; g.batch = nb;
        gemm(g, st);
        Gemm g2;  // U = H Rinv_n
        g2.M = Nt; g2.N = Nt; g2.K = Nt;
        g2.A = H; g2.lda = Nt; g2.sA = tt; g2.B = Rn; g2.ldb = Nt; g2.sB = tt;
        g2.C = b.U.p; g2.ldc = Nt; g2.sC = tt; g2.batch = nb;
        gemm(g2, st);
        Gemm g3;  // dRinv_n = Rinv_n U
        g3.M = Nt; g3.N = Nt; g3.K = Nt;
        g3.A = Rn; g3.lda = Nt; g3.sA = tt; g3.B = b.U.p; g3.ldb = Nt; g3.sB = tt;
        g3.C = dRout; g3.ldc = Nt; g3.sC = tt; g3.batch = nb;
        gemm(g3, st);
        symmetrize(dRout, Nt, nb, st);
    };
    SEHF_HIP_CHECK(hipStreamSynchronize(st));
    const double tw1 = sync ? sync->arrive(wall()) : wall();
    // Optional progress lines on stderr (SEHF_PROGRESS=<seconds>): the stream
    // is drained before each line so the reported radial index has finished.
    static const double prog_every = [] {
        const char* e = std::getenv("SEHF_PROGRESS");
        return e ? std::atof(e) : 0.0;
    }();
    double prog_last = tw1;
    auto progress = [&](const char* what, int n, int total) {
        if (prog_every <= 0.0 || wall() - prog_last < prog_every) return;
        SEHF_HIP_CHECK(hipStreamSynchronize(st));
        prog_last = wall();
        std::fprintf(stderr, "[sehf] batch of %d: %s %d/%d (%.1f s into batch)\n", nb, what, n, total, prog_last - tw0);
        std::fflush(stderr);
    };
    // ---------------- forward sweep ----------------
    if (ns0 >= 2) diag_matrix(b.Rp.p, b.anl.p + static_cast<size_t>(ns0 - 1) * nb * Nt, Nt, nb, st);
    else fill(b.Rp.p, tt * nb, 0.0, st);
    if (tan) {
        if (ns0 >= 2) diag_matrix(b.dRp.p, b.danl.p + static_cast<size_t>(ns0 - 1) * nb * Nt, Nt, nb, st);
        else fill(b.dRp.p, tt * nb, 0.0, st);
    }
    // The Schur inverses go to a ring of two halves of sgrp points each: the
    // side stream inverts group q + 1 while the main stream runs the ratio
    // steps of group q; FRP(n) waits for its group's inverses.
    // A segment cycle's forward part starts at the segment's first point from
    // the analytic start matrix (a ratio matrix of the right size and
    // conditioning; the cycle's results are not kept).
    const int n_first = seg >= 0 ? std::max(ns0, seg * L) : ns0;
    // Checkpoint sg (the start matrix Rinv_{sg L - 1} of segment sg) for
    // sg >= ck_lo; a segment cycle keeps its one start matrix in slot 0.
    const int ck_lo = ckpt_lo(s, L, tan);
    auto ckp = [&](double* base, int sg) {
        const int i = seg >= 0 ? 0 : sg - ck_lo;
        return base + static_cast<size_t>(i) * nb * tt;
    };
    // Forward accumulation of the dipole integrals (plain runs).  psi_n = P_n Z_n
    // with P_n = [Sinv_n | X_n] (X_n = -Sinv_n B D^{-1}, zero from n_tr on) and
    // Z_n = Rinv_n ... Rinv_{N-1} Z_N, so sum_{n<N} rows_n psi_n = Ad_{N-1} Z_N
    // with Ad_n = (Ad_{n-1} + rows_n P_n) Rinv_n, Ad_{-1} = 0 (Rinv_0 = 0 makes
    // Ad_0 = 0); Ad is (6 + n_occ) x N_tot per energy (ld ns).  Tangent runs
    // back-propagate instead: the tangent of these sums, dAd_n = (dAd_{n-1} +
    // d(rows_n P_n) + Ad_n H_n) Rinv_n, carries Rinv_n twice where a ratio
    // matrix passes near a pole (|dRinv_n| ~ |Rinv_n|^2), while the
    // back-propagated dZ_n = Rinv_n (dZ_{n+1} + H_n Z_n) carries it once.
    // Below n_start the ratio matrices are the analytic diagonal; the dipole
    // rows there still need those points' Sinv_n, so the loop starts at n = 1.
    const bool adj = !back_propagated_dipoles(s, tan) && !replay;
    const int n_lo = (adj && seg < 0 && probe == 0) ? std::min(1, n_first) : n_first;
    // B_n D_n^{-1} of the forward sweep: resident, or the ring slot of point n
    const int n_last = replay ? -1   // a replay solves nothing
                              : probe > 0 ? std::min(Nr - 2, ns0 + probe - 1)
                                          : (seg >= 0 ? std::min(seg * L + L, N) - 1 : Nr - 2);
    // point n is entry (n - n_lo) % sgrp of group q = (n - n_lo) / sgrp, in
    // ring half q & 1
    auto ring_idx = [&](int n) {
        const int q = (n - n_lo) / sgrp;
        return static_cast<size_t>((q & 1) * sgrp + (n - n_lo) % sgrp);
    };
    auto bd_fwd = [&](int n) -> double* {
        return streamed ? b.ringBD.p + ring_idx(n) * pf : s->BD.p + static_cast<size_t>(n) * pf;
    };
    auto sinv_slot = [&](int n) { return b.SinvR.p + ring_idx(n) * pp * nb; };
    // group q's Schur inverses, after the ratio steps of group q - 2 (the
    // previous users of its ring half) are done
    auto sinv_ahead = [&](int q) {
        const int a = n_lo + q * sgrp, e = std::min(a + sgrp, n_last + 1);
        if (q >= 2) SEHF_HIP_CHECK(hipStreamWaitEvent(st2, g.ev_f[q & 1], 0));
        for (int n = a; n < e; ++n) sinv_build(n, sinv_slot(n), bd_fwd(n));
        sinv_invert(sinv_slot(a), e - a);
        SEHF_HIP_CHECK(hipEventRecord(g.ev_s[q & 1], st2));
    };
    const int64_t sAd = static_cast<int64_t>(ns) * Nt;
    auto gm = [&](double* C, int64_t ldc, int64_t sC, const double* A, int64_t lda, int64_t sA, const double* Bm,
                  int64_t ldb, int64_t sB, int M, int Nn, int K, double beta) {
        Gemm q;
        q.M = M; q.N = Nn; q.K = K;
        q.A = A; q.lda = lda; q.sA = sA; q.B = Bm; q.ldb = ldb; q.sB = sB;
        q.C = C; q.ldc = ldc; q.sC = sC; q.beta = beta; q.batch = nb;
        gemm(q, st);
    };
    // Ad_{n-1} += rows_n P_n, then Ad_n = Ad_{n-1} Rinv_n
    auto adj_step = [&](int n, const double* Sv, const double* Rn) {
        if (n < s->n_pts) {
            const double* rows = s->Drows.p + static_cast<size_t>(n) * ns * Np;
            gm(b.Ad.p, ns, sAd, rows, ns, 0, Sv, Np, pp, ns, Np, Np, 1.0);
            if (n < s->n_tr && Nf > 0)
                gm(b.Ad.p + static_cast<size_t>(ns) * Np, ns, sAd, rows, ns, 0, Xd, Np, sXd, ns, Nf, Np, 1.0);
        }
        gm(b.Ad2.p, ns, sAd, b.Ad.p, ns, sAd, Rn, Nt, tt, ns, Nt, Nt, 0.0);
        std::swap(b.Ad.p, b.Ad2.p);
    };
    if (adj) fill(b.Ad.p, static_cast<size_t>(sAd) * nb, 0.0, st);
    if (n_lo <= n_last) sinv_ahead(0);
    for (int n = n_lo; n <= n_last; ++n) {
        const int q = (n - n_lo) / sgrp, slot = q & 1;
        const bool first = (n - n_lo) % sgrp == 0, last = (n - n_lo) % sgrp == sgrp - 1 || n == n_last;
        if (first) {
            if (n_lo + (q + 1) * sgrp <= n_last) sinv_ahead(q + 1);
            SEHF_HIP_CHECK(hipStreamWaitEvent(st, g.ev_s[slot], 0));
        }
        const double* Sslot = sinv_slot(n);
        if (n < ns0) {
            // analytic region (forward-accumulated dipoles only): Rinv_n =
            // diag(anl_n) built in Rc; Rp holds the start matrix Rinv_{n_start-1}
            if (n < s->n_tr && Nf > 0) {
                Gemm g;   // Xn = -Sinv (B D^{-1})
                g.M = Np; g.N = Nf; g.K = Np;
                g.A = Sslot; g.lda = Np; g.sA = pp;
                g.B = bd_fwd(n); g.ldb = Np; g.sB = 0;
                g.C = Xd; g.ldc = Np; g.sC = sXd; g.alpha = -1.0; g.batch = nb;
                gemm(g, st);
            }
            diag_matrix(b.Rc.p, b.anl.p + static_cast<size_t>(n) * nb * Nt, Nt, nb, st);
            adj_step(n, Sslot, b.Rc.p);
            if (last) SEHF_HIP_CHECK(hipEventRecord(g.ev_f[slot], st));
            continue;
        }
        frp_step(n, Sslot, b.Rp.p, b.Rc.p, bd_fwd(n));
        if (adj) adj_step(n, Sslot, b.Rc.p);
        if (tan) frp_tangent(n, Sslot, b.Rc.p, b.dRp.p, b.dRc.p);
        if (last) SEHF_HIP_CHECK(hipEventRecord(g.ev_f[slot], st));
        std::swap(b.Rp.p, b.Rc.p);
        if (tan) std::swap(b.dRp.p, b.dRc.p);
        progress("forward sweep n =", n, Nr - 2);
        if ((n + 1) % L == 0) {
            const int sidx = (n + 1) / L;
            if (sidx < n_seg && sidx >= ck_lo) {
                // (checkpoints may be in pinned host memory: hipMemcpyDefault)
                SEHF_HIP_CHECK(hipMemcpyAsync(ckp(b.ckpt.p, sidx), b.Rp.p, nb * tt * sizeof(double),
                                              hipMemcpyDefault, st));
                if (tan)
                    SEHF_HIP_CHECK(hipMemcpyAsync(ckp(b.dckpt.p, sidx), b.dRp.p, nb * tt * sizeof(double),
                                                  hipMemcpyDefault, st));
            }
        }
    }
    if (probe > 0) {
        SEHF_HIP_CHECK(hipStreamSynchronize(st2));
        SEHF_HIP_CHECK(hipStreamSynchronize(st));
        g.t_probe = wall() - tw1;
        return;
    }
    check_gj_info(b.gj, nb, st, "scan forward sweep");
    check_gj_info(b.gj2, nb * sgrp, st2, "scan forward sweep (Schur inverses)");
    const double tw2 = wall();
    if (seg_t) { seg_t[0] = tw1 - tw0; seg_t[1] = tw2 - tw1; }
    // ---------------- K matrix (KMatrixExtractor.cpp:75-198) ----------------
    if (seg < 0 && !replay) {
    SEHF_HIP_CHECK(hipMemcpyAsync(b.Rc.p, b.Rp.p, nb * tt * sizeof(double), hipMemcpyDeviceToDevice, st));
    gj(b.Rc.p, b.U.p, Nt, false);   // U <- R = Rinv_{N-1}^{-1}
    get_block(b.Am.p, b.U.p, Nt, tt, 0, 0, Np, Np, nb, st);   // S_schur starts as R_psi_psi
    bool schur = false;
    if (Nf > 0) {
        fill(b.nrm.p, 2 * nb, 0.0, st);
        sumsq_block(b.U.p, Nt, tt, 0, Np, Np, Nf, b.nrm.p, nb, st);
        sumsq_block(b.U.p, Nt, tt, Np, 0, Nf, Np, b.nrm.p + nb, nb, st);
        std::vector<double> nr(2 * nb);
        SEHF_HIP_CHECK(hipMemcpyAsync(nr.data(), b.nrm.p, 2 * nb * sizeof(double), hipMemcpyDeviceToHost, st));
        SEHF_HIP_CHECK(hipStreamSynchronize(st));
        for (int e = 0; e < nb; ++e) if (!(std::sqrt(nr[e]) < 1e-14 && std::sqrt(nr[nb + e]) < 1e-14)) schur = true;
    }
    if (schur) {
        get_block(b.Rff.p, b.U.p, Nt, tt, Np, Np, Nf, Nf, nb, st);
        gj(b.Rff.p, b.Rffi.p, Nf, false);
        Gemm g;   // T1 = R_psi_f R_ff^{-1}
        g.M = Np; g.N = Nf; g.K = Nf;
        g.A = b.U.p + static_cast<size_t>(Np) * Nt; g.lda = Nt; g.sA = tt;
        g.B = b.Rffi.p; g.ldb = Nf; g.sB = static_cast<int64_t>(Nf) * Nf;
        g.C = b.T1.p; g.ldc = Np; g.sC = pf; g.batch = nb;
        gemm(g, st);
        Gemm g2;  // S -= T1 R_f_psi
        g2.M = Np; g2.N = Np; g2.K = Nf;
        g2.A = b.T1.p; g2.lda = Np; g2.sA = pf;
        g2.B = b.U.p + Np; g2.ldb = Nt; g2.sB = tt;
        g2.C = b.Am.p; g2.ldc = Np; g2.sC = pp; g2.alpha = -1.0; g2.beta = 1.0; g2.batch = nb;
        gemm(g2, st);
    }
    symmetrize(b.Am.p, Np, nb, st);
    SEHF_HIP_CHECK(hipMemcpyAsync(b.T.p, b.Am.p, nb * pp * sizeof(double), hipMemcpyDeviceToDevice, st));
    kme_mats(b.Am.p, b.Bm.p, b.T.p, b.tab.p, Np, nb, st);
    gj(b.Bm.p, b.Psi.p, Np, false);          // Psi <- Bmat^{-1}
    {
        Gemm g;   // K = Bmat^{-1} Amat
        g.M = Np; g.N = Np; g.K = Np;
        g.A = b.Psi.p; g.lda = Np; g.sA = pp; g.B = b.Am.p; g.ldb = Np; g.sB = pp;
        g.C = b.Kd.p; g.ldc = Np; g.sC = pp; g.batch = nb;
        gemm(g, st);
    }
    symmetrize(b.Kd.p, Np, nb, st);
    if (tan) {
        // S_schur = ((Rinv_{N-1})_psi,psi)^{-1}, so dS = -S (dRinv_{N-1})_psi,psi S;
        // dK = sym(Bmat^{-1} (dAmat - dBmat K))
        get_block(b.Q1.p, b.dRp.p, Nt, tt, 0, 0, Np, Np, nb, st);
        Gemm g;   // Q2 = S dRinv_psi,psi
        g.M = Np; g.N = Np; g.K = Np;
        g.A = b.T.p; g.lda = Np; g.sA = pp; g.B = b.Q1.p; g.ldb = Np; g.sB = pp;
        g.C = b.Q2.p; g.ldc = Np; g.sC = pp; g.batch = nb;
        gemm(g, st);
        Gemm g2;  // Q1 = dS = -Q2 S
        g2.M = Np; g2.N = Np; g2.K = Np;
        g2.A = b.Q2.p; g2.lda = Np; g2.sA = pp; g2.B = b.T.p; g2.ldb = Np; g2.sB = pp;
        g2.C = b.Q1.p; g2.ldc = Np; g2.sC = pp; g2.alpha = -1.0; g2.batch = nb;
        gemm(g2, st);
        kme_dmats(b.dAm.p, b.dBm.p, b.Q1.p, b.T.p, b.tab.p, b.dtab.p, Np, nb, st);
        Gemm g3;  // dAmat -= dBmat K
        g3.M = Np; g3.N = Np; g3.K = Np;
        g3.A = b.dBm.p; g3.lda = Np; g3.sA = pp; g3.B = b.Kd.p; g3.ldb = Np; g3.sB = pp;
        g3.C = b.dAm.p; g3.ldc = Np; g3.sC = pp; g3.alpha = -1.0; g3.beta = 1.0; g3.batch = nb;
        gemm(g3, st);
        Gemm g4;  // dK = Bmat^{-1} (dAmat - dBmat K)
        g4.M = Np; g4.N = Np; g4.K = Np;
        g4.A = b.Psi.p; g4.lda = Np; g4.sA = pp; g4.B = b.dAm.p; g4.ldb = Np; g4.sB = pp;
        g4.C = b.dKd.p; g4.ldc = Np; g4.sC = pp; g4.batch = nb;
        gemm(g4, st);
        symmetrize(b.dKd.p, Np, nb, st);
    }
    // ---------------- outer boundary (BackPropagator.cpp:128-197, :538-551) ----------------
    psi_boundary(b.P.p, b.Kd.p, b.jy.p, Np, nb, st);
    psi_out(0, N, b.P.p);
    psi_out(1, N, b.P.p);
    }   // seg < 0
    // Z_psi = psi + h^2/6 (E psi - V_N psi) + B_N f_N with f_N = -D_N^{-1} B_N^T psi, Z_f = 0
    auto zmap = [&](const double* Pin, double* Zout) {
        Gemm g;   // Vpsi = V_N psi_N
        g.M = Np; g.N = Np; g.K = Np;
        g.A = s->VN.p; g.lda = Np; g.sA = 0; g.B = Pin; g.ldb = Np; g.sB = pp;
        g.C = b.Vpsi.p; g.ldc = Np; g.sC = pp; g.batch = nb;
        gemm(g, st);
        if (Nf > 0 && N < s->n_tr) {
            Gemm g1;  // G = B_N^T psi_N
            g1.M = Nf; g1.N = Np; g1.K = Np;
            g1.A = s->BN.p; g1.lda = Np; g1.sA = 0; g1.tA = 1; g1.B = Pin; g1.ldb = Np; g1.sB = pp;
            g1.C = b.G.p; g1.ldc = Nf; g1.sC = pf; g1.batch = nb;
            gemm(g1, st);
            row_scale(b.G.p, b.G.p, s->negDinvN.p, Nf, Np, nb, st);   // f_N = -D^{-1} B^T psi
            Gemm g2;  // Bf = B_N f_N
            g2.M = Np; g2.N = Np; g2.K = Nf;
            g2.A = s->BN.p; g2.lda = Np; g2.sA = 0; g2.B = b.G.p; g2.ldb = Nf; g2.sB = pf;
            g2.C = b.Bf.p; g2.ldc = Np; g2.sC = pp; g2.batch = nb;
            gemm(g2, st);
        } else {
            fill(b.Bf.p, pp * nb, 0.0, st);
        }
        z_boundary(Zout, Pin, b.Vpsi.p, b.Bf.p, b.cE.p, s->c6, Np, Nt, nb, st);
    };
    if (seg < 0 && !replay) zmap(b.P.p, b.Z.p);
    if (tan && seg < 0) {
        // d psi_N from dK and the tables; dZ = zmap(d psi_N) + h^2/6 w_N psi_N
        dpsi_boundary(b.dP.p, b.Kd.p, b.dKd.p, b.jy.p, b.djy.p, Np, nb, st);
        zmap(b.dP.p, b.dZ.p);
        axpy_b(b.dZ.p, Nt, static_cast<int64_t>(Nt) * Np, b.P.p, Np, pp, b.wN.p, 1.0, Np, Np, nb, st);
    }
    fill(b.acc.p, static_cast<size_t>(ns) * Np * nb, 0.0, st);
    fill(b.YJ.p, pp * nb, 0.0, st);
    fill(b.YY.p, pp * nb, 0.0, st);
    auto dip_accum = [&](int n, const double* psi, double* acc) {
        Gemm g;   // acc += rows_n psi_n
        g.M = ns; g.N = Np; g.K = Np;
        g.A = s->Drows.p + static_cast<size_t>(n) * ns * Np; g.lda = ns; g.sA = 0;
        g.B = psi; g.ldb = Np; g.sB = pp;
        g.C = acc; g.ldc = ns; g.sC = static_cast<int64_t>(ns) * Np;
        g.beta = 1.0; g.batch = nb;
        gemm(g, st);
    };
    auto fit_tangent = [&](int n, const double* psi, const double* dpsi) {
        const size_t o = static_cast<size_t>(n - n_fs) * Np;
        fit_accum_t(b.dYJ.p, b.dYY.p, psi, dpsi, b.Jc.p + o, b.Yc.p + o, b.dJc.p + o, b.dYc.p + o, sJc, Np, nb, st);
    };
    if (N < s->n_pts && seg < 0) {
        if (replay) psi_in(N, b.P.p);
        dip_accum(N, b.P.p, b.acc.p);
    }
    // The fit window [n_fs, n_fe] ends 3 points below the boundary N.  Its
    // sums YJ, YY come from a pre-pass over the top segments (below), so the
    // main pass knows A and B and takes the fit residual point by point; no
    // copy of the window's psi is kept.
    auto fit_sums = [&](int n, const double* psi) {
        const size_t o = static_cast<size_t>(n - n_fs) * Np;
        fit_accum(b.YJ.p, b.YY.p, nullptr, 0, psi, b.Jc.p + o, b.Yc.p + o, sJc, Np, nb, st);
    };
    auto fit_resid_at = [&](int n, const double* psi) {
        const size_t o = static_cast<size_t>(n - n_fs) * Np;
        fit_resid(psi, b.Am.p, b.Bm.p, b.Jc.p + o, b.Yc.p + o, b.coef.p, Np, 1, static_cast<int64_t>(pp), sJc, b.red.p,
                  nb, st);
    };
    if (tan) {
        fill(b.dacc.p, static_cast<size_t>(ns) * Np * nb, 0.0, st);
        fill(b.dYJ.p, pp * nb, 0.0, st);
        fill(b.dYY.p, pp * nb, 0.0, st);
        if (N < s->n_pts && seg < 0) dip_accum(N, b.dP.p, b.dacc.p);
        if (N >= n_fs && N <= n_fe && seg < 0) fit_tangent(N, b.P.p, b.dP.p);
    }
    if (adj && seg < 0)   // the forward-accumulated rows applied to the outer boundary: Ad_{N-1} Z_N
        gm(b.acc.p, ns, static_cast<int64_t>(ns) * Np, b.Ad.p, ns, sAd, b.Z.p, Nt, static_cast<int64_t>(Nt) * Np, ns,
           Np, Nt, 1.0);
    fill(b.red.p, 2 * nb, 0.0, st);
    SEHF_HIP_CHECK(hipStreamSynchronize(st));
    const double tw3 = wall();
    // psi-block of W_n^{-1} Zin (WInverseOperator.cpp:75-128):
    //   out = Sinv_n (Z_psi - (B D^{-1}) Z_f)
    const int64_t sZ = static_cast<int64_t>(Nt) * Np;
    auto winv_psi = [&](int n, const double* Zin, const double* Sn, double* outp, const double* BDn) {
        get_block(b.T.p, Zin, Nt, sZ, 0, 0, Np, Np, nb, st);
        if (n < s->n_tr && Nf > 0) {
            Gemm g2;  // rhs = Z_psi - (B D^{-1}) Z_f
            g2.M = Np; g2.N = Np; g2.K = Nf;
            g2.A = BDn; g2.lda = Np; g2.sA = 0;
            g2.B = Zin + Np; g2.ldb = Nt; g2.sB = sZ;
            g2.C = b.T.p; g2.ldc = Np; g2.sC = pp; g2.alpha = -1.0; g2.beta = 1.0; g2.batch = nb;
            gemm(g2, st);
        }
        Gemm g3;  // out = Sinv_n rhs
        g3.M = Np; g3.N = Np; g3.K = Np;
        g3.A = Sn; g3.lda = Np; g3.sA = pp; g3.B = b.T.p; g3.ldb = Np; g3.sB = pp;
        g3.C = outp; g3.ldc = Np; g3.sC = pp; g3.batch = nb;
        gemm(g3, st);
    };
    auto mm = [&](double* C, const double* A, const double* Bm, int M, int Nn, int K, int64_t lda, int64_t sA,
                  int64_t ldb, int64_t sB, int64_t ldc, int64_t sC, double beta) {
        Gemm g;
        g.M = M; g.N = Nn; g.K = K;
        g.A = A; g.lda = lda; g.sA = sA; g.B = Bm; g.ldb = ldb; g.sB = sB;
        g.C = C; g.ldc = ldc; g.sC = sC; g.beta = beta; g.batch = nb;
        gemm(g, st);
    };
    // ---------------- back-propagation by segments ----------------
    // Three kinds of pass, each over segments sg_hi down to sg_lo with every
    // segment's ratio matrices recomputed from its checkpoint:
    //   FitSums: from the top down to the fit window's first point n_fs, psi
    //     into the window sums;
    //   Resid:   the same points again once A, B are known, the fit residual;
    //   Full:    the whole grid with the dipole sums, the fit residual and the
    //     tangent (back-propagated dipoles; a segment cycle's pass).
    // A segment cycle runs a Full pass over its one segment (sg_hi = sg_lo)
    // and, with t_rec set, records when the segment's recompute has finished.
    enum class Pass { FitSums, Resid, Full };
    // the fit pre-pass also covers the psi tail when that is written to file
    const int fs_lo = (pw && pw->wants(0, plo.tail_lo)) ? std::min(n_fs, plo.tail_lo) : n_fs;
    bool seg_pending = false;   // a previous segment's backward products may still read segS
    double t_rec = 0.0;
    auto sweep = [&](Pass kind, int sg_hi, int sg_lo, bool time_rec) {
        const bool main = kind == Pass::Full;
        for (int sg = sg_hi; sg >= sg_lo; --sg) {
            const int n0 = sg * L, n1 = std::min(n0 + L, N);
            const double* prev = (sg == 0) ? nullptr : ckp(b.ckpt.p, sg);
            const bool tg = tan && main;
            const double* dprev0 = (sg == 0 || !tg) ? nullptr : ckp(b.dckpt.p, sg);
            const double* dprev = dprev0;
            // the segment's Schur inverses on the side stream, after the
            // previous segment's backward products are done with segS.  They
            // are enqueued a few points ahead of the recompute that consumes
            // them, so the host starts the main stream's ratio steps after a
            // handful of inverses instead of the whole segment's (each is a
            // few hundred launches).
            if (seg_pending) SEHF_HIP_CHECK(hipStreamWaitEvent(st2, g.ev_bwd, 0));
            // B_n D_n^{-1} of the segment's points: resident, or rebuilt with
            // the segment's Schur inverses into segBD
            auto bd_seg = [&](int n) -> double* {
                return streamed ? b.segBD.p + static_cast<size_t>(n - n0) * pf : s->BD.p + static_cast<size_t>(n) * pf;
            };
            constexpr int kAhead = 4;
            int n_enq = n0;
            auto enqueue_sinv_to = [&](int n_hi) {
                while (n_enq < std::min(n_hi, n1)) {   // whole groups of up to sgrp points
                    const int a = n_enq, e = std::min(a + sgrp, n1);
                    for (int n = a; n < e; ++n) sinv_build(n, b.segS.p + static_cast<size_t>(n - n0) * nb * pp, bd_seg(n));
                    sinv_invert(b.segS.p + static_cast<size_t>(a - n0) * nb * pp, e - a);
                    for (int n = a; n < e; ++n) SEHF_HIP_CHECK(hipEventRecord(g.ev_seg[n - n0], st2));
                    n_enq = e;
                }
            };
            enqueue_sinv_to(n0 + kAhead);
            for (int n = n0; n < n1; ++n) {
                enqueue_sinv_to(n + 1 + kAhead);
                double* Rn = b.segR.p + static_cast<size_t>(n - n0) * nb * tt;
                double* Sn = b.segS.p + static_cast<size_t>(n - n0) * nb * pp;
                double* dRn = tg ? b.dsegR.p + static_cast<size_t>(n - n0) * nb * tt : nullptr;
                SEHF_HIP_CHECK(hipStreamWaitEvent(st, g.ev_seg[n - n0], 0));
                if (n >= ns0) {
                    frp_step(n, Sn, prev, Rn, bd_seg(n));
                    if (tg) {   // dRinv_{n-1} is kept: the tangent works on a copy
                        SEHF_HIP_CHECK(hipMemcpyAsync(b.dRp.p, dprev, nb * tt * sizeof(double), hipMemcpyDefault, st));
                        frp_tangent(n, Sn, Rn, b.dRp.p, dRn);
                    }
                } else if (n == 0) {
                    fill(Rn, tt * nb, 0.0, st);
                    if (tg) fill(dRn, tt * nb, 0.0, st);
                } else {
                    diag_matrix(Rn, b.anl.p + static_cast<size_t>(n) * nb * Nt, Nt, nb, st);
                    if (tg) diag_matrix(dRn, b.danl.p + static_cast<size_t>(n) * nb * Nt, Nt, nb, st);
                }
                prev = Rn;
                dprev = dRn;
            }
            if (time_rec) {
                SEHF_HIP_CHECK(hipStreamSynchronize(st2));
                SEHF_HIP_CHECK(hipStreamSynchronize(st));
                t_rec = wall();
            }
            progress(main ? "back-propagation segment" : kind == Pass::FitSums ? "fit pre-pass segment"
                                                                                : "fit residual segment",
                     sg, n_seg);
            const int n_stop = main ? n0 : std::max(n0, kind == Pass::FitSums ? fs_lo : n_fs);
            for (int n = n1 - 1; n >= n_stop; --n) {
                const double* Rn = b.segR.p + static_cast<size_t>(n - n0) * nb * tt;
                const double* Sn = b.segS.p + static_cast<size_t>(n - n0) * nb * pp;
                mm(b.Zt.p, Rn, b.Z.p, Nt, Np, Nt, Nt, tt, Nt, sZ, Nt, sZ, 0.0);   // Z <- Rinv_n Z
                std::swap(b.Z.p, b.Zt.p);
                winv_psi(n, b.Z.p, Sn, b.Psi.p, bd_seg(n));                          // psi_n
                psi_out(kind == Pass::FitSums ? 0 : main ? 1 : -1, n, b.Psi.p);
                const bool in_win = n >= n_fs && n <= n_fe;
                if (main && n < s->n_pts) dip_accum(n, b.Psi.p, b.acc.p);
                if (in_win) {
                    if (kind == Pass::FitSums) fit_sums(n, b.Psi.p);
                    else fit_resid_at(n, b.Psi.p);
                }
                if (!tg) continue;
                // dZ_n = dRinv_n Z_{n+1} + Rinv_n dZ_{n+1}.  For n >= n_start
                // dRinv_n = Rinv_n H_n Rinv_n, H_n = dRinv_{n-1} + wF_n G G^T, and
                // G^T Z_n = psi_n, so
                //   dZ_n = Rinv_n (dZ_{n+1} + dRinv_{n-1} Z_n + wF_n G psi_n)
                // with G psi_n = [P1; -D^{-1} B^T P1], P1 = Sinv_n psi_n.  This never
                // multiplies by dRinv_n itself, which grows like 1/delta^2 where
                // Rinv_n passes near a pole.
                mm(b.P1.p, Sn, b.Psi.p, Np, Np, Np, Np, pp, Np, pp, Np, pp, 0.0);
                const double* wFn = b.wF.p + static_cast<size_t>(n) * nb;
                if (n >= ns0) {
                    const double* dRm1 = (n == n0) ? dprev0 : b.dsegR.p + static_cast<size_t>(n - 1 - n0) * nb * tt;
                    mm(b.dZ.p, dRm1, b.Z.p, Nt, Np, Nt, Nt, tt, Nt, sZ, Nt, sZ, 1.0);
                    axpy_b(b.dZ.p, Nt, sZ, b.P1.p, Np, pp, wFn, 1.0, Np, Np, nb, st);
                    if (n < s->n_tr && Nf > 0) {
                        Gemm g;   // dZ_f -= wF_n (B D^{-1})^T P1
                        g.M = Nf; g.N = Np; g.K = Np;
                        g.A = bd_seg(n); g.lda = Np; g.sA = 0; g.tA = 1;
                        g.B = b.P1.p; g.ldb = Np; g.sB = pp;
                        g.C = b.dZ.p + Np; g.ldc = Nt; g.sC = sZ;
                        g.alpha = -1.0; g.alpha_b = wFn; g.beta = 1.0; g.batch = nb;
                        gemm(g, st);
                    }
                    mm(b.dZt.p, Rn, b.dZ.p, Nt, Np, Nt, Nt, tt, Nt, sZ, Nt, sZ, 0.0);
                } else {
                    const double* dRn = b.dsegR.p + static_cast<size_t>(n - n0) * nb * tt;
                    mm(b.dZt.p, Rn, b.dZ.p, Nt, Np, Nt, Nt, tt, Nt, sZ, Nt, sZ, 0.0);
                    mm(b.dZt.p, dRn, b.Zt.p, Nt, Np, Nt, Nt, tt, Nt, sZ, Nt, sZ, 1.0);   // Zt holds Z_{n+1}
                }
                std::swap(b.dZ.p, b.dZt.p);
                // dpsi_n = (W_n^{-1} dZ_n)_psi + wP_n Sinv_n psi_n
                winv_psi(n, b.dZ.p, Sn, b.dPsi.p, bd_seg(n));
                axpy_b(b.dPsi.p, Np, pp, b.P1.p, Np, pp, b.wP.p + static_cast<size_t>(n) * nb, 1.0, Np, Np, nb, st);
                if (main && n < s->n_pts) dip_accum(n, b.dPsi.p, b.dacc.p);
                if (in_win) fit_tangent(n, b.Psi.p, b.dPsi.p);
            }
            SEHF_HIP_CHECK(hipEventRecord(g.ev_bwd, st));   // segS free for the next segment's Sinv
            seg_pending = true;
        }
    };
    if (seg >= 0) {
        // Segment cycle: the recompute starts from the same start matrix as
        // the forward part, and the back-propagation from Z = [I; 0] (the
        // boundary of a whole scan needs its K matrix); the fit residual uses
        // A = I, B = 0.  The products, inverses and fused dipole/fit work are
        // those of the scan's main pass over this segment.
        const int64_t sZs = static_cast<int64_t>(Nt) * Np;
        if (seg > 0) {
            double* c0 = ckp(b.ckpt.p, seg);
            if (ns0 >= 2) diag_matrix(c0, b.anl.p + static_cast<size_t>(ns0 - 1) * nb * Nt, Nt, nb, st);
            else fill(c0, tt * nb, 0.0, st);
            if (tan) fill(ckp(b.dckpt.p, seg), tt * nb, 0.0, st);
        }
        fill(b.Z.p, sZs * nb, 0.0, st);
        axpbyi(b.Z.p, Nt, sZs, b.Z.p, Nt, sZs, 0.0, nullptr, 0, 0, 0.0, 1.0, Np, Np, nb, st);
        if (tan) fill(b.dZ.p, sZs * nb, 0.0, st);
        fill(b.Am.p, pp * nb, 0.0, st);
        axpbyi(b.Am.p, Np, pp, b.Am.p, Np, pp, 0.0, nullptr, 0, 0, 0.0, 1.0, Np, Np, nb, st);
        fill(b.Bm.p, pp * nb, 0.0, st);
        SEHF_HIP_CHECK(hipStreamSynchronize(st));
        const double tb0 = wall();
        sweep(Pass::Full, seg, seg, true);
        SEHF_HIP_CHECK(hipStreamSynchronize(st2));
        SEHF_HIP_CHECK(hipStreamSynchronize(st));
        check_gj_info(b.gj, nb, st, "segment cycle");
        check_gj_info(b.gj2, nb * sgrp, st2, "segment cycle (Schur inverses)");
        const double tb1 = wall();
        if (seg_t) { seg_t[2] = t_rec - tb0; seg_t[3] = tb1 - t_rec; }
        return;
    }
    if (replay) {
        // the passes of a back-propagated scan, in the same order over the
        // same psi_n, read from the files: window sums, A and B, then the
        // dipole sums and the fit residual from the top down
        for (int n = n_fe; n >= n_fs; --n) {
            psi_in(n, b.Psi.p);
            fit_sums(n, b.Psi.p);
        }
        fit_solve(b.Am.p, b.Bm.p, b.YJ.p, b.YY.p, b.coef.p, Np, nb, st);
        for (int n = N - 1; n >= 0; --n) {
            const bool in_win = n >= n_fs && n <= n_fe, in_dip = n < s->n_pts;
            if (!in_win && !in_dip) continue;
            psi_in(n, b.Psi.p);
            if (in_dip) dip_accum(n, b.Psi.p, b.acc.p);
            if (in_win) fit_resid_at(n, b.Psi.p);
            progress("dipole replay n =", n, N);
        }
    } else {
        sweep(Pass::FitSums, n_seg - 1, fs_lo / L, false);
        // A, B of the fit from the completed window sums; the next pass starts
        // again from the outer boundary
        fit_solve(b.Am.p, b.Bm.p, b.YJ.p, b.YY.p, b.coef.p, Np, nb, st);
        zmap(b.P.p, b.Z.p);
        if (adj) sweep(Pass::Resid, n_seg - 1, n_fs / L, false);
        else sweep(Pass::Full, n_seg - 1, 0, false);
    }
    io_drain();
    check_gj_info(b.gj, nb, st, "scan back-propagation");
    check_gj_info(b.gj2, nb * sgrp, st2, "scan back-propagation (Schur inverses)");
    const double tw4 = wall();
    // ---------------- asymptotic fit ----------------
    if (tan) fit_dsolve(b.dAm.p, b.dBm.p, b.YJ.p, b.YY.p, b.dYJ.p, b.dYY.p, b.coef.p, b.dcoef.p, Np, nb, st);
    SEHF_HIP_CHECK(hipMemcpyAsync(b.T.p, b.Am.p, nb * pp * sizeof(double), hipMemcpyDeviceToDevice, st));
    gj(b.T.p, b.Ainv.p, Np, false);
    check_gj_info(b.gj, nb, st, "fitted A matrix");
    {
        Gemm g;   // K_fit = B A^{-1}
        g.M = Np; g.N = Np; g.K = Np;
        g.A = b.Bm.p; g.lda = Np; g.sA = pp; g.B = b.Ainv.p; g.ldb = Np; g.sB = pp;
        g.C = b.Kfit.p; g.ldc = Np; g.sC = pp; g.batch = nb;
        gemm(g, st);
    }
    // M = ((A - iB)^{-1})^dagger via the real 2Np system
    w2_assemble(b.W2.p, b.Am.p, b.Bm.p, Np, nb, st);
    gj(b.W2.p, b.W2i.p, 2 * Np, false);
    const int64_t w2s = 4 * static_cast<int64_t>(pp), w2ld = 2 * Np;
    auto mgemm = [&](double* C, const double* A, double alpha, const double* Bp, int64_t ldb, int ncol) {
        Gemm g;   // C (Np x ncol) = alpha * A^T * op(B)   (op(B) = rows of acc, transposed)
        g.M = Np; g.N = ncol; g.K = Np;
        g.A = A; g.lda = w2ld; g.sA = w2s; g.tA = 1;
        g.B = Bp; g.ldb = ldb; g.sB = static_cast<int64_t>(ns) * Np; g.tB = 1;
        g.C = C; g.ldc = Np; g.sC = static_cast<int64_t>(Np) * ncol; g.alpha = alpha; g.batch = nb;
        gemm(g, st);
    };
    mgemm(b.DRre.p, b.W2i.p, 1.0, b.acc.p, ns, 6);
    mgemm(b.DRim.p, b.W2i.p + Np, -1.0, b.acc.p, ns, 6);
    mgemm(b.ore.p, b.W2i.p, 1.0, b.acc.p + 6, ns, n_occ);
    mgemm(b.oim.p, b.W2i.p + Np, -1.0, b.acc.p + 6, ns, n_occ);
    dipole_final(b.Dre.p, b.Dim.p, b.DRre.p, b.DRim.p, b.ore.p, b.oim.p, s->dcorr.p, Np, n_occ, nb, st);
    if (tan) {
        // M = (C^{-1})^dagger with C = A - iB, so dM = -M dC^dagger M and, for
        // y = M x, d(M x) = M v with v = dx - dC^dagger y, dC^dagger = dA^T + i dB^T:
        //   v_re = dx - dA^T y_re + dB^T y_im,  v_im = -dB^T y_re - dA^T y_im;
        //   M v = (P^T - i Q^T) v with P + iQ = C^{-1} (the first block column of W2i).
        auto tg = [&](double* C, const double* A, int64_t lda, int64_t sA, const double* Bp, int ncol,
                      double alpha, double beta) {
            Gemm g;   // C (Np x ncol) = alpha A^T B + beta C
            g.M = Np; g.N = ncol; g.K = Np;
            g.A = A; g.lda = lda; g.sA = sA; g.tA = 1;
            g.B = Bp; g.ldb = Np; g.sB = static_cast<int64_t>(Np) * ncol;
            g.C = C; g.ldc = Np; g.sC = static_cast<int64_t>(Np) * ncol; g.alpha = alpha; g.beta = beta; g.batch = nb;
            gemm(g, st);
        };
        auto mtan = [&](int r0, int ncol, const double* yre, const double* yim, double* ore_, double* oim_) {
            const int64_t sv = static_cast<int64_t>(Np) * ncol;
            transpose_scaled(b.vre.p, Np, sv, b.dacc.p + r0, ns, static_cast<int64_t>(ns) * Np, 1.0, ncol, Np, nb, st);
            tg(b.vre.p, b.dAm.p, Np, pp, yre, ncol, -1.0, 1.0);
            tg(b.vre.p, b.dBm.p, Np, pp, yim, ncol, 1.0, 1.0);
            tg(b.vim.p, b.dBm.p, Np, pp, yre, ncol, -1.0, 0.0);
            tg(b.vim.p, b.dAm.p, Np, pp, yim, ncol, -1.0, 1.0);
            tg(ore_, b.W2i.p, w2ld, w2s, b.vre.p, ncol, 1.0, 0.0);
            tg(ore_, b.W2i.p + Np, w2ld, w2s, b.vim.p, ncol, 1.0, 1.0);
            tg(oim_, b.W2i.p, w2ld, w2s, b.vim.p, ncol, 1.0, 0.0);
            tg(oim_, b.W2i.p + Np, w2ld, w2s, b.vre.p, ncol, -1.0, 1.0);
        };
        mtan(0, 6, b.DRre.p, b.DRim.p, b.dDRre.p, b.dDRim.p);
        if (n_occ > 0) mtan(6, n_occ, b.ore.p, b.oim.p, b.dore.p, b.doim.p);
        dipole_final(b.dDre.p, b.dDim.p, b.dDRre.p, b.dDRim.p, b.dore.p, b.doim.p, s->dcorr.p, Np, n_occ, nb, st);
    }
    check_gj_info(b.gj, nb, st, "dipole normalisation");
    // ---------------- download ----------------
    auto dl = [&](double* dst, const double* src, size_t per) {
        if (dst) SEHF_HIP_CHECK(hipMemcpyAsync(dst + static_cast<size_t>(out_off) * per, src, per * nb * sizeof(double),
                                               hipMemcpyDeviceToHost, st));
    };
    if (!replay) dl(out->K, b.Kd.p, pp);   // a replay does not form K
    dl(out->A, b.Am.p, pp);
    dl(out->B, b.Bm.p, pp);
    dl(out->K_fit, b.Kfit.p, pp);
    std::vector<double> acc(static_cast<size_t>(ns) * Np * nb), dre(6 * static_cast<size_t>(Np) * nb), dim(dre.size()),
        drre(dre.size()), drim(dre.size()), red(2 * nb);
    SEHF_HIP_CHECK(hipMemcpyAsync(acc.data(), b.acc.p, acc.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    SEHF_HIP_CHECK(hipMemcpyAsync(dre.data(), b.Dre.p, dre.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    SEHF_HIP_CHECK(hipMemcpyAsync(dim.data(), b.Dim.p, dim.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    SEHF_HIP_CHECK(hipMemcpyAsync(drre.data(), b.DRre.p, drre.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    SEHF_HIP_CHECK(hipMemcpyAsync(drim.data(), b.DRim.p, drim.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    SEHF_HIP_CHECK(hipMemcpyAsync(red.data(), b.red.p, red.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    std::vector<double> Ah(pp * nb), Kh(pp * nb);
    SEHF_HIP_CHECK(hipMemcpyAsync(Ah.data(), b.Am.p, Ah.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    SEHF_HIP_CHECK(hipMemcpyAsync(Kh.data(), b.Kfit.p, Kh.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    std::vector<double> dacc, ddre, ddim;
    if (tan && tout) {
        dl(tout->dK, b.dKd.p, pp);
        dl(tout->dA, b.dAm.p, pp);
        dl(tout->dB, b.dBm.p, pp);
        dacc.resize(acc.size()); ddre.resize(dre.size()); ddim.resize(dre.size());
        SEHF_HIP_CHECK(hipMemcpyAsync(dacc.data(), b.dacc.p, dacc.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        SEHF_HIP_CHECK(hipMemcpyAsync(ddre.data(), b.dDre.p, ddre.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        SEHF_HIP_CHECK(hipMemcpyAsync(ddim.data(), b.dDim.p, ddim.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    SEHF_HIP_CHECK(hipStreamSynchronize(st));
    const double tw5 = wall();
    g.t_stage[1] += tw2 - tw1; g.t_stage[2] += tw3 - tw2; g.t_stage[3] += tw4 - tw3;
    g.t_stage[4] += tw5 - tw4; g.t_stage[5] += tw1 - tw0;
    for (int e = 0; e < nb; ++e) {
        const int E_ = out_off + e;
        const double* a = acc.data() + static_cast<size_t>(e) * ns * Np;
        for (int slot = 0; slot < 6; ++slot)
            for (int be = 0; be < Np; ++be) {
                const size_t o = static_cast<size_t>(e) * 6 * Np + static_cast<size_t>(slot) * Np + be;
                const size_t od = (static_cast<size_t>(E_) * 6 + slot) * Np + be;
                if (out->d_raw) out->d_raw[od] = a[slot + static_cast<size_t>(be) * ns];
                if (out->D_raw) { out->D_raw[2 * od] = drre[o]; out->D_raw[2 * od + 1] = drim[o]; }
                if (out->D) { out->D[2 * od] = dre[o]; out->D[2 * od + 1] = dim[o]; }
            }
        if (out->b_overlap)
            for (int al = 0; al < n_occ; ++al)
                for (int be = 0; be < Np; ++be)
                    out->b_overlap[(static_cast<size_t>(E_) * n_occ + al) * Np + be] = a[6 + al + static_cast<size_t>(be) * ns];
        if (tan && tout) {
            const double* da = dacc.data() + static_cast<size_t>(e) * ns * Np;
            for (int slot = 0; slot < 6; ++slot)
                for (int be = 0; be < Np; ++be) {
                    const size_t o = static_cast<size_t>(e) * 6 * Np + static_cast<size_t>(slot) * Np + be;
                    const size_t od = (static_cast<size_t>(E_) * 6 + slot) * Np + be;
                    if (tout->dd_raw) tout->dd_raw[od] = da[slot + static_cast<size_t>(be) * ns];
                    if (tout->dD) { tout->dD[2 * od] = ddre[o]; tout->dD[2 * od + 1] = ddim[o]; }
                }
            if (tout->db_overlap)
                for (int al = 0; al < n_occ; ++al)
                    for (int be = 0; be < Np; ++be)
                        tout->db_overlap[(static_cast<size_t>(E_) * n_occ + al) * Np + be] = da[6 + al + static_cast<size_t>(be) * ns];
        }
        if (out->diag) {
            double* d = out->diag + static_cast<size_t>(E_) * 16;
            std::fill(d, d + 16, 0.0);
            const double* A = Ah.data() + static_cast<size_t>(e) * pp;
            const double* K = Kh.data() + static_cast<size_t>(e) * pp;
            double ami = 0.0, ks_ = 0.0;
            for (int j = 0; j < Np; ++j)
                for (int i = 0; i < Np; ++i) {
                    ami = std::max(ami, std::fabs(A[i + static_cast<size_t>(j) * Np] - (i == j ? 1.0 : 0.0)));
                    ks_ = std::max(ks_, std::fabs(K[i + static_cast<size_t>(j) * Np] - K[j + static_cast<size_t>(i) * Np]));
                }
            d[0] = red[2 * e] / std::max(red[2 * e + 1], 1e-30);
            d[1] = red[2 * e];
            d[2] = ks_;
            d[3] = ami;
            d[4] = s->n_tr; d[5] = s->n_start; d[6] = n_fs; d[7] = n_fe;
            d[8] = shA[e]; d[9] = shS[e]; d[10] = n_closed[e];
        }
    }
    if (pw) pw->commit();
}

}  // namespace

extern "C" {

int sehf_scan_create(sehf_ctx* ctx, const sehf_problem* p, sehf_scan** out) {
    const char* e = std::getenv("SEHF_TABLES");
    const int tables = (e && std::string(e) == "streamed") ? SEHF_TABLES_STREAMED : SEHF_TABLES_RESIDENT;
    return sehf_scan_create_with(ctx, p, tables, out);
}

int sehf_scan_create_with(sehf_ctx* ctx, const sehf_problem* p, int tables, sehf_scan** out) {
    return guarded([&] {
        if (tables != SEHF_TABLES_RESIDENT && tables != SEHF_TABLES_STREAMED)
            throw std::runtime_error("sehf_scan_create_with: tables must be SEHF_TABLES_RESIDENT or _STREAMED");
        auto s = std::make_unique<sehf_scan>();
        s->tables = tables;
        s->ctx = ctx;
        s->st = sehf_ctx_stream(ctx);
        s->ctx_device = sehf_ctx_device(ctx);
        if (const char* e = std::getenv("SEHF_SCAN_GROUPS")) s->n_groups = std::max(1, std::min(4, std::atoi(e)));
        if (const char* e = std::getenv("SEHF_SINV_INVERSE")) s->sinv_spd = std::string(e) != "gj";
        if (const char* e = std::getenv("SEHF_DIPOLE_PATH"))
            s->dipole_path = std::atoi(e) == SEHF_DIPOLES_BACKWARD ? SEHF_DIPOLES_BACKWARD : SEHF_DIPOLES_FORWARD;
        s->S = std::make_unique<Setup>(*p);
        const Setup& S = *s->S;
        s->Np = S.Np; s->Nf = S.Nf; s->Nt = S.Nt; s->Nr = S.Nr; s->n_tr = S.n_tr; s->n_start = S.n_start;
        s->h = S.dr; s->c6 = S.dr * S.dr / 6.0; s->nslot = 6 + S.n_occ;
        const double t0 = wall();
        build_shared(s.get());
        s->t_stage[0] = wall() - t0;
        *out = s.release();
    });
}

int sehf_scan_set_checkpoints(sehf_scan* s, int where) {
    return guarded([&] {
        if (where != SEHF_CKPT_DEVICE && where != SEHF_CKPT_HOST)
            throw std::runtime_error("sehf_scan_set_checkpoints: where must be SEHF_CKPT_DEVICE or _HOST");
        s->ckpt_where = where;
    });
}

int sehf_scan_memory(const sehf_scan* s, size_t* bytes) {
    return guarded([&] {
        size_t sh = s->S0.bytes() + s->BD.bytes() + s->Dinv.bytes() + s->VN.bytes() + s->BN.bytes() +
                    s->negDinvN.bytes() + s->Drows.bytes() + s->dcorr.bytes() + s->cvd.bytes() + s->bdb.bytes();
        const auto& T = s->tin;
        sh += T.pi.bytes() + T.pj.bytes() + T.vsig.bytes() + T.lmu.bytes() + T.xlam.bytes() + T.vptr.bytes() +
              T.xptr.bytes() + T.vval.bytes() + T.xval.bytes() + T.vs.bytes() + T.vp.bytes() + T.ir2.bytes() +
              T.chi.bytes() + T.fac.bytes() + T.live.bytes();
        size_t dev = 0, host = 0;
        for (const auto& g : s->groups) {
            Batch& b = const_cast<Batch&>(g->bt);
            BatchDims d = dims_of(s);
            batch_arrays(b, d, 1, 1, 1, 1, b.has_tan, [&](auto& arr, size_t, bool) {
                (arr.host ? host : dev) += arr.bytes();
            });
            dev += gj_buffer_bytes(b.gj.n, b.gj.batch) + gj_buffer_bytes(b.gj2.n, b.gj2.batch);
            host += g->io_stage.bytes();   // pinned staging of the psi files
        }
        bytes[0] = sh; bytes[1] = dev; bytes[2] = host;
    });
}

int sehf_scan_info(const sehf_scan* s, int* info) {
    return guarded([&] {
        info[0] = s->Np; info[1] = s->Nf; info[2] = s->Nt; info[3] = s->n_tr;
        info[4] = s->n_start; info[5] = s->n_pts; info[6] = s->Nr; info[7] = s->S->l_exch;
    });
}

}  // extern "C"

namespace {

void scan_run(sehf_scan* s, const double* k, const double* dir, int nk, int batch, int seg_len, sehf_scan_out* out,
              sehf_scan_tangent_out* tout) {
    if (nk <= 0) return;
    if (!out) throw std::runtime_error("sehf_scan_run: out must not be NULL");
    for (int i = 0; i < nk; ++i)
        if (!(k[i] > 0.0)) throw std::runtime_error("sehf_scan_run: every k must be > 0");
    if (dir)
        for (int i = 0; i < 2 * nk; ++i)
            if (!std::isfinite(dir[i])) throw std::runtime_error("sehf_scan_run_tangent: directions must be finite");
    const int N = s->Nr - 1;
    int L = seg_len > 0 ? seg_len : static_cast<int>(std::ceil(std::sqrt(static_cast<double>(N))));
    L = std::max(L, s->n_start + 1);
    const int n_seg = (N + L - 1) / L;
    const int n_fit = 200;
    int B = batch > 0 ? std::min(batch, nk) : nk;
    const int n_batches = (nk + B - 1) / B;
    const int G = std::max(1, std::min(s->n_groups, n_batches));
    while (static_cast<int>(s->groups.size()) < G) {
        s->groups.emplace_back(new Group());
        s->groups.back()->init();
    }
    // a replay from psi files needs no segment buffers or checkpoints
    const bool replay = !s->replay_root.empty();
    for (int gi = 0; gi < G; ++gi)
        alloc_batch(s, *s->groups[gi], B, replay ? 1 : L, replay ? 1 : n_seg - ckpt_lo(s, L, dir != nullptr), n_fit,
                    dir != nullptr);
    if (out->d_corr) std::memcpy(out->d_corr, s->dcorr_h.data(), s->dcorr_h.size() * sizeof(double));
    // batch i runs on group i % G; each group is driven by its own host thread
    std::vector<std::string> errs(G);
    auto drive = [&](int gi) {
        try {
            SEHF_HIP_CHECK(hipSetDevice(s->ctx_device));
            for (int bi = gi; bi < n_batches; bi += G) {
                const int e0 = bi * B, nb = std::min(B, nk - e0);
                run_batch(s, *s->groups[gi], k + e0, dir ? dir + 2 * e0 : nullptr, nb, L, out, tout, e0);
            }
        } catch (const std::exception& e) {
            errs[gi] = e.what();
        }
    };
    if (G == 1) {
        drive(0);
    } else {
        std::vector<std::thread> th;
        for (int gi = 0; gi < G; ++gi) th.emplace_back(drive, gi);
        for (auto& x : th) x.join();
    }
    for (int gi = 0; gi < G; ++gi)
        if (!errs[gi].empty()) throw std::runtime_error(errs[gi]);
}

// One segment cycle of the scan for all nk energies in one batch (group 0).
void scan_segment(sehf_scan* s, const double* k, const double* dir, int nk, int seg, int seg_len, double* t,
                  int* layout) {
    const int N = s->Nr - 1;
    int L = seg_len > 0 ? seg_len : static_cast<int>(std::ceil(std::sqrt(static_cast<double>(N))));
    L = std::max(L, s->n_start + 1);
    const int n_seg = (N + L - 1) / L;
    int n_fs = 0, n_fe = 0;
    fit_window(s->Nr, n_fs, n_fe);
    if (layout) { layout[0] = L; layout[1] = n_seg; layout[2] = n_fs; layout[3] = n_fs / L; }
    if (seg < 0) return;   // layout query
    if (nk <= 0 || !k) throw std::runtime_error("sehf_scan_run_segment: empty k list");
    for (int i = 0; i < nk; ++i)
        if (!(k[i] > 0.0)) throw std::runtime_error("sehf_scan_run_segment: every k must be > 0");
    if (dir)
        for (int i = 0; i < 2 * nk; ++i)
            if (!std::isfinite(dir[i])) throw std::runtime_error("sehf_scan_run_segment: directions must be finite");
    if (seg >= n_seg)
        throw std::runtime_error("sehf_scan_run_segment: segment " + std::to_string(seg) + " outside [0, " +
                                 std::to_string(n_seg) + ")");
    if (s->groups.empty()) {
        s->groups.emplace_back(new Group());
        s->groups.back()->init();
    }
    SEHF_HIP_CHECK(hipSetDevice(s->ctx_device));
    alloc_batch(s, *s->groups[0], nk, L, n_seg - ckpt_lo(s, L, dir != nullptr), 200, dir != nullptr);
    double tt[4] = {0, 0, 0, 0};
    run_batch(s, *s->groups[0], k, dir, nk, L, nullptr, nullptr, 0, 0, nullptr, seg, tt);
    if (t) std::copy(tt, tt + 4, t);
}

}  // namespace

extern "C" {

int sehf_scan_run_segment(sehf_scan* s, const double* k, const double* dir, int nk, int seg, int seg_len,
                          double* seconds, int* layout) {
    return guarded([&] { scan_segment(s, k, dir, nk, seg, seg_len, seconds, layout); });
}

int sehf_scan_run(sehf_scan* s, const double* k, int nk, int batch, int seg_len, sehf_scan_out* out) {
    return guarded([&] { scan_run(s, k, nullptr, nk, batch, seg_len, out, nullptr); });
}

int sehf_scan_run_tangent(sehf_scan* s, const double* k, const double* dir, int nk, int batch, int seg_len,
                          sehf_scan_out* out, sehf_scan_tangent_out* tout) {
    return guarded([&] {
        if (!dir || !tout) throw std::runtime_error("sehf_scan_run_tangent: dir and tout must not be NULL");
        scan_run(s, k, dir, nk, batch, seg_len, out, tout);
    });
}

int sehf_scan_autotune(sehf_scan* s, const double* k, int nk, int probe_steps, int tangent, int* batch_out,
                       int* groups_out, int* tiles_out, double* report, int* n_report) {
    return guarded([&] {
        if (nk <= 0 || !k) throw std::runtime_error("sehf_scan_autotune: empty k list");
        for (int i = 0; i < nk; ++i)
            if (!(k[i] > 0.0)) throw std::runtime_error("sehf_scan_autotune: every k must be > 0");
        const int steps = std::max(2, std::min(probe_steps > 0 ? probe_steps : 8, s->Nr - 1 - s->n_start));
        const int N = s->Nr - 1;
        int L = static_cast<int>(std::ceil(std::sqrt(static_cast<double>(N))));
        L = std::max(L, s->n_start + 1);
        const int n_seg = (N + L - 1) / L, n_fit = 200;
        std::vector<double> dir(tangent ? 2 * static_cast<size_t>(nk) : 0);
        for (int i = 0; i < nk && tangent; ++i) { dir[2 * i] = 1.0; dir[2 * i + 1] = 0.0; }
        const int cand[] = {1, 2, 4, 8, 12, 16, 24, 32, 48, 64, 96, 128};
        // one probe: allocate G groups of batch B, time `steps` forward steps
        // of all groups started together; < 0 when the buffers do not fit
        auto probe = [&](int B, int G) -> double {
            s->groups.clear();   // release every batch buffer before sizing the candidate
            try {
                while (static_cast<int>(s->groups.size()) < G) {
                    s->groups.emplace_back(new Group());
                    s->groups.back()->init();
                }
                for (int gi = 0; gi < G; ++gi)
                    alloc_batch(s, *s->groups[gi], B, L, n_seg - ckpt_lo(s, L, tangent != 0), n_fit, tangent != 0);
            } catch (const std::exception&) {
                (void)hipGetLastError();
                return -1.0;
            }
            // energies of the probe: group gi takes k[(gi*B + e) % nk]
            std::vector<std::vector<double>> kg(G, std::vector<double>(B)), dg(G, std::vector<double>(2 * B));
            for (int gi = 0; gi < G; ++gi)
                for (int e = 0; e < B; ++e) {
                    const int src = (gi * B + e) % nk;
                    kg[gi][e] = k[src];
                    if (tangent) { dg[gi][2 * e] = dir[2 * src]; dg[gi][2 * e + 1] = dir[2 * src + 1]; }
                }
            double wall_s = 0.0;
            for (int rep = 0; rep < 2; ++rep) {   // rep 0 warms the streams and kernels
                std::vector<std::string> errs(G);
                ProbeSync sync;
                sync.total = G;
                auto drive = [&](int gi) {
                    try {
                        SEHF_HIP_CHECK(hipSetDevice(s->ctx_device));
                        run_batch(s, *s->groups[gi], kg[gi].data(), tangent ? dg[gi].data() : nullptr, B, L, nullptr,
                                  nullptr, 0, rep == 0 ? 2 : steps, &sync);
                    } catch (const std::exception& e) {
                        errs[gi] = e.what();
                        sync.arrive(wall());   // never leave the other groups waiting (no-op once all arrived)
                    }
                };
                std::vector<std::thread> th;
                for (int gi = 0; gi < G; ++gi) th.emplace_back(drive, gi);
                for (auto& x : th) x.join();
                for (int gi = 0; gi < G; ++gi)
                    if (!errs[gi].empty()) throw std::runtime_error(errs[gi]);
                double tp = 0.0;
                for (int gi = 0; gi < G; ++gi) tp = std::max(tp, s->groups[gi]->t_probe);
                wall_s = std::max(tp, 1e-9);
            }
            return static_cast<double>(G) * B * steps / wall_s;   // energy-steps per second
        };
        double best = -1.0;
        int bestB = 1, bestG = 1, bestT = gemm_tiles(), nrep = 0;
        auto note = [&](int B, int G, int T, double rate) {
            if (report && nrep < 32) {
                report[4 * nrep] = B; report[4 * nrep + 1] = G; report[4 * nrep + 2] = T; report[4 * nrep + 3] = rate;
            }
            ++nrep;
        };
        // batch width and groups at the current tile policy
        for (int G = 1; G <= 2; ++G) {
            double prev = -1.0;
            for (int B : cand) {
                if (G * B > nk) break;
                const double rate = probe(B, G);
                if (rate < 0.0) break;   // does not fit: larger batches do not either
                note(B, G, bestT, rate);
                if (rate > best) { best = rate; bestB = B; bestG = G; }
                if (prev > 0.0 && rate < 1.03 * prev) break;   // saturated: larger batches only cost memory
                prev = rate;
            }
        }
        // GEMM tile policy at that batch width
        const int T0 = bestT;
        for (int T = 0; T <= 2; ++T) {
            if (T == T0) continue;
            set_gemm_tiles(T);
            const double rate = probe(bestB, bestG);
            note(bestB, bestG, T, rate);
            if (rate > best) { best = rate; bestT = T; }
        }
        set_gemm_tiles(bestT);
        s->groups.clear();
        s->n_groups = bestG;
        if (batch_out) *batch_out = bestB;
        if (groups_out) *groups_out = bestG;
        if (tiles_out) *tiles_out = bestT;
        if (n_report) *n_report = std::min(nrep, 32);
    });
}

int sehf_set_gemm_tiles(int policy) {
    return guarded([&] { set_gemm_tiles(policy); });
}

int sehf_set_gj_pivoting(int mode) {
    return guarded([&] { set_gj_pivoting(mode); });
}

int sehf_scan_stats(const sehf_scan* s, double* t) {
    return guarded([&] {
        for (int i = 0; i < 8; ++i) {
            t[i] = s->t_stage[i];
            if (i > 0)
                for (auto& g : s->groups) t[i] += g->t_stage[i];
        }
    });
}

int sehf_prof_begin(int every) {
    return guarded([&] { prof_begin(every); });
}

int sehf_prof_end(double* out) {
    return guarded([&] { prof_end(out); });
}

int sehf_scan_set_groups(sehf_scan* s, int groups) {
    return guarded([&] {
        if (groups < 1 || groups > 4)
            throw std::runtime_error("sehf_scan_set_groups: groups must be in [1, 4] (one stream each)");
        s->n_groups = groups;
    });
}

int sehf_scan_set_psi_store(sehf_scan* s, const char* root, int keep) {
    return guarded([&] {
        if (!s) throw std::runtime_error("sehf_scan_set_psi_store: NULL scan");
        if (keep & ~(SEHF_PSI_TAIL | SEHF_PSI_MAIN))
            throw std::runtime_error("sehf_scan_set_psi_store: keep must be a combination of SEHF_PSI_TAIL and SEHF_PSI_MAIN");
        if (!root || !*root || keep == 0) {
            s->psi_root.clear();
            s->psi_keep = 0;
            return;
        }
        s->psi_root = root;
        s->psi_keep = keep;
    });
}

int sehf_scan_dipole_only(sehf_scan* s, const char* root, const double* k, int nk, int batch, sehf_scan_out* out) {
    return guarded([&] {
        if (!s) throw std::runtime_error("sehf_scan_dipole_only: NULL scan");
        if (!root || !*root) throw std::runtime_error("sehf_scan_dipole_only: root must name the psi store");
        struct Clear {
            sehf_scan* s;
            ~Clear() { s->replay_root.clear(); }
        } clear{s};
        s->replay_root = root;
        scan_run(s, k, nullptr, nk, batch, 0, out, nullptr);
    });
}

int sehf_scan_set_dipole_path(sehf_scan* s, int path) {
    return guarded([&] {
        if (path != SEHF_DIPOLES_FORWARD && path != SEHF_DIPOLES_BACKWARD)
            throw std::runtime_error("sehf_scan_set_dipole_path: path must be SEHF_DIPOLES_FORWARD (0) or "
                                     "SEHF_DIPOLES_BACKWARD (1)");
        s->dipole_path = path;
    });
}

void sehf_scan_destroy(sehf_scan* s) { delete s; }

}  // extern "C"
