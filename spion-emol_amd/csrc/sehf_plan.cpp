// sehf_plan.cpp -- device storage plan of the batched scan (sehf_plan_storage).
//
// The reference's StoragePlanner (src/scatt/StoragePlanner.cpp:31-200)
// places pot, sinv, rinv and psi in RAM or in chunked disk files, greedily in
// that priority order, under (RAM - runtime - non-storage) x (1 - reserve).
// Here the scarce level is HBM and the quantity to buy with it is batch
// width (energies per batch set the throughput), so the decisions are:
//   pot    resident tables (N_psi^2 Nr + N_psi N_f n_transition doubles) or
//          their compact inputs with a per-step rebuild on the device;
//   sinv / rinv  checkpoint spacing L (segment buffers L (N_total^2 +
//          N_psi^2) per energy against the checkpoints) and the checkpoint
//          level, HBM or pinned host memory;
//   psi    never stored.
// Candidates are tried in order of increasing cost per step: resident
// tables, then streamed tables (a per-step rebuild on the side stream), with
// checkpoints every ceil(sqrt(Nr - 1)) points in HBM; then the checkpoint
// spacing that gives the widest batch, in HBM and then in pinned host
// memory.  The first candidate whose batch reaches the wanted width wins,
// else the one with the widest batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sehf.h"
#include "sehf_plan.h"

namespace sehf {
std::string& last_error_ref();
}

using namespace sehf;

namespace {

constexpr int kFitPoints = 200;      // fit window length (AsymptoticAmplitudes.cpp:70-91)
constexpr int kDefaultBatchCap = 64;

std::string gb(size_t b) {
    char buf[32];
    std::snprintf(buf, sizeof buf, "%.2f GB", static_cast<double>(b) / 1e9);
    return buf;
}

// First fit-window point (the scan's fit_window, clamped to the kept tail).
int fit_lo(int Nr) {
    const int N = Nr - 1;
    const int asym_off = std::max(0, N - 300 + 1);
    return std::max(std::max(Nr - 200, static_cast<int>(0.9 * Nr)), asym_off);
}

struct Candidate {
    int tables = SEHF_TABLES_RESIDENT, ckpt = SEHF_CKPT_DEVICE, L = 1, n_ck = 1, batch = 0;
    size_t shared = 0, per_energy = 0, host_per_energy = 0;
};

}  // namespace

extern "C" int sehf_plan_storage(const sehf_plan_in* in, sehf_plan* out) {
    try {
        if (!in || !out) throw std::runtime_error("sehf_plan_storage: NULL argument");
        const sehf_plan_in& q = *in;
        if (q.Nr < 3 || q.n_psi < 1 || q.n_f < 0 || q.n_occ < 0 || q.n_exp < 1 || q.n_lambda < 0 ||
            q.n_energies < 1 || q.n_transition < 0)
            throw std::runtime_error("sehf_plan_storage: sizes out of range");
        if (q.n_f > 0 && q.n_occ < 1) throw std::runtime_error("sehf_plan_storage: exchange block without orbitals");
        if (!(q.reserve_fraction >= 0.0 && q.reserve_fraction < 0.9))
            throw std::runtime_error("sehf_plan_storage: reserve_fraction must be in [0, 0.9)");
        if (q.dipole_path != SEHF_DIPOLES_FORWARD && q.dipole_path != SEHF_DIPOLES_BACKWARD)
            throw std::runtime_error("sehf_plan_storage: dipole_path must be SEHF_DIPOLES_FORWARD or _BACKWARD");
        size_t dev = q.device_bytes;
        if (dev == 0) {
            size_t fr = 0, tot = 0;
            if (hipMemGetInfo(&fr, &tot) != hipSuccess) throw std::runtime_error("sehf_plan_storage: hipMemGetInfo failed");
            dev = fr;
        }
        const size_t avail = dev > q.runtime_bytes ? dev - q.runtime_bytes : 0;
        const size_t budget = static_cast<size_t>(static_cast<double>(avail) * (1.0 - q.reserve_fraction));
        const int cap = q.batch_cap > 0 ? q.batch_cap : kDefaultBatchCap;
        const int want = std::min(q.n_energies, cap);

        const size_t Np = q.n_psi, Nf = q.n_f, Nt = Np + Nf, Nr = q.Nr, N = Nr - 1;
        const size_t n_bd = std::min<size_t>(q.n_transition, Nr);
        const size_t pp = Np * Np, pf = Np * std::max<size_t>(Nf, 1), d8 = sizeof(double);
        const size_t ns = 6 + q.n_occ;
        // energy-independent bytes common to both table modes: D^{-1}, the
        // two diagonals, the dipole rows, V_N and B_N
        const size_t common = d8 * (std::max<size_t>(Nf, 1) * Nr + 2 * Np * Nr + ns * Np * Nr + pp + pf + Nf);
        const size_t resident = common + d8 * (pp * Nr + (Nf > 0 ? pf * n_bd : 0));
        // streamed: V_s(r), chi(r), four per-point factors and the sparse
        // Gaunt tables (bounded by 16 bytes per coupling: N_psi^2 / 2 pairs
        // with up to l_cont + 1 couplings each, N_psi n_sigma exchange rows
        // with up to l_exch + 1)
        const size_t l_cont = static_cast<size_t>(std::lround(std::sqrt(static_cast<double>(Np)))) - 1;
        const size_t n_sigma = q.n_occ > 0 ? Nf / q.n_occ : 0;
        const size_t l_exch = static_cast<size_t>(std::lround(std::sqrt(static_cast<double>(n_sigma))));
        const size_t gaunt = 16 * ((Np * (Np + 1) / 2) * (l_cont + 1) + Np * n_sigma * (l_exch + 1));
        const size_t streamed = common + d8 * (static_cast<size_t>(q.n_exp) * Nr +
                                               static_cast<size_t>(q.n_occ) * q.n_lambda * n_bd + 4 * Nr) + gaunt;

        BatchDims d;
        d.Np = Np; d.Nf = Nf; d.Nt = Nt; d.ns = ns; d.n_occ = q.n_occ;
        d.nstab = std::min<size_t>(60, Nr); d.n_start = 1; d.Nr = Nr;
        d.sinv_group = sinv_group_for(Np, true);   // the scan's default Schur inverse (SPD block elimination)
        const bool tan = q.tangent != 0;
        const bool fwd = q.dipole_path == SEHF_DIPOLES_FORWARD && !tan;
        const int nfs = fit_lo(q.Nr);
        auto n_ck_of = [&](int L) {
            const int n_seg = static_cast<int>((N + L - 1) / L);
            return fwd ? n_seg - nfs / L : n_seg;
        };
        auto evaluate = [&](int tables, int ckpt, int L) {
            Candidate c;
            c.tables = tables; c.ckpt = ckpt; c.L = L; c.n_ck = n_ck_of(L);
            d.streamed = tables == SEHF_TABLES_STREAMED;
            d.ckpt_host = ckpt == SEHF_CKPT_HOST;
            c.shared = d.streamed ? streamed : resident;
            const BatchBytes one = batch_bytes(d, 1, L, c.n_ck, kFitPoints, tan);
            c.per_energy = one.device;
            c.host_per_energy = one.host;
            // widest batch that fits (the bytes grow monotonically with it)
            int lo = 0, hi = q.n_energies;
            while (lo < hi) {
                const int mid = (lo + hi + 1) / 2;
                const BatchBytes bb = batch_bytes(d, mid, L, c.n_ck, kFitPoints, tan);
                const bool fits = c.shared + bb.device <= budget && bb.host <= q.host_bytes;
                if (fits) lo = mid;
                else hi = mid - 1;
            }
            c.batch = std::min(lo, cap);
            return c;
        };
        const int L0 = std::max(2, static_cast<int>(std::ceil(std::sqrt(static_cast<double>(N)))));
        // the spacing with the widest batch (ties: the longer, which
        // recomputes less)
        auto best_of = [&](int tables, int ckpt) {
            Candidate best = evaluate(tables, ckpt, L0);
            for (int L = L0 - 1; L >= 2 && best.batch < want; --L) {
                const Candidate c = evaluate(tables, ckpt, L);
                if (c.batch > best.batch) best = c;
            }
            return best;
        };
        std::vector<Candidate> cands;
        cands.push_back(evaluate(SEHF_TABLES_RESIDENT, SEHF_CKPT_DEVICE, L0));
        cands.push_back(evaluate(SEHF_TABLES_STREAMED, SEHF_CKPT_DEVICE, L0));
        cands.push_back(best_of(SEHF_TABLES_STREAMED, SEHF_CKPT_DEVICE));
        if (q.host_bytes > 0) cands.push_back(best_of(SEHF_TABLES_STREAMED, SEHF_CKPT_HOST));
        const Candidate* pick = nullptr;
        for (const auto& c : cands)
            if (c.batch >= want) { pick = &c; break; }
        if (!pick)
            for (const auto& c : cands)
                if (c.batch > 0 && (!pick || c.batch > pick->batch)) pick = &c;
        std::string rep;
        rep += "device storage plan: N_psi=" + std::to_string(Np) + " N_f=" + std::to_string(Nf) + " N_total=" +
               std::to_string(Nt) + " Nr=" + std::to_string(Nr) + (tan ? ", tangent" : "") +
               (fwd ? ", forward-accumulated dipoles" : ", back-propagated dipoles") + "\n";
        rep += "  budget " + gb(budget) + " = (" + gb(dev) + " - runtime " + gb(q.runtime_bytes) + ") x (1 - " +
               std::to_string(q.reserve_fraction).substr(0, 4) + "); host for checkpoints " + gb(q.host_bytes) + "\n";
        rep += "  tables resident " + gb(resident) + ", streamed " + gb(streamed) + "\n";
        for (const auto& c : cands)
            rep += std::string("  candidate: tables ") + (c.tables == SEHF_TABLES_STREAMED ? "streamed" : "resident") +
                   ", checkpoints " + (c.ckpt == SEHF_CKPT_HOST ? "host" : "device") + ", L=" + std::to_string(c.L) +
                   " (" + std::to_string(c.n_ck) + " checkpoints), " + gb(c.per_energy) + " + host " +
                   gb(c.host_per_energy) + " per energy -> batch " + std::to_string(c.batch) + "\n";
        if (!pick) throw std::runtime_error(rep + "  no placement fits one energy");
        std::memset(out, 0, sizeof *out);
        out->tables = pick->tables; out->ckpt = pick->ckpt; out->seg_len = pick->L; out->n_ckpt = pick->n_ck;
        out->batch = pick->batch; out->budget_bytes = budget; out->shared_bytes = pick->shared;
        out->per_energy_bytes = pick->per_energy; out->host_per_energy_bytes = pick->host_per_energy;
        rep += std::string("  chosen: tables ") + (pick->tables == SEHF_TABLES_STREAMED ? "streamed" : "resident") +
               ", checkpoints " + (pick->ckpt == SEHF_CKPT_HOST ? "host" : "device") + ", L=" + std::to_string(pick->L) +
               ", batch " + std::to_string(pick->batch) + " of " + std::to_string(q.n_energies) + " (cap " +
               std::to_string(cap) + ")\n";
        std::snprintf(out->report, sizeof out->report, "%s", rep.c_str());
        return 0;
    } catch (const std::exception& e) {
        last_error_ref() = e.what();
        return 1;
    }
}
