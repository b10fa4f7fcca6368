// sehf_host_api.cpp -- host-only entry points (no device needed): the
// energy-independent inputs and scalar tables the scan uploads, exposed so
// they can be checked on a machine without a GPU.
#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sehf.h"
#include "sehf_host_math.h"
#include "sehf_setup.h"

namespace sehf {
std::string& last_error_ref();
}

using namespace sehf;

template <class F>
static int host_guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        last_error_ref() = e.what();
        return 1;
    }
}

extern "C" {

int sehf_host_potential(const sehf_problem* p, int ir, double* V) {
    return host_guarded([&] { Setup S(*p); S.potential(ir, V); });
}
int sehf_host_exchange_B(const sehf_problem* p, int ir, double* B) {
    return host_guarded([&] { Setup S(*p); S.exchange_B(ir, B); });
}
int sehf_host_sizes(const sehf_problem* p, int* info) {
    return host_guarded([&] {
        Setup S(*p);
        info[0] = S.Np; info[1] = S.Nf; info[2] = S.Nt; info[3] = S.n_tr;
        info[4] = S.n_start; info[5] = S.Nr; info[6] = S.l_exch; info[7] = S.n_lambda;
    });
}
int sehf_host_dipole_tables(const sehf_problem* p, double* rows, double* dcorr, int* n_pts) {
    return host_guarded([&] {
        Setup S(*p);
        std::vector<double> r, d;
        S.dipole_tables(r, d, *n_pts);
        if (rows) std::memcpy(rows, r.data(), r.size() * sizeof(double));
        if (dcorr) std::memcpy(dcorr, d.data(), d.size() * sizeof(double));
    });
}
double sehf_host_ric_j(int l, double x) { return ric_j(l, x); }
double sehf_host_ric_y(int l, double x) { return ric_y(l, x); }
double sehf_host_real_gaunt(int l1, int m1, int l2, int m2, int l3, int m3) {
    // the quadrature table the setup builds its coupling lists from
    static std::mutex mu;
    static std::unique_ptr<GauntTable> table;
    const int need = std::max({l1, l2, l3, 0});
    std::lock_guard<std::mutex> lock(mu);
    if (!table || table->lmax() < need) table.reset(new GauntTable(std::max(need, 64)));
    return (*table)(l1, m1, l2, m2, l3, m3);
}
double sehf_host_min_eig(int n, const double* a) { return sym_min_eig(n, a); }

}  // extern "C"
