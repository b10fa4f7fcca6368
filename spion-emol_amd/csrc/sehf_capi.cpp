// sehf_capi.cpp -- per-step C ABI (the reference's GPU stepper classes).
//
// Each handle owns its device buffers (allocated once at creation, freed at
// destroy) and runs on its context's stream; every call is synchronous like
// the reference's step() methods (src/scatt/gpu/cuda/GpuPropagate.cu:133-139).
#include "../../include/sehf.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>

#include "sehf_common.h"

namespace sehf {

static thread_local std::string g_last_error;
std::string& last_error_ref() { return g_last_error; }

void check_gj_info(const GjBuffers& gb, int batch, hipStream_t s, const char* what) {
    std::vector<int> info(batch);
    SEHF_HIP_CHECK(hipMemcpyAsync(info.data(), gb.info.p, sizeof(int) * batch, hipMemcpyDeviceToHost, s));
    SEHF_HIP_CHECK(hipStreamSynchronize(s));
    // the flags are sticky between checks: clear them once read, so a later
    // call on the same buffers reports only its own zero pivots
    bool any = false;
    for (int b = 0; b < batch; ++b) any = any || info[b] != 0;
    if (any) {
        fill_int(gb.info.p, static_cast<size_t>(batch), 0, s);
        SEHF_HIP_CHECK(hipStreamSynchronize(s));
    }
    for (int b = 0; b < batch; ++b) {
        if (info[b] > 0)
            throw std::runtime_error(std::string(what) + ": singular matrix (zero pivot at index " +
                                     std::to_string(info[b] - 1) + ", batch entry " + std::to_string(b) + ")");
        if (info[b] < 0)
            throw std::runtime_error(std::string(what) + ": matrix not positive definite (non-positive pivot at "
                                     "index " + std::to_string(-info[b] - 1) + ", batch entry " + std::to_string(b) +
                                     ")");
    }
}

template <class F>
static int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return 1;
    } catch (...) {
        g_last_error = "unknown error";
        return 1;
    }
}

}  // namespace sehf

using namespace sehf;

struct sehf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string name;
    size_t mem = 0;
    int is_gpu = 1;
};

hipStream_t sehf_ctx_stream(const sehf_ctx* c) { return c->stream; }
int sehf_ctx_device(const sehf_ctx* c) { return c->device; }

struct sehf_frp {
    sehf_ctx* ctx;
    int N;
    DevArray<double> W, U, R, asym;
    GjBuffers gj;
    int sym_request = 0;
    sehf_frp_stats_t stats{};
};

struct sehf_sinv {
    sehf_ctx* ctx;
    int Np, Nf;
    DevArray<double> A, B, Bs, Dinv, Sinv, asym;
    GjBuffers gj;
    sehf_sinv_stats_t stats{};
};

struct sehf_bp {
    sehf_ctx* ctx;
    int Nt, Np, Nf;
    DevArray<double> Z, Ztmp, Y, Rinv, Winv;
    sehf_bp_stats_t stats{};
};

struct sehf_dme {
    sehf_ctx* ctx;
    int Np, ns;
    DevArray<double> V, psi, res;
    sehf_dme_stats_t stats{};
};

// Per-phase wall time of a synchronous step: each mark waits for the stream
// and adds the time since the previous mark to a counter (nanoseconds).
struct PhaseClock {
    hipStream_t st;
    std::chrono::steady_clock::time_point t0;
    explicit PhaseClock(hipStream_t s) : st(s), t0(std::chrono::steady_clock::now()) {}
    void mark(uint64_t& acc) {
        SEHF_HIP_CHECK(hipStreamSynchronize(st));
        const auto t = std::chrono::steady_clock::now();
        acc += static_cast<uint64_t>(std::chrono::duration_cast<std::chrono::nanoseconds>(t - t0).count());
        t0 = t;
    }
};

static double download_asym(DevArray<double>& a, const double* X, int n, hipStream_t s) {
    fill(a.p, 1, 0.0, s);
    max_asym(X, n, 0, n, 1, a.p, s);
    double v = 0.0;
    SEHF_HIP_CHECK(hipMemcpyAsync(&v, a.p, sizeof(double), hipMemcpyDeviceToHost, s));
    SEHF_HIP_CHECK(hipStreamSynchronize(s));
    return v;
}

extern "C" {

const char* sehf_last_error(void) { return g_last_error.c_str(); }

int sehf_gpu_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n > 0 ? 1 : 0;
}

int sehf_ctx_create(int prefer_gpu, sehf_ctx** out) {
    return guarded([&] {
        if (!sehf_gpu_available())
            throw std::runtime_error("sehf_ctx_create: no HIP device visible (this backend has no host fallback)");
        (void)prefer_gpu;
        auto* c = new sehf_ctx();
        SEHF_HIP_CHECK(hipGetDevice(&c->device));
        hipDeviceProp_t prop;
        SEHF_HIP_CHECK(hipGetDeviceProperties(&prop, c->device));
        c->name = prop.name[0] ? std::string(prop.name) : std::string(prop.gcnArchName);
        c->mem = prop.totalGlobalMem;
        SEHF_HIP_CHECK(hipStreamCreate(&c->stream));
        *out = c;
    });
}

void sehf_ctx_destroy(sehf_ctx* ctx) {
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int sehf_ctx_info(const sehf_ctx* ctx, char* device_name, int name_len, size_t* global_mem_bytes,
                  int* is_gpu) {
    return guarded([&] {
        if (device_name && name_len > 0) {
            std::strncpy(device_name, ctx->name.c_str(), name_len - 1);
            device_name[name_len - 1] = '\0';
        }
        if (global_mem_bytes) *global_mem_bytes = ctx->mem;
        if (is_gpu) *is_gpu = ctx->is_gpu;
    });
}

// ---------------- forward stepper ----------------
int sehf_frp_create(sehf_ctx* ctx, int N_total, sehf_frp** out) {
    return guarded([&] {
        if (N_total <= 0) throw std::runtime_error("sehf_frp_create: N_total must be > 0");
        auto* s = new sehf_frp{ctx, N_total};
        const size_t nn = static_cast<size_t>(N_total) * N_total;
        s->W.alloc(nn); s->U.alloc(nn); s->R.alloc(nn); s->asym.alloc(1);
        s->gj.ensure(N_total, 1);
        fill(s->R.p, nn, 0.0, ctx->stream);
        *out = s;
    });
}
int sehf_frp_init_R_prev_inv(sehf_frp* s, const double* R) {
    return guarded([&] {
        const size_t nn = static_cast<size_t>(s->N) * s->N;
        SEHF_HIP_CHECK(hipMemcpyAsync(s->R.p, R, nn * sizeof(double), hipMemcpyHostToDevice, s->ctx->stream));
        SEHF_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
    });
}
int sehf_frp_step(sehf_frp* s, const double* W_inv, double* Rinv_out, double* sym_err) {
    return guarded([&] {
        hipStream_t st = s->ctx->stream;
        const int N = s->N;
        const size_t nn = static_cast<size_t>(N) * N;
        PhaseClock pc(st);
        SEHF_HIP_CHECK(hipMemcpyAsync(s->W.p, W_inv, nn * sizeof(double), hipMemcpyHostToDevice, st));
        pc.mark(s->stats.t_upload_ns);
        // U = 12 W_inv - 10 I - Rinv_{n-1}
        axpbyi(s->U.p, N, 0, s->W.p, N, 0, 12.0, s->R.p, N, 0, -1.0, -10.0, N, N, 1, st);
        pc.mark(s->stats.t_u_combine_ns);
        fill_int(s->gj.info.p, 1, 0, st);
        gj_inverse(s->U.p, s->R.p, N, N, 0, 1, true, s->gj.work(), st);
        check_gj_info(s->gj, 1, st, "sehf_frp_step");
        pc.mark(s->stats.t_inverse_ns);
        if (Rinv_out)
            SEHF_HIP_CHECK(hipMemcpyAsync(Rinv_out, s->R.p, nn * sizeof(double), hipMemcpyDeviceToHost, st));
        const double a = download_asym(s->asym, s->R.p, N, st);
        if (sym_err) *sym_err = a;
        pc.mark(s->stats.t_download_ns);
        ++s->stats.n_steps;
    });
}
int sehf_frp_enable_symmetric_inverse(sehf_frp* s, int enable) {
    return guarded([&] { s->sym_request = enable ? 1 : 0; });
}
int sehf_frp_is_symmetric_inverse_active(const sehf_frp*) { return 0; }
int sehf_frp_stats(const sehf_frp* s, sehf_frp_stats_t* out) {
    return guarded([&] { *out = s->stats; });
}
int sehf_frp_reset_stats(sehf_frp* s) {
    return guarded([&] { s->stats = sehf_frp_stats_t{}; });
}
void sehf_frp_destroy(sehf_frp* s) { delete s; }

// ---------------- Sinv stepper ----------------
int sehf_sinv_create(sehf_ctx* ctx, int N_psi, int N_f, sehf_sinv** out) {
    return guarded([&] {
        if (N_psi <= 0 || N_f < 0) throw std::runtime_error("sehf_sinv_create: bad sizes");
        auto* s = new sehf_sinv{ctx, N_psi, N_f};
        const size_t pp = static_cast<size_t>(N_psi) * N_psi, pf = static_cast<size_t>(N_psi) * std::max(N_f, 1);
        s->A.alloc(pp); s->Sinv.alloc(pp); s->B.alloc(pf); s->Bs.alloc(pf);
        s->Dinv.alloc(std::max(N_f, 1)); s->asym.alloc(1);
        s->gj.ensure(N_psi, 1);
        *out = s;
    });
}
static void sinv_common(sehf_sinv* s, double* Sinv_out, double* sym_err, PhaseClock& pc) {
    hipStream_t st = s->ctx->stream;
    const int Np = s->Np;
    fill_int(s->gj.info.p, 1, 0, st);
    gj_inverse(s->A.p, s->Sinv.p, Np, Np, 0, 1, true, s->gj.work(), st);
    check_gj_info(s->gj, 1, st, "sehf_sinv_step");
    pc.mark(s->stats.t_inverse_ns);
    if (Sinv_out)
        SEHF_HIP_CHECK(hipMemcpyAsync(Sinv_out, s->Sinv.p, sizeof(double) * Np * Np, hipMemcpyDeviceToHost, st));
    const double a = download_asym(s->asym, s->Sinv.p, Np, st);
    if (sym_err) *sym_err = a;
    pc.mark(s->stats.t_download_ns);
    ++s->stats.n_steps;
}
int sehf_sinv_step(sehf_sinv* s, const double* A, const double* B, const double* Dinv, double* Sinv_out,
                   double* sym_err) {
    return guarded([&] {
        hipStream_t st = s->ctx->stream;
        const int Np = s->Np, Nf = s->Nf;
        PhaseClock pc(st);
        SEHF_HIP_CHECK(hipMemcpyAsync(s->A.p, A, sizeof(double) * Np * Np, hipMemcpyHostToDevice, st));
        if (Nf > 0) {
            SEHF_HIP_CHECK(hipMemcpyAsync(s->B.p, B, sizeof(double) * Np * Nf, hipMemcpyHostToDevice, st));
            SEHF_HIP_CHECK(hipMemcpyAsync(s->Dinv.p, Dinv, sizeof(double) * Nf, hipMemcpyHostToDevice, st));
        }
        pc.mark(s->stats.t_upload_ns);
        if (Nf > 0) {
            col_scale(s->Bs.p, Np, s->B.p, Np, s->Dinv.p, Np, Nf, st);
            Gemm g;   // S = A - (B D^{-1}) B^T
            g.M = Np; g.N = Np; g.K = Nf;
            g.A = s->Bs.p; g.lda = Np; g.B = s->B.p; g.ldb = Np; g.tB = 1;
            g.C = s->A.p; g.ldc = Np; g.alpha = -1.0; g.beta = 1.0;
            gemm(g, st);
        }
        pc.mark(s->stats.t_schur_ns);
        sinv_common(s, Sinv_out, sym_err, pc);
    });
}
int sehf_sinv_step_inverse_only(sehf_sinv* s, const double* A, double* Sinv_out, double* sym_err) {
    return guarded([&] {
        hipStream_t st = s->ctx->stream;
        PhaseClock pc(st);
        SEHF_HIP_CHECK(hipMemcpyAsync(s->A.p, A, sizeof(double) * s->Np * s->Np, hipMemcpyHostToDevice, st));
        pc.mark(s->stats.t_upload_ns);
        sinv_common(s, Sinv_out, sym_err, pc);
    });
}
int sehf_sinv_stats(const sehf_sinv* s, sehf_sinv_stats_t* out) {
    return guarded([&] { *out = s->stats; });
}
int sehf_sinv_reset_stats(sehf_sinv* s) {
    return guarded([&] { s->stats = sehf_sinv_stats_t{}; });
}
void sehf_sinv_destroy(sehf_sinv* s) { delete s; }

// ---------------- back stepper ----------------
int sehf_bp_create(sehf_ctx* ctx, int N_total, int N_psi, int N_f, sehf_bp** out) {
    return guarded([&] {
        if (N_total != N_psi + N_f || N_psi <= 0) throw std::runtime_error("sehf_bp_create: N_total != N_psi + N_f");
        auto* s = new sehf_bp{ctx, N_total, N_psi, N_f};
        const size_t tp = static_cast<size_t>(N_total) * N_psi, tt = static_cast<size_t>(N_total) * N_total;
        s->Z.alloc(tp); s->Ztmp.alloc(tp); s->Y.alloc(tp); s->Rinv.alloc(tt); s->Winv.alloc(tt);
        fill(s->Z.p, tp, 0.0, ctx->stream);
        *out = s;
    });
}
int sehf_bp_init_Z(sehf_bp* s, const double* Z) {
    return guarded([&] {
        SEHF_HIP_CHECK(hipMemcpyAsync(s->Z.p, Z, sizeof(double) * s->Nt * s->Np, hipMemcpyHostToDevice, s->ctx->stream));
        SEHF_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
    });
}
int sehf_bp_step(sehf_bp* s, const double* Rinv, const double* W_inv, double* psi_out, double* f_out,
                 int compute_f) {
    return guarded([&] {
        hipStream_t st = s->ctx->stream;
        const int Nt = s->Nt, Np = s->Np, Nf = s->Nf;
        const size_t tt = static_cast<size_t>(Nt) * Nt;
        PhaseClock pc(st);
        SEHF_HIP_CHECK(hipMemcpyAsync(s->Rinv.p, Rinv, tt * sizeof(double), hipMemcpyHostToDevice, st));
        SEHF_HIP_CHECK(hipMemcpyAsync(s->Winv.p, W_inv, tt * sizeof(double), hipMemcpyHostToDevice, st));
        pc.mark(s->stats.t_upload_ns);
        Gemm g;
        g.M = Nt; g.N = Np; g.K = Nt;
        g.A = s->Rinv.p; g.lda = Nt; g.B = s->Z.p; g.ldb = Nt; g.C = s->Ztmp.p; g.ldc = Nt;
        gemm(g, st);
        std::swap(s->Z.p, s->Ztmp.p);
        pc.mark(s->stats.t_gemm_z_ns);
        g.A = s->Winv.p; g.B = s->Z.p; g.C = s->Y.p;
        gemm(g, st);
        pc.mark(s->stats.t_gemm_y_ns);
        if (psi_out)
            SEHF_HIP_CHECK(hipMemcpy2DAsync(psi_out, sizeof(double) * Np, s->Y.p, sizeof(double) * Nt,
                                            sizeof(double) * Np, Np, hipMemcpyDeviceToHost, st));
        if (compute_f && f_out && Nf > 0)
            SEHF_HIP_CHECK(hipMemcpy2DAsync(f_out, sizeof(double) * Nf, s->Y.p + Np, sizeof(double) * Nt,
                                            sizeof(double) * Nf, Np, hipMemcpyDeviceToHost, st));
        pc.mark(s->stats.t_extract_ns);
        ++s->stats.n_steps;
    });
}
int sehf_bp_stats(const sehf_bp* s, sehf_bp_stats_t* out) {
    return guarded([&] { *out = s->stats; });
}
int sehf_bp_reset_stats(sehf_bp* s) {
    return guarded([&] { s->stats = sehf_bp_stats_t{}; });
}
int sehf_bp_get_Z(const sehf_bp* s, double* Z_out) {
    return guarded([&] {
        SEHF_HIP_CHECK(hipMemcpyAsync(Z_out, s->Z.p, sizeof(double) * s->Nt * s->Np, hipMemcpyDeviceToHost, s->ctx->stream));
        SEHF_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
    });
}
void sehf_bp_destroy(sehf_bp* s) { delete s; }

// ---------------- dipole engine ----------------
int sehf_dme_create(sehf_ctx* ctx, int N_psi, int n_slots, sehf_dme** out) {
    return guarded([&] {
        if (N_psi <= 0 || n_slots <= 0) throw std::runtime_error("sehf_dme_create: bad sizes");
        auto* s = new sehf_dme{ctx, N_psi, n_slots};
        s->V.alloc(static_cast<size_t>(n_slots) * N_psi);
        s->psi.alloc(static_cast<size_t>(N_psi) * N_psi);
        s->res.alloc(static_cast<size_t>(n_slots) * N_psi);
        *out = s;
    });
}
int sehf_dme_step(sehf_dme* s, const double* V, const double* psi_ir, double* result_out) {
    return guarded([&] {
        hipStream_t st = s->ctx->stream;
        const int Np = s->Np, ns = s->ns;
        PhaseClock pc(st);
        SEHF_HIP_CHECK(hipMemcpyAsync(s->V.p, V, sizeof(double) * ns * Np, hipMemcpyHostToDevice, st));
        SEHF_HIP_CHECK(hipMemcpyAsync(s->psi.p, psi_ir, sizeof(double) * Np * Np, hipMemcpyHostToDevice, st));
        pc.mark(s->stats.t_upload_ns);
        Gemm g;
        g.M = ns; g.N = Np; g.K = Np;
        g.A = s->V.p; g.lda = ns; g.B = s->psi.p; g.ldb = Np; g.C = s->res.p; g.ldc = ns;
        gemm(g, st);
        pc.mark(s->stats.t_gemm_ns);
        SEHF_HIP_CHECK(hipMemcpyAsync(result_out, s->res.p, sizeof(double) * ns * Np, hipMemcpyDeviceToHost, st));
        pc.mark(s->stats.t_download_ns);
        ++s->stats.n_steps;
    });
}
int sehf_dme_stats(const sehf_dme* s, sehf_dme_stats_t* out) {
    return guarded([&] { *out = s->stats; });
}
int sehf_dme_reset_stats(sehf_dme* s) {
    return guarded([&] { s->stats = sehf_dme_stats_t{}; });
}
void sehf_dme_destroy(sehf_dme* s) { delete s; }

}  // extern "C"
