// sehf_timing.cpp -- timing entry points for the path's two inner kernels
// (the batched FP64 GEMM and the batched Gauss-Jordan inverse) on
// synthetic operands of chosen shapes, and a device copy of known size
// (the reference of HBM byte counters).  They run on the default stream,
// with HIP events around the timed repetitions, and are what the benchmark
// and the hardware-counter runs use to look at one kernel shape at a time.
#include "../../include/sehf.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "sehf_common.h"

namespace sehf {
std::string& last_error_ref();
}

using namespace sehf;

namespace {

template <class F>
int guarded_t(F&& f) {
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

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    EventPair() {
        SEHF_HIP_CHECK(hipEventCreate(&a));
        SEHF_HIP_CHECK(hipEventCreate(&b));
    }
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    double ms() const {
        SEHF_HIP_CHECK(hipEventSynchronize(b));
        float t = 0.0f;
        SEHF_HIP_CHECK(hipEventElapsedTime(&t, a, b));
        return static_cast<double>(t);
    }
};

// Uniform pseudo-random entries in [-0.5, 0.5) (64-bit LCG).
void lcg_fill(std::vector<double>& h, uint64_t seed) {
    uint64_t st = seed;
    for (double& v : h) {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        v = static_cast<double>(st >> 11) / 9007199254740992.0 - 0.5;
    }
}

}  // namespace

extern "C" {

int sehf_time_gemm(int M, int N, int K, int lda, int ldb, int ldc, int batch, int beta_one, int reps, double* ms_out) {
    return guarded_t([&] {
        if (M <= 0 || N <= 0 || K <= 0 || batch <= 0 || reps <= 0)
            throw std::runtime_error("sehf_time_gemm: sizes, batch and reps must be > 0");
        if (lda < M || ldb < K || ldc < M) throw std::runtime_error("sehf_time_gemm: leading dimension below the rows");
        const size_t a = static_cast<size_t>(lda) * K, b = static_cast<size_t>(ldb) * N, c = static_cast<size_t>(ldc) * N;
        DevArray<double> A(a * batch), B(b * batch), C(c * batch);
        fill(A.p, A.n, 0.5, nullptr);
        fill(B.p, B.n, 0.25, nullptr);
        fill(C.p, C.n, 0.0, nullptr);
        Gemm g;
        g.M = M; g.N = N; g.K = K;
        g.A = A.p; g.lda = lda; g.sA = static_cast<int64_t>(a);
        g.B = B.p; g.ldb = ldb; g.sB = static_cast<int64_t>(b);
        g.C = C.p; g.ldc = ldc; g.sC = static_cast<int64_t>(c);
        g.beta = beta_one ? 1.0 : 0.0; g.batch = batch;
        gemm(g, nullptr);   // warm-up
        EventPair ev;
        SEHF_HIP_CHECK(hipEventRecord(ev.a, nullptr));
        for (int r = 0; r < reps; ++r) gemm(g, nullptr);
        SEHF_HIP_CHECK(hipEventRecord(ev.b, nullptr));
        *ms_out = ev.ms() / reps;
    });
}

int sehf_time_inverse(int n, int batch, int reps, double* ms_out, double* resid_out) {
    return guarded_t([&] {
        if (n <= 0 || batch <= 0 || reps <= 0) throw std::runtime_error("sehf_time_inverse: n, batch and reps must be > 0");
        const size_t per = static_cast<size_t>(n) * n;
        // random entries with the diagonal raised by 2: generic pivoting,
        // condition ~ n
        std::vector<double> h(per * batch);
        lcg_fill(h, 12345);
        for (int bb = 0; bb < batch; ++bb)
            for (int i = 0; i < n; ++i) h[bb * per + i + static_cast<size_t>(i) * n] += 2.0;
        DevArray<double> A(per * batch), W(per * batch), X(per * batch);
        SEHF_HIP_CHECK(hipMemcpy(A.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
        GjBuffers gb;
        gb.ensure(n, batch);
        double total = 0.0;
        for (int r = 0; r <= reps; ++r) {   // r = 0 warms up
            SEHF_HIP_CHECK(hipMemcpy(W.p, A.p, per * batch * sizeof(double), hipMemcpyDeviceToDevice));
            EventPair ev;
            SEHF_HIP_CHECK(hipEventRecord(ev.a, nullptr));
            gj_inverse(W.p, X.p, n, n, static_cast<int64_t>(per), batch, false, gb.work(), nullptr);
            SEHF_HIP_CHECK(hipEventRecord(ev.b, nullptr));
            const double ms = ev.ms();
            if (r > 0) total += ms;
        }
        check_gj_info(gb, batch, nullptr, "sehf_time_inverse");
        if (resid_out) {   // sampled max |A X - I| of the first matrix
            std::vector<double> x(per);
            SEHF_HIP_CHECK(hipMemcpy(x.data(), X.p, per * sizeof(double), hipMemcpyDeviceToHost));
            double worst = 0.0;
            for (int j = 0; j < n; j += 7)
                for (int i = 0; i < n; i += 5) {
                    double acc = 0.0;
                    for (int q = 0; q < n; ++q) acc += h[i + static_cast<size_t>(q) * n] * x[q + static_cast<size_t>(j) * n];
                    worst = std::max(worst, std::fabs(acc - (i == j ? 1.0 : 0.0)));
                }
            *resid_out = worst;
        }
        *ms_out = total / reps;
    });
}

int sehf_time_spd_inverse(int n, int batch, int reps, double* ms_out, double* resid_out) {
    return guarded_t([&] {
        if (n <= 0 || batch <= 0 || reps <= 0)
            throw std::runtime_error("sehf_time_spd_inverse: n, batch and reps must be > 0");
        const size_t per = static_cast<size_t>(n) * n;
        // I + 0.05 (R + R^T) / sqrt(n), R uniform in [-0.5, 0.5): symmetric
        // positive definite with condition ~1.1, like the Schur matrices
        std::vector<double> r(per), h(per * batch);
        lcg_fill(r, 777);
        const double sc = 0.05 / std::sqrt(static_cast<double>(n));
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                h[i + static_cast<size_t>(j) * n] = (i == j ? 1.0 : 0.0) +
                                                    sc * (r[i + static_cast<size_t>(j) * n] + r[j + static_cast<size_t>(i) * n]);
        for (int bb = 1; bb < batch; ++bb) std::copy(h.begin(), h.begin() + per, h.begin() + bb * per);
        DevArray<double> A(per * batch), W(per * batch);
        SEHF_HIP_CHECK(hipMemcpy(A.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
        GjBuffers gb;
        gb.ensure(n, batch);
        double total = 0.0;
        for (int rep = 0; rep <= reps; ++rep) {   // rep 0 warms up
            SEHF_HIP_CHECK(hipMemcpy(W.p, A.p, per * batch * sizeof(double), hipMemcpyDeviceToDevice));
            EventPair ev;
            SEHF_HIP_CHECK(hipEventRecord(ev.a, nullptr));
            spd_inverse(W.p, n, n, static_cast<int64_t>(per), batch, true, gb.work(), nullptr);
            SEHF_HIP_CHECK(hipEventRecord(ev.b, nullptr));
            const double ms = ev.ms();
            if (rep > 0) total += ms;
        }
        check_gj_info(gb, batch, nullptr, "sehf_time_spd_inverse");
        if (resid_out) {   // sampled max |A X - I| of the last matrix
            std::vector<double> x(per);
            SEHF_HIP_CHECK(hipMemcpy(x.data(), W.p + (batch - 1) * per, per * sizeof(double), hipMemcpyDeviceToHost));
            double worst = 0.0;
            for (int j = 0; j < n; j += 7)
                for (int i = 0; i < n; i += 5) {
                    double acc = 0.0;
                    for (int q = 0; q < n; ++q) acc += h[i + static_cast<size_t>(q) * n] * x[q + static_cast<size_t>(j) * n];
                    worst = std::max(worst, std::fabs(acc - (i == j ? 1.0 : 0.0)));
                }
            *resid_out = worst;
        }
        *ms_out = total / reps;
    });
}

int sehf_copy_probe(size_t bytes, double* ms_out) {
    return guarded_t([&] {
        if (bytes == 0) throw std::runtime_error("sehf_copy_probe: bytes must be > 0");
        DevArray<char> a(bytes), b(bytes);
        SEHF_HIP_CHECK(hipMemset(a.p, 1, bytes));
        SEHF_HIP_CHECK(hipDeviceSynchronize());
        EventPair ev;
        SEHF_HIP_CHECK(hipEventRecord(ev.a, nullptr));
        SEHF_HIP_CHECK(hipMemcpyAsync(b.p, a.p, bytes, hipMemcpyDeviceToDevice, nullptr));
        SEHF_HIP_CHECK(hipEventRecord(ev.b, nullptr));
        const double ms = ev.ms();
        if (ms_out) *ms_out = ms;
    });
}

}  // extern "C"
