// gemm_probe.cpp -- profiling harness for k_dgemm on the scan's shapes.
//
// Runs the product GEMM (sehf::gemm, the same launcher the scan uses) on the
// block-update shapes of the Gauss-Jordan inverse of order 1566 and on the
// back-propagation products at l_cont = 30, batch 32, a few times each, after
// a device-to-device copy of a known size.  Under rocprofv3 --pmc the copy is
// the calibration reference of the HBM byte counters (it reads and writes
// exactly `calib_mib` MiB); under --kernel-trace the dispatch durations give
// each shape's rate.  Prints one line per shape with its HIP-event rate.
//
//   gemm_probe [batch=32] [reps=3] [calib_mib=1024]
//   gemm_probe gj [batch=32] [reps=3]   -- the batched Gauss-Jordan inverse of
//                                          orders 1566 and 961 instead
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "sehf_common.h"
#include "sehf_kernels.h"

using namespace sehf;

// Batched Gauss-Jordan inverse of random matrices (uniform entries, diagonal
// raised by 2: generic pivoting, condition ~ n), timed with HIP events; a
// sample of |A X - I| of the first matrix is printed as a check.
static int gj_main(int batch, int reps) {
    for (int n : {1566, 961}) {
        const size_t per = static_cast<size_t>(n) * n;
        std::vector<double> h(per * batch);
        uint64_t st = 12345;
        for (size_t i = 0; i < h.size(); ++i) {
            st = st * 6364136223846793005ull + 1442695040888963407ull;
            h[i] = static_cast<double>(st >> 11) / 9007199254740992.0 - 0.5;
        }
        for (int b = 0; b < batch; ++b)
            for (int i = 0; i < n; ++i) h[b * per + i + static_cast<size_t>(i) * n] += 2.0;
        DevArray<double> A(per * batch), W(per * batch), X(per * batch);
        SEHF_HIP_CHECK(hipMemcpy(A.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
        GjBuffers gb;
        gb.ensure(n, batch);
        hipEvent_t e0, e1;
        SEHF_HIP_CHECK(hipEventCreate(&e0));
        SEHF_HIP_CHECK(hipEventCreate(&e1));
        float total = 0.0f;
        for (int r = 0; r <= reps; ++r) {   // r = 0 warms up
            SEHF_HIP_CHECK(hipMemcpy(W.p, A.p, per * batch * sizeof(double), hipMemcpyDeviceToDevice));
            SEHF_HIP_CHECK(hipEventRecord(e0, nullptr));
            gj_inverse(W.p, X.p, n, n, static_cast<int64_t>(per), batch, false, gb.work(), nullptr);
            SEHF_HIP_CHECK(hipEventRecord(e1, nullptr));
            SEHF_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.0f;
            SEHF_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            if (r > 0) total += ms;
        }
        std::vector<double> x(per);
        SEHF_HIP_CHECK(hipMemcpy(x.data(), X.p, per * sizeof(double), hipMemcpyDeviceToHost));
        double worst = 0.0;
        for (int j = 0; j < n; j += 7)
            for (int i = 0; i < n; i += 5) {
                double acc = 0.0;
                for (int q = 0; q < n; ++q) acc += h[i + static_cast<size_t>(q) * n] * x[q + static_cast<size_t>(j) * n];
                worst = std::max(worst, std::fabs(acc - (i == j ? 1.0 : 0.0)));
            }
        const double ms = total / reps;
        std::printf("Gauss-Jordan inverse n=%d batch %d: %.3f ms  %.1f TFLOP/s (2n^3)  max|AX-I| sampled %.1e\n", n,
                    batch, ms, 2.0 * n * static_cast<double>(n) * n * batch / (ms * 1e-3) / 1e12, worst);
        std::fflush(stdout);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    return 0;
}

int main(int argc, char** argv) {
    try {
        if (argc > 1 && std::string(argv[1]) == "gj") {
            const int batch = argc > 2 ? std::atoi(argv[2]) : 32;
            const int reps = argc > 3 ? std::atoi(argv[3]) : 3;
            if (batch < 1 || batch > 64 || reps < 1) throw std::runtime_error("usage: gemm_probe gj [batch<=64] [reps]");
            return gj_main(batch, reps);
        }
        const int batch = argc > 1 ? std::atoi(argv[1]) : 32;
        const int reps = argc > 2 ? std::atoi(argv[2]) : 3;
        const size_t calib = static_cast<size_t>(argc > 3 ? std::atoi(argv[3]) : 1024) << 20;
        const int n = 1566, np = 961;
        const size_t per = static_cast<size_t>(n) * n;
        if (batch < 1 || batch > 64 || reps < 1) throw std::runtime_error("usage: gemm_probe [batch<=64] [reps] [calib_mib]");
        if (calib > 0) {
            DevArray<char> a(calib), b(calib);
            SEHF_HIP_CHECK(hipMemset(a.p, 1, calib));
            SEHF_HIP_CHECK(hipDeviceSynchronize());
            SEHF_HIP_CHECK(hipMemcpy(b.p, a.p, calib, hipMemcpyDeviceToDevice));
            SEHF_HIP_CHECK(hipDeviceSynchronize());
            std::printf("calibration copy: %zu MiB read + %zu MiB written\n", calib >> 20, calib >> 20);
        }
        DevArray<double> A(per * batch), B(per * batch), C(per * batch);
        fill(A.p, A.n, 0.5, nullptr);
        fill(B.p, B.n, 0.25, nullptr);
        fill(C.p, C.n, 0.0, nullptr);
        SEHF_HIP_CHECK(hipDeviceSynchronize());
        struct Shape { int M, N, K; int tA; int64_t lda, ldb, ldc; double beta; const char* name; };
        const Shape shapes[] = {
            {783, n, 783, 0, 783, n, n, 1.0, "Gauss-Jordan update 783 x 1566 x 783"},
            {392, n, 391, 0, 392, n, n, 1.0, "Gauss-Jordan update 392 x 1566 x 391"},
            {196, n, 196, 0, 196, n, n, 1.0, "Gauss-Jordan update 196 x 1566 x 196"},
            {98, n, 98, 0, 98, n, n, 1.0, "Gauss-Jordan update 98 x 1566 x 98"},
            {48, n, 48, 0, 48, n, n, 1.0, "Gauss-Jordan update 48 x 1566 x 48"},
            {n, np, n, 0, n, n, n, 0.0, "back-propagation Z <- Rinv Z 1566 x 961 x 1566"},
            {np, np, np, 0, np, np, np, 0.0, "back-propagation psi = Sinv rhs 961^3"},
        };
        hipEvent_t e0, e1;
        SEHF_HIP_CHECK(hipEventCreate(&e0));
        SEHF_HIP_CHECK(hipEventCreate(&e1));
        for (const Shape& s : shapes) {
            Gemm g;
            g.M = s.M; g.N = s.N; g.K = s.K;
            g.A = A.p; g.lda = s.lda; g.sA = static_cast<int64_t>(per); g.tA = s.tA;
            g.B = B.p; g.ldb = s.ldb; g.sB = static_cast<int64_t>(per);
            g.C = C.p; g.ldc = s.ldc; g.sC = static_cast<int64_t>(per);
            g.beta = s.beta; g.batch = batch;
            gemm(g, nullptr);   // warm-up
            SEHF_HIP_CHECK(hipEventRecord(e0, nullptr));
            for (int r = 0; r < reps; ++r) gemm(g, nullptr);
            SEHF_HIP_CHECK(hipEventRecord(e1, nullptr));
            SEHF_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.0f;
            SEHF_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            const double us = ms * 1e3 / reps;
            const double fl = 2.0 * s.M * s.N * static_cast<double>(s.K) * batch;
            const double bytes = 8.0 * batch * (static_cast<double>(s.M) * s.K + static_cast<double>(s.K) * s.N +
                                                (s.beta != 0.0 ? 2.0 : 1.0) * s.M * s.N);
            std::printf("%-50s batch %d: %9.1f us  %6.1f TFLOP/s  algorithmic bytes %.3e\n", s.name, batch, us,
                        fl / (us * 1e-6) / 1e12, bytes);
            std::fflush(stdout);
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "gemm_probe: %s\n", e.what());
        return 1;
    }
}
