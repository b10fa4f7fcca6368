// sehf_common.h -- host-side helpers shared by the C ABI and the scan engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "sehf_kernels.h"

#define SEHF_HIP_CHECK(expr)                                                                      \
    do {                                                                                          \
        hipError_t _e = (expr);                                                                   \
        if (_e != hipSuccess)                                                                     \
            throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(_e) + " at " +  \
                                     __FILE__ + ":" + std::to_string(__LINE__) + " (" #expr ")"); \
    } while (0)

namespace sehf {

// Owning device allocation, or pinned host memory (alloc_host: page-locked,
// mapped into the device's address space, so kernels and copies reach it).
template <class T>
struct DevArray {
    T* p = nullptr;
    size_t n = 0;
    bool host = false;
    DevArray() = default;
    explicit DevArray(size_t count) { alloc(count); }
    DevArray(const DevArray&) = delete;
    DevArray& operator=(const DevArray&) = delete;
    DevArray(DevArray&& o) noexcept : p(o.p), n(o.n), host(o.host) { o.p = nullptr; o.n = 0; }
    DevArray& operator=(DevArray&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; host = o.host; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevArray() { release(); }
    void alloc(size_t count) {
        release();
        host = false;
        if (count == 0) return;
        SEHF_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
        n = count;
    }
    void alloc_host(size_t count) {
        release();
        host = true;
        if (count == 0) return;
        SEHF_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&p), count * sizeof(T), hipHostMallocDefault));
        n = count;
    }
    void release() {
        if (p) (void)(host ? hipHostFree(p) : hipFree(p));
        p = nullptr; n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Gauss-Jordan workspace for `batch` matrices of order n.
struct GjBuffers {
    DevArray<int> piv, perm, info, nom, stgi, cnt;
    DevArray<double> tbuf, stg;
    int n = 0, batch = 0;
    void ensure(int n_, int batch_) {
        if (n_ <= n && batch_ <= batch) return;
        n = std::max(n, n_); batch = std::max(batch, batch_);
        piv.alloc(static_cast<size_t>(n) * batch);
        perm.alloc(static_cast<size_t>(n) * batch);
        info.alloc(static_cast<size_t>(batch));
        tbuf.alloc(gj_tbuf_elems(n) * batch);
        nom.alloc(gj_nom_elems(n) * batch);
        stg.alloc(gj_stg_elems() * batch);
        stgi.alloc(gj_stgi_elems() * batch);
        cnt.alloc(static_cast<size_t>(batch));
        fill_int(cnt.p, static_cast<size_t>(batch), 0, nullptr);
        fill_int(info.p, static_cast<size_t>(batch), 0, nullptr);
        SEHF_HIP_CHECK(hipDeviceSynchronize());
    }
    GjWork work() const { return GjWork{piv.p, perm.p, tbuf.p, info.p, nom.p, stg.p, stgi.p, cnt.p}; }
};

// Throw when any matrix of the last inversion met an exactly-zero pivot
// (the reference throws on dgetrf info != 0, LapackInverse.hpp:69-73).
void check_gj_info(const GjBuffers& gb, int batch, hipStream_t s, const char* what);

}  // namespace sehf
