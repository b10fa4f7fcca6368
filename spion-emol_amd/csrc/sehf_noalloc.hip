// sehf_noalloc.hip -- the forward ratio step for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc).
//
// sehf_frp_step_noalloc runs the kernels of sehf_frp_step (U = 12 W_inv - 10 I
// - R_prev, then the Gauss-Jordan inverse with its symmetrisation) on
// caller-owned device buffers.  Its workspace -- the matrix U that the
// inverse works on in place, the Gauss-Jordan block copies, pivot records,
// leaf nominees, staging, completion counters and zero-pivot flag, and the
// symmetry-error cell -- lives in one file-scope device array, `scratch`, of
// fixed size (orders up to kNoallocMaxN).  Each device holds its own copy of
// the array for the life of the process; its address on a device comes from
// hipGetSymbolAddress, looked up once per device with that device current.
// All calls on one device share that copy: calls that overlap on one device
// race (they must be serialised by the caller).
#include "../../include/sehf.h"

#include <algorithm>
#include <mutex>
#include <string>

#include "sehf_common.h"

namespace sehf {
std::string& last_error_ref();
}

using namespace sehf;

namespace {

constexpr int kNoallocMaxN = SEHF_FRP_NOALLOC_MAX_N;
constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Byte offsets of the workspace pieces for an order-n step.
struct Layout {
    size_t U, tbuf, piv, perm, nom, stg, stgi, cnt, info, asym, end;
    explicit Layout(int n) {
        const size_t nn = static_cast<size_t>(n) * n;
        size_t o = 0;
        U = o;    o = align_up(o + nn * sizeof(double));
        tbuf = o; o = align_up(o + gj_tbuf_elems(n) * sizeof(double));
        piv = o;  o = align_up(o + static_cast<size_t>(n) * sizeof(int));
        perm = o; o = align_up(o + static_cast<size_t>(n) * sizeof(int));
        nom = o;  o = align_up(o + gj_nom_elems(n) * sizeof(int));
        stg = o;  o = align_up(o + gj_stg_elems() * sizeof(double));
        stgi = o; o = align_up(o + gj_stgi_elems() * sizeof(int));
        cnt = o;  o = align_up(o + sizeof(int));
        info = o; o = align_up(o + sizeof(int));
        asym = o; o = align_up(o + sizeof(double));
        end = o;
    }
};

// Size of the array: the layout of the largest order, computed the same way.
constexpr size_t scratch_bytes() {
    const size_t n = kNoallocMaxN, nn = n * n;
    size_t o = 0;
    o = align_up(o + nn * sizeof(double));                            // U
    o = align_up(o + nn * sizeof(double));                            // tbuf (n^2 doubles)
    o = align_up(o + n * sizeof(int));                                // piv
    o = align_up(o + n * sizeof(int));                                // perm
    o = align_up(o + ((n + 255) / 256) * 16 * sizeof(int));          // leaf nominees
    o = align_up(o + (4 * 16 * 16 + 2 * 2 * 16 * 16) * sizeof(double));   // leaf staging
    o = align_up(o + (4 * 16 + 4) * sizeof(int));                     // leaf move list
    o = align_up(o + sizeof(int));                                    // counter
    o = align_up(o + sizeof(int));                                    // zero-pivot flag
    o = align_up(o + sizeof(double));                                 // symmetry error
    return o;
}

}  // namespace

// The workspace array (one copy per device).
static __device__ __align__(256) unsigned char scratch[scratch_bytes()];

namespace {

std::mutex g_mu;
constexpr int kMaxDevices = 64;
unsigned char* g_scratch[kMaxDevices] = {};

// Device address of this device's copy of `scratch` (looked up once per
// device, with that device current).  The array's name in host code is only
// a host-side stand-in: every use below goes through this address.
unsigned char* scratch_on(int dev) {
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("sehf_frp_step_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        SEHF_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw std::runtime_error("sehf_frp_step_noalloc: no device address for the scratch array");
        g_scratch[dev] = static_cast<unsigned char*>(p);
    }
    return g_scratch[dev];
}

template <class F>
int guarded_na(F&& f) {
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

}  // namespace

extern "C" {

int sehf_frp_step_noalloc(int N_total, const double* W_inv, const double* R_prev, double* Rinv_out, double* sym_err) {
    return guarded_na([&] {
        static_assert(scratch_bytes() >= 2 * sizeof(double) * kNoallocMaxN * kNoallocMaxN, "scratch size");
        if (N_total <= 0 || N_total > kNoallocMaxN)
            throw std::runtime_error("sehf_frp_step_noalloc: N_total = " + std::to_string(N_total) +
                                     " outside [1, " + std::to_string(kNoallocMaxN) +
                                     "] (the size of the static workspace)");
        if (!W_inv || !R_prev || !Rinv_out)
            throw std::runtime_error("sehf_frp_step_noalloc: W_inv, R_prev and Rinv_out must be device pointers");
        const int N = N_total;
        const Layout lay(N);
        if (lay.end > scratch_bytes()) throw std::runtime_error("sehf_frp_step_noalloc: workspace layout overflow");
        int dev = 0;
        SEHF_HIP_CHECK(hipGetDevice(&dev));
        unsigned char* base = scratch_on(dev);
        hipStream_t st = nullptr;   // the default stream: inputs must be complete before the call
        double* U = reinterpret_cast<double*>(base + lay.U);
        int* cnt = reinterpret_cast<int*>(base + lay.cnt);
        int* info = reinterpret_cast<int*>(base + lay.info);
        double* asym = reinterpret_cast<double*>(base + lay.asym);
        GjWork w{reinterpret_cast<int*>(base + lay.piv), reinterpret_cast<int*>(base + lay.perm),
                 reinterpret_cast<double*>(base + lay.tbuf), info, reinterpret_cast<int*>(base + lay.nom),
                 reinterpret_cast<double*>(base + lay.stg), reinterpret_cast<int*>(base + lay.stgi), cnt};
        // the array keeps the last call's contents: the counters and the flag
        // must start at zero
        fill_int(cnt, 1, 0, st);
        fill_int(info, 1, 0, st);
        fill(asym, 1, 0.0, st);
        // U = 12 W_inv - 10 I - R_prev, then Rinv = sym(U^{-1}) (sehf_frp_step's kernels)
        axpbyi(U, N, 0, W_inv, N, 0, 12.0, R_prev, N, 0, -1.0, -10.0, N, N, 1, st);
        gj_inverse(U, Rinv_out, N, N, 0, 1, true, w, st);
        max_asym(Rinv_out, N, 0, N, 1, asym, st);
        int flag = 0;
        double a = 0.0;
        SEHF_HIP_CHECK(hipMemcpyAsync(&flag, info, sizeof(int), hipMemcpyDeviceToHost, st));
        SEHF_HIP_CHECK(hipMemcpyAsync(&a, asym, sizeof(double), hipMemcpyDeviceToHost, st));
        SEHF_HIP_CHECK(hipStreamSynchronize(st));
        if (flag != 0)
            throw std::runtime_error("sehf_frp_step_noalloc: singular matrix (zero pivot at index " +
                                     std::to_string(flag - 1) + ")");
        if (sym_err) *sym_err = a;
    });
}

}  // extern "C"
