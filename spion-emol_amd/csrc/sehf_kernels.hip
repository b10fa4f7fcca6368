// sehf_kernels.hip -- gfx950 FP64 kernels of the static-exchange solver.
//
//  * k_dgemm       : batched FP64 GEMM on v_mfma_f64_16x16x4f64, 128x128 output
//                    tile per 8-wave workgroup for deep products, 64x64 per
//                    4 waves for shallow ones, 16x128 / 32x128 for thin ones;
//                    K staged through LDS in slabs of 16; XCD-aware tile order.
//  * Gauss-Jordan  : in-place inversion with partial pivoting, recursive over
//                    column blocks.  Leaves (<=16 columns) pick their pivots by
//                    tournament over blocks of 256 rows (k_leaf_nominate, whose
//                    last block merges) and apply them in a second launch
//                    (k_leaf_apply); the pivot rows' interchanges are applied
//                    in full at once, so no deferred row interchanges exist.
//                    Interior nodes are one MFMA GEMM over all rows (rank-w
//                    block update).
//
// Gauss-Jordan works on the row-major view of the column-major buffer
// (element (i,j) at base[i*ld + j], i.e. on M^T).  Row pivoting of M^T is
// column pivoting of M, and the row-major result of inv(M^T) is inv(M) in
// column-major order, so callers pass and receive ordinary column-major data.
#include "sehf_kernels.h"

#include <algorithm>
#include <atomic>
#include <climits>
#include <mutex>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

namespace sehf {

typedef double d4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// GEMM
// ---------------------------------------------------------------------------
constexpr int GBK = 16;

struct GemmK {
    int M, N, K;
    const double* A; int64_t lda, sA; int tA;
    const double* B; int64_t ldb, sB; int tB;
    double* C; int64_t ldc, sC;
    double alpha, beta;
    const double* alpha_b;
    int zc0, zc1;
};

// BM x BN output tile per workgroup of WM x WN waves; each wave owns a
// (BM/WM) x (BN/WN) sub-tile of 16 x 16 MFMA blocks.  K runs in slabs of 16
// staged through LDS with the next slab's global loads in flight.
// Workgroups are dispatched round-robin over the 8 XCDs; the linear block
// index is remapped so that each XCD works on one contiguous range of tiles
// (neighbouring tiles of the same matrix share A row- and B column-panels in
// that XCD's L2).
// PC: C's old values (beta != 0) are loaded into registers before the K loop
// instead of after it, so their latency overlaps the products; for the small
// tiles of shallow products, whose K loop is a few slabs.
template <int BM, int BN, int WM, int WN, bool PC = false>
__global__ __launch_bounds__(64 * WM * WN) void k_dgemm(GemmK p) {
    constexpr int NT = 64 * WM * WN;
    constexpr int TM = BM / WM, TN = BN / WN, FM = TM / 16, FN = TN / 16;
    constexpr int RA = BM * GBK / NT, RB = BN * GBK / NT;
    static_assert(TM % 16 == 0 && TN % 16 == 0 && RA * NT == BM * GBK && RB * NT == BN * GBK, "tile");
    int bx, by, bz;
    {
        const int gx = gridDim.x, gy = gridDim.y;
        const int T = gx * gy * gridDim.z;
        const int L = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
        const int x = L % 8, idx = L / 8, q = T / 8, r = T % 8;
        const int nl = (x < r) ? x * (q + 1) + idx : r * (q + 1) + (x - r) * q + idx;
        bx = nl % gx; by = (nl / gx) % gy; bz = nl / (gx * gy);
    }
    const int64_t b = bz;
    const double* __restrict__ A = p.A + b * p.sA;
    const double* __restrict__ B = p.B + b * p.sB;
    double* __restrict__ C = p.C + b * p.sC;
    const int m0 = bx * BM, n0 = by * BN;
    __shared__ double As[GBK][BM + 2];   // [k][i] of op(A)
    __shared__ double Bs[GBK][BN + 2];   // [k][j] of op(B)
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int wm = (wv % WM) * TM, wn = (wv / WM) * TN;
    d4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
    double ra[RA], rb[RB];
    double cpre[PC ? FM : 1][PC ? FN : 1][PC ? 4 : 1];
    if constexpr (PC) {
#pragma unroll
        for (int bi = 0; bi < FM; ++bi)
#pragma unroll
            for (int bj = 0; bj < FN; ++bj)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = m0 + wm + bi * 16 + (lane & 15);
                    const int col = n0 + wn + bj * 16 + 4 * v + (lane >> 4);
                    cpre[bi][bj][v] = (p.beta != 0.0 && row < p.M && col < p.N) ? C[row + (int64_t)col * p.ldc] : 0.0;
                }
    }
    // Element assignment of the slabs: coalesced along the contiguous
    // dimension of the stored operand.
    auto a_ik = [&](int r, int& i, int& k) {
        const int e = t + r * NT;
        if (!p.tA) { i = e % BM; k = e / BM; } else { k = e % GBK; i = e / GBK; }
    };
    auto b_kj = [&](int r, int& k, int& j) {
        const int e = t + r * NT;
        if (!p.tB) { k = e % GBK; j = e / GBK; } else { j = e % BN; k = e / BN; }
    };
    auto load = [&](int k0) {
#pragma unroll
        for (int r = 0; r < RA; ++r) {
            int i, k; a_ik(r, i, k);
            const int gi = m0 + i, gk = k0 + k;
            double v = 0.0;
            if (gi < p.M && gk < p.K) v = p.tA ? A[gk + (int64_t)gi * p.lda] : A[gi + (int64_t)gk * p.lda];
            ra[r] = v;
        }
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            int k, j; b_kj(r, k, j);
            const int gj = n0 + j, gk = k0 + k;
            double v = 0.0;
            if (gj < p.N && gk < p.K) v = p.tB ? B[gj + (int64_t)gk * p.ldb] : B[gk + (int64_t)gj * p.ldb];
            rb[r] = v;
        }
    };
    load(0);
    for (int k0 = 0; k0 < p.K; k0 += GBK) {
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RA; ++r) { int i, k; a_ik(r, i, k); As[k][i] = ra[r]; }
#pragma unroll
        for (int r = 0; r < RB; ++r) { int k, j; b_kj(r, k, j); Bs[k][j] = rb[r]; }
        __syncthreads();
        if (k0 + GBK < p.K) load(k0 + GBK);
#pragma unroll
        for (int kk = 0; kk < GBK; kk += 4) {
            const int kq = kk + (lane >> 4);
            double av[FM], bv[FN];
#pragma unroll
            for (int bi = 0; bi < FM; ++bi) av[bi] = As[kq][wm + bi * 16 + (lane & 15)];
#pragma unroll
            for (int bj = 0; bj < FN; ++bj) bv[bj] = Bs[kq][wn + bj * 16 + (lane & 15)];
            // MFMA computes the transposed 16x16 block: its A operand is the
            // op(B) column slab, its B operand the op(A) row slab.  Result
            // register v of lane l holds D(4v + l/16, l%16), i.e.
            // C(row l%16, col 4v + l/16), so stores run along C's contiguous
            // row index.
#pragma unroll
            for (int bi = 0; bi < FM; ++bi)
#pragma unroll
                for (int bj = 0; bj < FN; ++bj)
                    acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[bj], av[bi], acc[bi][bj], 0, 0, 0);
        }
    }
    const double al = p.alpha_b ? p.alpha * p.alpha_b[b] : p.alpha;
#pragma unroll
    for (int bi = 0; bi < FM; ++bi)
#pragma unroll
        for (int bj = 0; bj < FN; ++bj)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = m0 + wm + bi * 16 + (lane & 15);
                const int col = n0 + wn + bj * 16 + 4 * v + (lane >> 4);
                if (row < p.M && col < p.N) {
                    double* c = C + row + (int64_t)col * p.ldc;
                    const double be = (col >= p.zc0 && col < p.zc1) ? 0.0 : p.beta;
                    double val = al * acc[bi][bj][v];
                    if constexpr (PC) {
                        if (be != 0.0) val += be * cpre[bi][bj][v];
                    } else {
                        if (be != 0.0) val += be * (*c);
                    }
                    *c = val;
                }
            }
}

// ---------------------------------------------------------------------------
// Optional per-launch timing (HIP events on the launch stream).
// ---------------------------------------------------------------------------
namespace {
struct ProfRec { hipEvent_t a, b; double flops; int cls; hipStream_t s; };
std::atomic<bool> g_prof{false};
std::mutex g_prof_mu;
int g_every = 1;
long long g_count[5] = {0, 0, 0, 0, 0};
std::vector<ProfRec> g_recs;
std::vector<hipEvent_t> g_pool;
hipEvent_t take_event() {
    if (!g_pool.empty()) { hipEvent_t e = g_pool.back(); g_pool.pop_back(); return e; }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
}
}  // namespace

void prof_begin(int every) {
    std::lock_guard<std::mutex> lock(g_prof_mu);
    g_every = every > 0 ? every : 1;
    for (auto& x : g_count) x = 0;
    g_recs.clear();
    g_prof = true;
}

__global__ void k_prof_empty() {}

// Event-pair cost of an empty launch on a stream: what each sampled
// duration carries on top of its kernel.
static double event_pair_overhead_ms(hipStream_t s) {
    std::vector<float> d;
    hipEvent_t a = take_event(), b = take_event();
    for (int i = 0; i < 33; ++i) {
        (void)hipEventRecord(a, s);
        hipLaunchKernelGGL(k_prof_empty, dim3(1), dim3(64), 0, s);
        (void)hipEventRecord(b, s);
        (void)hipEventSynchronize(b);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, a, b);
        if (i > 0) d.push_back(ms);
    }
    g_pool.push_back(a);
    g_pool.push_back(b);
    std::sort(d.begin(), d.end());
    return d[d.size() / 2];
}

void prof_end(double* out) {
    g_prof = false;
    std::lock_guard<std::mutex> lock(g_prof_mu);
    for (int i = 0; i < 16; ++i) out[i] = 0.0;
    // the sampled launches' stream (any recorded one)
    const double ovh = g_recs.empty() ? 0.0 : event_pair_overhead_ms(g_recs.front().s);
    out[15] = ovh;
    for (auto& r : g_recs) {
        (void)hipEventSynchronize(r.b);
        float ms_raw = 0.0f;
        (void)hipEventElapsedTime(&ms_raw, r.a, r.b);
        const double ms = std::max(0.0, static_cast<double>(ms_raw) - ovh);
        const int slots[2] = {r.cls == 1 ? 1 : 0, r.cls};   // class 0 sums every GEMM bucket
        for (int u = 0; u < (r.cls >= 2 ? 2 : 1); ++u) {
            out[3 * slots[u] + 0] += ms;
            out[3 * slots[u] + 1] += r.flops;
            out[3 * slots[u] + 2] += 1.0;
        }
        g_pool.push_back(r.a);
        g_pool.push_back(r.b);
    }
    g_recs.clear();
}

struct ProfScope {
    bool on;
    ProfRec r;
    hipStream_t s;
    ProfScope(hipStream_t st, double flops, int cls) : on(g_prof), s(st) {
        if (!on) return;
        std::lock_guard<std::mutex> lock(g_prof_mu);
        on = (g_count[cls]++ % g_every) == 0;
        if (on) { r.a = take_event(); r.b = take_event(); r.flops = flops; r.cls = cls; r.s = s; (void)hipEventRecord(r.a, s); }
    }
    ~ProfScope() {
        if (!on) return;
        std::lock_guard<std::mutex> lock(g_prof_mu);
        (void)hipEventRecord(r.b, s);
        g_recs.push_back(r);
    }
};

namespace {
std::atomic<int> g_tiles{0};
}
void set_gemm_tiles(int policy) {
    if (policy < 0 || policy > 2) throw std::runtime_error("GEMM tile policy must be 0, 1 or 2");
    g_tiles.store(policy);
}
int gemm_tiles() { return g_tiles.load(); }

void gemm(const Gemm& g, hipStream_t s) {
    if (g.M <= 0 || g.N <= 0 || g.batch <= 0) return;
    GemmK p{g.M, g.N, g.K, g.A, g.lda, g.sA, g.tA, g.B, g.ldb, g.sB, g.tB, g.C, g.ldc, g.sC,
            g.alpha, g.beta, g.alpha_b, g.zc0, g.zc1};
    // sampled timing buckets by depth: 2 (K < 64), 3 (64 <= K < 256), 4 (K >= 256)
    ProfScope ps(s, 2.0 * g.M * g.N * static_cast<double>(g.K) * g.batch, g.K < 64 ? 2 : (g.K < 256 ? 3 : 4));
    // Tile by shape (measured on the Gauss-Jordan updates and back-propagation
    // products at batch 32): 128 x 128 over 8 waves (64 x 32 each) for deep
    // products that fill it (45-51 TFLOP/s where 64 x 64 gives 36-42);
    // 64 x 64 over 4 waves for shallow ones, which are bound by latency and
    // memory and want the larger grid; 16- and 32-row tiles for the thin
    // block updates of the innermost Gauss-Jordan levels (M = 16 or 32
    // columns), 20-30% faster there than 64-row tiles that are mostly empty.
    // Policy 0 (default): 128 x 128 for deep products (K >= 256); 64 x 128
    // over 8 waves (32 x 32 each) for 64 < M with K < 256, the lower levels of
    // the Gauss-Jordan recursion, where the 128-row tile leaves too few
    // workgroups (M = 192, K = 192: 229 against 297 us at batch 64).
    // The small tiles of shallow products load C's old values before their
    // K loop (PC): the order-1566 inverse at batch 64 went from 23.2 to 22.3 ms.
    const int pol = g_tiles.load(std::memory_order_relaxed);
    const bool big = g.M > 64 && (pol == 2 || (pol == 0 && g.K >= 256));
    if (!big && pol == 0 && g.M > 64) {
        dim3 grid((g.M + 63) / 64, (g.N + 127) / 128, g.batch);
        hipLaunchKernelGGL((k_dgemm<64, 128, 2, 4, true>), grid, dim3(512), 0, s, p);
        return;
    }
    if (big) {
        // Rows past the last full 128-row tile: when at most 64 remain, they
        // go to a second launch with a thin tile instead of a mostly empty
        // 128-row tile (392 rows: 400 rows of MFMA work instead of 512).
        // Every tile shape sums K in the same order, so C is bitwise the same.
        const int M1 = (g.M / 128) * 128, R = g.M - M1;
        const bool split = M1 > 0 && R > 0 && R <= 64;
        GemmK p1 = p;
        if (split) p1.M = M1;
        dim3 grid((p1.M + 127) / 128, (g.N + 127) / 128, g.batch);
        hipLaunchKernelGGL((k_dgemm<128, 128, 2, 4>), grid, dim3(512), 0, s, p1);
        if (split) {
            GemmK p2 = p;
            p2.M = R;
            p2.A = g.tA ? g.A + static_cast<int64_t>(M1) * g.lda : g.A + M1;
            p2.C = g.C + M1;
            if (R <= 16) {
                hipLaunchKernelGGL((k_dgemm<16, 128, 1, 4>), dim3(1, (g.N + 127) / 128, g.batch), dim3(256), 0, s, p2);
            } else if (R <= 32) {
                hipLaunchKernelGGL((k_dgemm<32, 128, 1, 4>), dim3(1, (g.N + 127) / 128, g.batch), dim3(256), 0, s, p2);
            } else {
                hipLaunchKernelGGL((k_dgemm<64, 64, 2, 2>), dim3(1, (g.N + 63) / 64, g.batch), dim3(256), 0, s, p2);
            }
        }
    } else if (pol != 1 && g.M <= 16) {
        dim3 grid((g.M + 15) / 16, (g.N + 127) / 128, g.batch);
        hipLaunchKernelGGL((k_dgemm<16, 128, 1, 4, true>), grid, dim3(256), 0, s, p);
    } else if (pol != 1 && g.M <= 32) {
        dim3 grid((g.M + 31) / 32, (g.N + 127) / 128, g.batch);
        hipLaunchKernelGGL((k_dgemm<32, 128, 1, 4, true>), grid, dim3(256), 0, s, p);
    } else {
        dim3 grid((g.M + 63) / 64, (g.N + 63) / 64, g.batch);
        if (g.K < 256) hipLaunchKernelGGL((k_dgemm<64, 64, 2, 2, true>), grid, dim3(256), 0, s, p);
        else hipLaunchKernelGGL((k_dgemm<64, 64, 2, 2>), grid, dim3(256), 0, s, p);
    }
}

// ---------------------------------------------------------------------------
// Gauss-Jordan leaf: pivots c .. c+w-1 (w <= 16) on columns [c, c+w) of the
// row-major view, as three kernels that spread each matrix over many
// workgroups.  Pivot rows come from tournament pivoting (the panel step of
// communication-avoiding LU):
//
//  k_leaf_nominate  grid (row blocks of 256, batch).  Optionally applies the
//      left sibling's pivot block to this panel first (pre-update).  Every
//      wave runs partial pivoting on its 64 rows with wave-only traffic (DPP
//      max, readlanes) and nominates at most w rows; wave 0 repeats the
//      selection on the 64 nominees' panel values and keeps w per block.
//      The block that finishes last (per-matrix completion counter) then
//      runs the merge with one wave: partial pivoting on the block nominees
//      fixes the w pivot rows in order; it inverts the pivot block A_P
//      (Gauss-Jordan in that order), turns the choice into the LAPACK-style
//      interchange record piv[c+jj] and the composite move list of the
//      <= 2w rows involved, and stages the old values the apply step must
//      not race on (moved rows' panel and side-range values, and T = chosen
//      rows of the side range).
//  k_leaf_apply     grid (blocks of 256 rows / 256 columns, batch).  Rows:
//      X(P,K) = A_P^-1, X(O,K) = -A_O A_P^-1 (the w pivot steps restricted to
//      the panel), and optionally the post-update of the side range Jq with
//      this pivot block.  Columns outside panel and Jq: the row moves.
//
// With the pivot rows fixed the panel result equals w Gauss-Jordan steps on
// the panel columns; the columns outside the panel only receive the row
// interchanges (their block updates are the recursion's GEMMs or the fused
// sibling updates below).
//
// Fused sibling updates for a node of two leaves [c, c+w1) and
// [c+w1, c+w1+w2):  the left leaf's merge stages T = X(K_L, J_R) (after
// its interchanges) and the right leaf's nominate applies X(O,J_R) +=
// X(O,K_L) T, X(K_L,J_R) = X(K_L,K_L) T; the right leaf's merge stages
// T2 = X(K_R, J_L) and its apply performs X(O,J_L) += X(O,K_R) T2,
// X(K_R,J_L) = X(K_R,K_R) T2.
// ---------------------------------------------------------------------------
constexpr int GJW = 16;
constexpr int LEAF_BLK = 256;          // rows (and columns) per nominate/apply block
constexpr int STG_D = 4 * GJW * GJW + 2 * 2 * GJW * GJW;   // doubles of staging per matrix
constexpr int STG_I = 4 * GJW + 4;     // ints of staging per matrix

// staging layout (doubles): Ainv[16][16] | T[16][16] | mvK[32][16] | mvX[32][16]
template <class P> __device__ __forceinline__ P* stg_ainv(P* s) { return s; }
template <class P> __device__ __forceinline__ P* stg_T(P* s) { return s + GJW * GJW; }
template <class P> __device__ __forceinline__ P* stg_mvK(P* s) { return s + 2 * GJW * GJW; }
template <class P> __device__ __forceinline__ P* stg_mvX(P* s) { return s + 4 * GJW * GJW; }
// staging ints: dst[32] | src[32] | cnt

__device__ __forceinline__ double lane_read(double x, int l) {
    const long long bits = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(bits), l);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(bits >> 32), l);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
    const long long bits = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(bits), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(bits >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// max over the 64 lanes, returned uniform: quad_perm [1,0,3,2], [2,3,0,1],
// row_half_mirror and row_mirror reduce each 16-lane row; four readlanes
// combine the rows.
__device__ __forceinline__ double wave_max(double x) {
    x = fmax(x, dpp_d<0xB1>(x));
    x = fmax(x, dpp_d<0x4E>(x));
    x = fmax(x, dpp_d<0x141>(x));
    x = fmax(x, dpp_d<0x140>(x));
    return fmax(fmax(lane_read(x, 0), lane_read(x, 16)), fmax(lane_read(x, 32), lane_read(x, 48)));
}

// Partial-pivoting selection of up to w rows among the wave's NC candidate
// slots per lane (row index crow, panel values cv; crow < 0 = empty).  cv is
// destroyed.  Lane jj < w receives the row chosen for column jj, or -1 when
// no candidate has a non-zero entry left in that column.
template <int NC>
__device__ __forceinline__ int wave_select(double (&cv)[NC][GJW], const int (&crow)[NC], int w, int lane) {
    unsigned int csel = 0;
    int chosen = -1;
#pragma unroll
    for (int jj = 0; jj < GJW; ++jj) {
        if (jj < w) {
            double mine = -1.0;
            int eb = 0;
#pragma unroll
            for (int e = 0; e < NC; ++e) {
                const double a = fabs(cv[e][jj]);
                if (crow[e] >= 0 && !((csel >> e) & 1u) && a > mine) { mine = a; eb = e; }
            }
            const double m = wave_max(mine);
            const unsigned long long bal = __ballot(mine == m && mine > 0.0);
            if (bal != 0ull) {
                const int L = __ffsll(static_cast<long long>(bal)) - 1;
                const int ep = __builtin_amdgcn_readlane(eb, L);
                // the winning slot ep is uniform: every lane selects that
                // slot (a chain over the static slot indices, so the register
                // array is never indexed dynamically), then lane L's value is
                // read out once
                int cr = crow[0];
#pragma unroll
                for (int e = 1; e < NC; ++e) cr = (ep == e) ? crow[e] : cr;
                const int rowp = __builtin_amdgcn_readlane(cr, L);
                if (lane == L) csel |= 1u << ep;
                if (lane == jj) chosen = rowp;
                double prow[GJW];
#pragma unroll
                for (int j = 0; j < GJW; ++j) {
                    if (j >= jj) {
                        double x = cv[0][j];
#pragma unroll
                        for (int e = 1; e < NC; ++e) x = (ep == e) ? cv[e][j] : x;
                        prow[j] = lane_read(x, L);
                    } else {
                        prow[j] = 0.0;
                    }
                }
                const double pinv = 1.0 / prow[jj];
#pragma unroll
                for (int e = 0; e < NC; ++e) {
                    if (crow[e] >= 0 && !((csel >> e) & 1u)) {
                        const double f = cv[e][jj] * pinv;
#pragma unroll
                        for (int j = 0; j < GJW; ++j)
                            if (j > jj) cv[e][j] = fma(-f, prow[j], cv[e][j]);
                    }
                }
            }
        }
    }
    return chosen;
}

// Merge step, run by one wave: partial pivoting on the block nominees fixes
// the w pivot rows in order; inverts the pivot block A_P (Gauss-Jordan in
// that order); writes the LAPACK-style interchange record and the composite
// move list; stages T = chosen rows x side range and the moved rows' old
// panel / side-range values (all read before anything moves).
// NC = nominee slots per lane: nblk * 16 <= 64 * NC.
template <int NC>
__device__ __forceinline__ void leaf_merge_wave(const double* __restrict__ M, int64_t ld, int n, int c, int w,
                                                int xa0, int xw, const int* nom_b, int nblk, double* S, int* SI,
                                                int* piv_b, int* info_b, int lane) {
    double cv[NC][GJW];
    int crow[NC];
#pragma unroll
    for (int e = 0; e < NC; ++e) {
        const int idx = lane + 64 * e;
        crow[e] = (idx < nblk * GJW) ? nom_b[idx] : -1;
#pragma unroll
        for (int j = 0; j < GJW; ++j)
            cv[e][j] = (crow[e] >= 0 && j < w) ? M[(int64_t)crow[e] * ld + c + j] : 0.0;
    }
    int chosen = wave_select<NC>(cv, crow, w, lane);
    // rank-deficient panel: fill empty steps with rows of [c, c+w) not yet
    // chosen (one exists for every empty step) and flag the matrix
    int sing = 0;
    for (int jj = 0; jj < w; ++jj) {
        if (__builtin_amdgcn_readlane(chosen, jj) >= 0) continue;
        if (sing == 0) sing = c + jj + 1;
        bool fr = lane < w;
        for (int j2 = 0; j2 < w; ++j2)
            if (__builtin_amdgcn_readlane(chosen, j2) == c + lane) fr = false;
        const unsigned long long fb = __ballot(fr);
        const int rowp = c + __ffsll(static_cast<long long>(fb)) - 1;
        if (lane == jj) chosen = rowp;
    }
    // interchange record: slot s < w is position c+s, slot w+i is position
    // chosen_i when it lies below the panel rows; content = original row now
    // at the slot's position.
    const int prs = __shfl(chosen, (lane >= w && lane < 2 * w) ? lane - w : 0);
    int pos = -1;
    if (lane < w) pos = c + lane;
    else if (lane < 2 * w) pos = (prs >= c + w) ? prs : -1;
    int content = pos;
    int mypiv = 0;
    for (int jj = 0; jj < w; ++jj) {
        const int pr = __builtin_amdgcn_readlane(chosen, jj);
        const unsigned long long hb = __ballot(pos >= 0 && content == pr);
        const int bs = __ffsll(static_cast<long long>(hb)) - 1;
        const int pp = __builtin_amdgcn_readlane(pos, bs);
        const int va = __builtin_amdgcn_readlane(content, jj);
        const int vb = __builtin_amdgcn_readlane(content, bs);
        if (lane == jj) { content = vb; mypiv = pp; }
        else if (lane == bs) content = va;
    }
    const bool mv = pos >= 0 && content != pos;
    // every value this step reads after selection, issued together: A_P
    // rows, T = chosen rows x side range, moved rows' old values
    double a[GJW], tv[GJW], mk[GJW], mx[GJW];
    {
        const int pr = (lane < w) ? chosen : -1;
        const int src = mv ? content : -1;
#pragma unroll
        for (int j = 0; j < GJW; ++j) {
            a[j] = (pr >= 0 && j < w) ? M[(int64_t)pr * ld + c + j] : ((lane == j) ? 1.0 : 0.0);
            tv[j] = (pr >= 0 && j < xw) ? M[(int64_t)pr * ld + xa0 + j] : 0.0;
            mk[j] = (src >= 0 && j < w) ? M[(int64_t)src * ld + c + j] : 0.0;
            mx[j] = (src >= 0 && j < xw) ? M[(int64_t)src * ld + xa0 + j] : 0.0;
        }
    }
    // inverse of the pivot block, identity padded to 16 x 16
#pragma unroll
    for (int k = 0; k < GJW; ++k) {
        double pk[GJW];
#pragma unroll
        for (int j = 0; j < GJW; ++j) pk[j] = lane_read(a[j], k);
        const double pinv = 1.0 / pk[k];
        if (lane == k) {
#pragma unroll
            for (int j = 0; j < GJW; ++j) a[j] = (j == k) ? pinv : pk[j] * pinv;
        } else {
            const double f = a[k] * pinv;
#pragma unroll
            for (int j = 0; j < GJW; ++j) a[j] = (j == k) ? -f : fma(-f, pk[j], a[j]);
        }
    }
    if (lane < GJW) {
#pragma unroll
        for (int j = 0; j < GJW; ++j) {
            stg_ainv(S)[lane * GJW + j] = a[j];
            if (xw > 0) stg_T(S)[lane * GJW + j] = tv[j];
        }
    }
    const unsigned long long bal = __ballot(mv);
    if (mv) {
        const int at = __popcll(bal & ((1ull << lane) - 1ull));
        SI[at] = pos;                 // dst
        SI[2 * GJW + at] = content;   // src
#pragma unroll
        for (int j = 0; j < GJW; ++j) {
            stg_mvK(S)[at * GJW + j] = mk[j];
            stg_mvX(S)[at * GJW + j] = mx[j];
        }
    }
    if (lane == 0) {
        SI[4 * GJW] = __popcll(bal);
        if (sing != 0 && *info_b == 0) *info_b = sing;
    }
    if (lane < w) piv_b[c + lane] = mypiv;
}

// Nominate (all blocks) + merge (the block that finishes last, found with a
// per-matrix completion counter; no block waits for another).  Each block
// publishes its nominees and pre-updated panel rows with a device-scope
// release fence before counting in; only the last block, which reads them,
// takes the acquire fence (it invalidates the XCD's L2, which the other
// blocks have no reason to pay for), and it resets the counter for the next
// leaf.
template <int NC>
__global__ __launch_bounds__(LEAF_BLK) void k_leaf_nominate(double* mat, int64_t ld, int64_t stride, int n,
                                                            int c, int w, int pre_c, int pre_w, int xa0,
                                                            int xw, double* stg, int* stgi, int* nom, int nblk,
                                                            int* cnt, int* piv, int* info) {
    const int64_t b = blockIdx.y;
    double* __restrict__ M = mat + b * stride;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int r = blockIdx.x * LEAF_BLK + t;
    __shared__ double s_T[GJW][GJW];
    __shared__ double s_rows[LEAF_BLK][GJW + 1];   // the block's panel rows (+1: bank spread)
    __shared__ int s_nom[4][GJW];
    __shared__ int s_last;
    // The block's 256 x 16 panel tile (and, for the pre-update, its tile in
    // the left sibling's columns) moves between HBM and registers with 16
    // lanes per row segment, and is turned into one row per thread through
    // LDS: a per-thread row access would touch 64 lines per instruction.
    const int r0 = blockIdx.x * LEAF_BLK;
    const int tj = t % GJW, tr = t / GJW;
    constexpr int TRS = LEAF_BLK / GJW;   // tile rows per pass
    double pv[GJW], px[GJW];
#pragma unroll
    for (int it = 0; it < GJW; ++it) {
        const int rr = r0 + tr + it * TRS;
        pv[it] = (rr < n && tj < w) ? M[(int64_t)rr * ld + c + tj] : 0.0;
        px[it] = (rr < n && tj < pre_w) ? M[(int64_t)rr * ld + pre_c + tj] : 0.0;
    }
    if (pre_w > 0) s_T[t / GJW][t % GJW] = stg_T(stg + b * STG_D)[t];
#pragma unroll
    for (int it = 0; it < GJW; ++it) s_rows[tr + it * TRS][tj] = pv[it];
    __syncthreads();
    double v[GJW];
#pragma unroll
    for (int j = 0; j < GJW; ++j) v[j] = s_rows[t][j];
    if (pre_w > 0) {
        double xv[GJW];
        __syncthreads();
#pragma unroll
        for (int it = 0; it < GJW; ++it) s_rows[tr + it * TRS][tj] = px[it];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < GJW; ++j) xv[j] = s_rows[t][j];
        if (r < n) {
            const bool inK = (r >= pre_c && r < pre_c + pre_w);
            double acc[GJW];
#pragma unroll
            for (int j = 0; j < GJW; ++j) acc[j] = inK ? 0.0 : v[j];
#pragma unroll
            for (int kk = 0; kk < GJW; ++kk) {
                if (kk < pre_w) {
#pragma unroll
                    for (int j = 0; j < GJW; ++j) acc[j] = fma(xv[kk], s_T[kk][j], acc[j]);
                }
            }
#pragma unroll
            for (int j = 0; j < GJW; ++j) v[j] = (j < w) ? acc[j] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < GJW; ++j) s_rows[t][j] = v[j];
        __syncthreads();
#pragma unroll
        for (int it = 0; it < GJW; ++it) {
            const int rr = r0 + tr + it * TRS;
            if (rr < n && tj < w) M[(int64_t)rr * ld + c + tj] = s_rows[tr + it * TRS][tj];
        }
    }
    // wave-local partial pivoting over the wave's rows (r >= c eligible)
    {
        double cv[1][GJW];
#pragma unroll
        for (int j = 0; j < GJW; ++j) cv[0][j] = v[j];
        const int crow[1] = {(r >= c && r < n) ? r : -1};
        const int ch = wave_select<1>(cv, crow, w, lane);
        if (lane < GJW) s_nom[wv][lane] = (lane < w) ? ch : -1;
    }
    __syncthreads();
    // block merge: wave 0, one nominee per lane (64 = 4 waves x 16), values from LDS
    if (wv == 0) {
        double cv[1][GJW];
        int crow[1];
        crow[0] = s_nom[lane >> 4][lane & 15];
        const int loc = (crow[0] >= 0) ? crow[0] - blockIdx.x * LEAF_BLK : 0;
#pragma unroll
        for (int j = 0; j < GJW; ++j) cv[0][j] = (crow[0] >= 0) ? s_rows[loc][j] : 0.0;
        const int ch = wave_select<1>(cv, crow, w, lane);
        if (lane < GJW) nom[(b * nblk + blockIdx.x) * GJW + lane] = (lane < w) ? ch : -1;
    }
    // count in; the last block of this matrix runs the merge
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (t == 0) s_last = (atomicAdd(cnt + b, 1) == nblk - 1) ? 1 : 0;
    __syncthreads();
    if (!s_last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (wv == 0)
        leaf_merge_wave<NC>(M, ld, n, c, w, xa0, xw, nom + b * nblk * GJW, nblk, stg + b * STG_D,
                            stgi + b * STG_I, piv + b * n, info + b, lane);
    if (t == 0) cnt[b] = 0;
}

// post: 1 = update side range Jq = [xa0, xa0+xw) with this pivot block
// (T2 staged by the merge); 0 = the side range is only moved.
__global__ __launch_bounds__(LEAF_BLK) void k_leaf_apply(double* mat, int64_t ld, int64_t stride, int n,
                                                         int c, int w, int xa0, int xw, int post,
                                                         const double* stg, const int* stgi) {
    const int64_t b = blockIdx.y;
    double* __restrict__ M = mat + b * stride;
    const int t = threadIdx.x;
    const double* S = stg + b * STG_D;
    const int* SI = stgi + b * STG_I;
    __shared__ double s_inv[GJW][GJW];
    __shared__ double s_T2[GJW][GJW];
    __shared__ double s_mv[2][2 * GJW][GJW];   // moved rows' old panel / side-range values
    __shared__ double s_tile[LEAF_BLK][GJW + 1];
    // Staging first: the loads below are waited for in issue order, so the
    // small staging reads must not queue behind the tile and row-move loads.
    constexpr int NMV = 4 * GJW * GJW / LEAF_BLK;
    const double inv_v = stg_ainv(S)[t];
    const double t2_v = post ? stg_T(S)[t] : 0.0;
    double mvv[NMV];
#pragma unroll
    for (int q = 0; q < NMV; ++q) {
        const int idx = t + q * LEAF_BLK, h = idx / (2 * GJW * GJW), e = idx % (2 * GJW * GJW);
        mvv[q] = (h == 0 ? stg_mvK(S) : stg_mvX(S))[e];
    }
    // The move list is uniform.  Slots at or beyond cnt repeat move 0 (the
    // same value to the same address), which keeps the row-move loads and
    // stores below free of per-slot branches.
    const int cnt = SI[4 * GJW];
    int mdst[2 * GJW], msrc[2 * GJW];
#pragma unroll
    for (int e = 0; e < 2 * GJW; ++e) { mdst[e] = SI[e]; msrc[e] = SI[2 * GJW + e]; }
#pragma unroll
    for (int e = 1; e < 2 * GJW; ++e)
        if (e >= cnt) { mdst[e] = mdst[0]; msrc[e] = msrc[0]; }
    // The block's panel tile and side-range tile move between HBM and
    // registers with 16 lanes per row segment and are turned into one row
    // per thread (and back) through LDS, as in k_leaf_nominate.
    const int r0 = blockIdx.x * LEAF_BLK, r = r0 + t;
    const int tj = t % GJW, tr = t / GJW;
    constexpr int TRS = LEAF_BLK / GJW;
    double pa[GJW], px[GJW];
#pragma unroll
    for (int it = 0; it < GJW; ++it) {
        const int rr = r0 + tr + it * TRS;
        pa[it] = (rr < n && tj < w) ? M[(int64_t)rr * ld + c + tj] : 0.0;
        px[it] = (post && rr < n && tj < xw) ? M[(int64_t)rr * ld + xa0 + tj] : 0.0;
    }
    // Columns outside the panel (and outside Jq when post) only receive the
    // row moves: their loads go out with the tile loads and their stores
    // follow the first barrier (they touch columns disjoint from the tiles
    // this kernel reads and writes).
    const int jc = blockIdx.x * LEAF_BLK + t;
    const bool colmv = cnt > 0 && jc < n && !(jc >= c && jc < c + w) && !(post && jc >= xa0 && jc < xa0 + xw);
    double xm[2 * GJW];
    if (colmv) {
#pragma unroll
        for (int e = 0; e < 2 * GJW; ++e) xm[e] = M[(int64_t)msrc[e] * ld + jc];
    }
    s_inv[t / GJW][t % GJW] = inv_v;
    if (post) s_T2[t / GJW][t % GJW] = t2_v;
#pragma unroll
    for (int q = 0; q < NMV; ++q) {
        const int idx = t + q * LEAF_BLK, h = idx / (2 * GJW * GJW), e = idx % (2 * GJW * GJW);
        s_mv[h][e / GJW][e % GJW] = mvv[q];
    }
#pragma unroll
    for (int it = 0; it < GJW; ++it) s_tile[tr + it * TRS][tj] = pa[it];
    __syncthreads();
    if (colmv) {
#pragma unroll
        for (int e = 0; e < 2 * GJW; ++e) M[(int64_t)mdst[e] * ld + jc] = xm[e];
    }
    // own row's panel and side-range values (replaced from the staging area
    // below when the row receives a move)
    double a0[GJW], x0[GJW];
#pragma unroll
    for (int j = 0; j < GJW; ++j) a0[j] = s_tile[t][j];
    if (post) {
        __syncthreads();
#pragma unroll
        for (int it = 0; it < GJW; ++it) s_tile[tr + it * TRS][tj] = px[it];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < GJW; ++j) x0[j] = s_tile[t][j];
    }
    // ---- rows: o = new panel values, x = new side-range values
    double o[GJW], x[GJW];
#pragma unroll
    for (int j = 0; j < GJW; ++j) { o[j] = 0.0; x[j] = 0.0; }
    if (r < n) {
        int e_src = -1;
#pragma unroll
        for (int e = 0; e < 2 * GJW; ++e)
            if (e < cnt && mdst[e] == r) e_src = e;
        const bool inK = (r >= c && r < c + w);
        if (inK) {
#pragma unroll
            for (int j = 0; j < GJW; ++j) o[j] = s_inv[r - c][j];
        } else {
            double a[GJW];
            if (e_src >= 0) {
#pragma unroll
                for (int j = 0; j < GJW; ++j) a[j] = s_mv[0][e_src][j];
            } else {
#pragma unroll
                for (int j = 0; j < GJW; ++j) a[j] = a0[j];
            }
#pragma unroll
            for (int kk = 0; kk < GJW; ++kk) {
#pragma unroll
                for (int j = 0; j < GJW; ++j) o[j] = fma(-a[kk], s_inv[kk][j], o[j]);
            }
        }
        if (post) {
            if (!inK) {
                if (e_src >= 0) {
#pragma unroll
                    for (int j = 0; j < GJW; ++j) x[j] = s_mv[1][e_src][j];
                } else {
#pragma unroll
                    for (int j = 0; j < GJW; ++j) x[j] = x0[j];
                }
            }
#pragma unroll
            for (int kk = 0; kk < GJW; ++kk) {
                if (kk < w) {
#pragma unroll
                    for (int j = 0; j < GJW; ++j) x[j] = fma(o[kk], s_T2[kk][j], x[j]);
                }
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < GJW; ++j) s_tile[t][j] = o[j];
    __syncthreads();
#pragma unroll
    for (int it = 0; it < GJW; ++it) {
        const int rr = r0 + tr + it * TRS;
        if (rr < n && tj < w) M[(int64_t)rr * ld + c + tj] = s_tile[tr + it * TRS][tj];
    }
    if (post) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < GJW; ++j) s_tile[t][j] = x[j];
        __syncthreads();
#pragma unroll
        for (int it = 0; it < GJW; ++it) {
            const int rr = r0 + tr + it * TRS;
            if (rr < n && tj < xw) M[(int64_t)rr * ld + xa0 + tj] = s_tile[tr + it * TRS][tj];
        }
    }
}

// Copy the block rows K (nk rows from k0) x columns J (nj from j0) of the
// row-major view into T (row-major nk x nj): one element per thread over the
// flattened nk * nj block, so T is written contiguously and a row of M is
// read contiguously whatever nj is.
__global__ void k_copy_block(const double* mat, int64_t ld, int64_t stride, int k0, int nk, int j0,
                             int nj, double* tb, int64_t tstride) {
    const int64_t b = blockIdx.y;
    const double* M = mat + b * stride;
    double* T = tb + b * tstride;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nk * nj) return;
    const int kk = static_cast<int>(e / nj), jj = static_cast<int>(e % nj);
    T[e] = M[(int64_t)(k0 + kk) * ld + j0 + jj];
}

// Column-interchange composition: perm = S_{n-1} ... S_0 applied to identity.
// With `staged` the interchange record is copied to LDS first, so the serial
// chain of swaps waits on LDS rather than on a global load per step (orders
// up to 8192, 64 KB of LDS).
__global__ void k_gj_perm(const int* piv, int* perm, int n, int staged) {
    extern __shared__ int sp[];   // [0, n): permutation, [n, 2n): interchange record when staged
    const int64_t b = blockIdx.x;
    const int* pm_in = piv + b * n;
    int* pm = perm + b * n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        sp[i] = i;
        if (staged) sp[n + i] = pm_in[i];
    }
    const int* sv = staged ? sp + n : pm_in;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = n - 1; k >= 0; --k) {
            const int q = sv[k];
            const int x = sp[k]; sp[k] = sp[q]; sp[q] = x;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) pm[i] = sp[i];
}

// out_rm(i, j) = X_rm(i, perm[j])  [ or 1/2 (X_rm(i, perm[j]) + X_rm(j, perm[i])) ]
constexpr int PT = 32;
__global__ __launch_bounds__(256) void k_gj_final(const double* mat, double* out, int64_t ld,
                                                  int64_t stride, int n, const int* perm, int sym) {
    const int64_t b = blockIdx.z;
    const double* X = mat + b * stride;
    double* O = out + b * stride;
    const int* pm = perm + b * n;
    __shared__ double t1[PT][PT + 1];
    __shared__ double t2[PT][PT + 1];
    const int i0 = blockIdx.y * PT, j0 = blockIdx.x * PT;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    for (int yy = ty; yy < PT; yy += 8) {
        const int i = i0 + yy, j = j0 + tx;
        if (i < n && j < n) t1[yy][tx] = X[(int64_t)i * ld + pm[j]];
        if (sym) {
            const int jr = j0 + yy, ic = i0 + tx;   // row jr, column perm[ic]
            if (jr < n && ic < n) t2[yy][tx] = X[(int64_t)jr * ld + pm[ic]];
        }
    }
    __syncthreads();
    for (int yy = ty; yy < PT; yy += 8) {
        const int i = i0 + yy, j = j0 + tx;
        if (i < n && j < n) {
            double v = t1[yy][tx];
            if (sym) v = 0.5 * (v + t2[tx][yy]);
            O[(int64_t)i * ld + j] = v;
        }
    }
}

// Symmetrised form over tile pairs: the workgroup of upper-triangle tile pair
// (I, J), I <= J, gathers G1 = X_rm(I rows, perm[J cols]) and G2 = X_rm(J
// rows, perm[I cols]) once and writes both out(I, J) = 1/2 (G1 + G2^T) and
// out(J, I) = 1/2 (G2 + G1^T): half the gathered reads of one tile per
// workgroup, and the same sums in the same order.
__global__ __launch_bounds__(256) void k_gj_final_sym(const double* mat, double* out, int64_t ld,
                                                      int64_t stride, int n, const int* perm) {
    const int64_t b = blockIdx.z;
    const double* X = mat + b * stride;
    double* O = out + b * stride;
    const int* pm = perm + b * n;
    __shared__ double g1[PT][PT + 1];
    __shared__ double g2[PT][PT + 1];
    const int t = blockIdx.x;   // t = J (J + 1) / 2 + I
    int J = static_cast<int>((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((J + 1) * (J + 2) / 2 <= t) ++J;
    while (J * (J + 1) / 2 > t) --J;
    const int I = t - J * (J + 1) / 2;
    const int i0 = I * PT, j0 = J * PT;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    for (int yy = ty; yy < PT; yy += 8) {
        const int i = i0 + yy, j = j0 + tx;
        if (i < n && j < n) g1[yy][tx] = X[(int64_t)i * ld + pm[j]];
        const int jr = j0 + yy, ic = i0 + tx;
        if (jr < n && ic < n) g2[yy][tx] = X[(int64_t)jr * ld + pm[ic]];
    }
    __syncthreads();
    for (int yy = ty; yy < PT; yy += 8) {
        const int i = i0 + yy, j = j0 + tx;
        if (i < n && j < n) O[(int64_t)i * ld + j] = 0.5 * (g1[yy][tx] + g2[tx][yy]);
        const int jr = j0 + yy, ic = i0 + tx;
        if (I != J && jr < n && ic < n) O[(int64_t)jr * ld + ic] = 0.5 * (g2[yy][tx] + g1[tx][yy]);
    }
}

// ---------------------------------------------------------------------------
// Gauss-Jordan leaf in one workgroup per matrix, with partial pivoting
// (orders up to WG_THREADS * R).  Thread t owns rows t + q * WG_THREADS, q < R,
// and keeps their panel values in registers, so the pivot search of each
// panel column runs over the whole matrix with one barrier per column:
//
//  phase 0  load the panel rows (and, with pre_w > 0, apply the left
//           sibling's pivot block: P = [r in K_L ? 0 : P] + X(:, K_L) T);
//           the pre-updated panel is written back for the staging reads.
//  phase 1  for each panel column jj: every wave finds its largest |value|
//           among its rows r >= c not yet chosen (lowest row on ties) and
//           publishes it with that row's values; after the barrier every
//           thread picks the overall largest (lowest row on ties), which is
//           the row LU with partial pivoting chooses, and eliminates column
//           jj from its candidate rows (values used for the selection only).
//  phase 2  wave 0 turns the choice into the LAPACK interchange record and
//           the move list, inverts the pivot block A_P and stages T (chosen
//           rows x side range, for the right sibling) and the moved rows'
//           old values in LDS, as the tournament merge does.
//  phase 3  rows: X(P,K) = A_P^-1, X(O,K) = -A_O A_P^-1, and with post the
//           side range update, exactly as k_leaf_apply.
//  phase 4  the other columns receive the row moves (one column per thread).
//
// The arithmetic of the pre-update, the pivot block inverse and the apply is
// that of the tournament kernels, so where both choose the same pivot rows
// the results are bitwise the same.  Selected by SEHF_GJ_LEAF=partial: with
// one compute unit per matrix and a barrier per pivot column it is slower
// than the tournament leaf (about 120 against 60 us per leaf at order 1566),
// and serves as the reference pivot order in the parity tests.
// ---------------------------------------------------------------------------
constexpr int WG_THREADS = 1024;
constexpr int WG_WAVES = WG_THREADS / 64;

__device__ __forceinline__ int wave_min_int(int x) {
    x = min(x, __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false));
    x = min(x, __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false));
    x = min(x, __builtin_amdgcn_update_dpp(0, x, 0x141, 0xF, 0xF, false));
    x = min(x, __builtin_amdgcn_update_dpp(0, x, 0x140, 0xF, 0xF, false));
    return min(min(__builtin_amdgcn_readlane(x, 0), __builtin_amdgcn_readlane(x, 16)),
               min(__builtin_amdgcn_readlane(x, 32), __builtin_amdgcn_readlane(x, 48)));
}

template <int R>
__global__ __launch_bounds__(WG_THREADS) void k_leaf_wg(double* mat, int64_t ld, int64_t stride, int n, int c,
                                                        int w, int pre_c, int pre_w, int xa0, int xw, int post,
                                                        double* stg, int* piv, int* info) {
    const int64_t b = blockIdx.x;
    double* __restrict__ M = mat + b * stride;
    double* S = stg + b * STG_D;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    // Row panels move between HBM and registers 16 lanes per row segment
    // (coalesced) and are turned into one row per thread through this tile.
    __shared__ double s_tile[WG_THREADS][GJW + 1];
    __shared__ double s_T[GJW][GJW];                 // pre-update block, then T / T2
    __shared__ double s_cv[2][WG_WAVES];             // per-wave best |value| (step parity)
    __shared__ int s_cr[2][WG_WAVES];                // per-wave best row
    __shared__ double s_crow[2][WG_WAVES][GJW];      // per-wave best row's values
    __shared__ int s_chosen[GJW];
    __shared__ double s_inv[GJW][GJW];
    __shared__ double s_mv[2][2 * GJW][GJW];         // moved rows' old panel / side-range values
    __shared__ int s_dst[2 * GJW], s_src[2 * GJW], s_cnt;
    const int tj = t % GJW, tr = t / GJW;
    constexpr int TRS = WG_THREADS / GJW;            // tile rows per pass
    // coalesced load / store of rows [r0, r0 + WG_THREADS) x columns [col0, col0 + nc)
    auto tile_load = [&](int r0, int col0, int nc, double (&g)[GJW]) {
#pragma unroll
        for (int it = 0; it < GJW; ++it) {
            const int rr = r0 + tr + it * TRS;
            g[it] = (rr < n && tj < nc) ? M[(int64_t)rr * ld + col0 + tj] : 0.0;
        }
    };
    auto tile_store = [&](int r0, int col0, int nc, const double (&g)[GJW]) {
#pragma unroll
        for (int it = 0; it < GJW; ++it) {
            const int rr = r0 + tr + it * TRS;
            if (rr < n && tj < nc) M[(int64_t)rr * ld + col0 + tj] = g[it];
        }
    };
    // registers (coalesced layout) -> own row, and back; the caller's barriers
    // separate consecutive uses of the tile
    auto to_row = [&](const double (&g)[GJW], double (&row)[GJW]) {
#pragma unroll
        for (int it = 0; it < GJW; ++it) s_tile[tr + it * TRS][tj] = g[it];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < GJW; ++j) row[j] = s_tile[t][j];
        __syncthreads();
    };
    auto from_row = [&](const double (&row)[GJW], double (&g)[GJW]) {
#pragma unroll
        for (int j = 0; j < GJW; ++j) s_tile[t][j] = row[j];
        __syncthreads();
#pragma unroll
        for (int it = 0; it < GJW; ++it) g[it] = s_tile[tr + it * TRS][tj];
        __syncthreads();
    };

    // ---- phase 0: panel rows (+ pre-update, written back) ----
    double v[R][GJW];
    if (pre_w > 0 && t < GJW * GJW) s_T[t / GJW][t % GJW] = stg_T(S)[t];
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int r0 = q * WG_THREADS, r = r0 + t;
        double g[GJW];
        tile_load(r0, c, w, g);
        to_row(g, v[q]);
        if (pre_w > 0) {
            tile_load(r0, pre_c, pre_w, g);
            double xv[GJW];
            to_row(g, xv);
            const bool inK = (r >= pre_c && r < pre_c + pre_w);
            double acc[GJW];
#pragma unroll
            for (int j = 0; j < GJW; ++j) acc[j] = inK ? 0.0 : v[q][j];
#pragma unroll
            for (int kk = 0; kk < GJW; ++kk) {
                if (kk < pre_w) {
#pragma unroll
                    for (int j = 0; j < GJW; ++j) acc[j] = fma(xv[kk], s_T[kk][j], acc[j]);
                }
            }
#pragma unroll
            for (int j = 0; j < GJW; ++j) v[q][j] = (r < n && j < w) ? acc[j] : 0.0;
            from_row(v[q], g);
            tile_store(r0, c, w, g);
        }
    }
    // ---- phase 1: partial pivoting over all rows r >= c ----
    unsigned int csel = 0;
    // unrolled over the 16 columns so that every register index is static
#pragma unroll
    for (int jj = 0; jj < GJW; ++jj) {
        if (jj < w) {
            const int par = jj & 1;
            double mine = -1.0;
            int qb = 0;
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const int r = t + q * WG_THREADS;
                const double a = fabs(v[q][jj]);
                if (r >= c && r < n && !((csel >> q) & 1u) && a > mine) { mine = a; qb = q; }
            }
            const double wm = wave_max(mine);
            const int my_row = (wm > 0.0 && mine == wm) ? t + qb * WG_THREADS : INT_MAX;
            const int wr = wave_min_int(my_row);
            if (wr == INT_MAX) {
                if (lane == 0) { s_cv[par][wv] = -1.0; s_cr[par][wv] = INT_MAX; }
            } else if (my_row == wr) {
                s_cv[par][wv] = wm;
                s_cr[par][wv] = wr;
#pragma unroll
                for (int j = 0; j < GJW; ++j) {
                    double x = v[0][j];
#pragma unroll
                    for (int q = 1; q < R; ++q) x = (qb == q) ? v[q][j] : x;
                    s_crow[par][wv][j] = x;
                }
            }
            __syncthreads();
            double best = -1.0;
            int brow = INT_MAX, bw = 0;
#pragma unroll
            for (int x = 0; x < WG_WAVES; ++x) {
                const double cvx = s_cv[par][x];
                const int crx = s_cr[par][x];
                if (cvx > best || (cvx == best && crx < brow)) { best = cvx; brow = crx; bw = x; }
            }
            if (t == 0) s_chosen[jj] = (best > 0.0) ? brow : -1;
            if (best > 0.0) {
                double prow[GJW];
#pragma unroll
                for (int j = 0; j < GJW; ++j) prow[j] = (j >= jj) ? s_crow[par][bw][j] : 0.0;
                const double pinv = 1.0 / prow[jj];
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    const int r = t + q * WG_THREADS;
                    if (r == brow) { csel |= 1u << q; continue; }
                    if (r >= c && r < n && !((csel >> q) & 1u)) {
                        const double f = v[q][jj] * pinv;
#pragma unroll
                        for (int j = 0; j < GJW; ++j)
                            if (j > jj) v[q][j] = fma(-f, prow[j], v[q][j]);
                    }
                }
            }
        }
    }
    __syncthreads();
    // ---- phase 2: interchange record, pivot block inverse, staging (wave 0) ----
    if (wv == 0) {
        int chosen = (lane < w) ? s_chosen[lane] : -1;
        // rank-deficient panel: fill empty steps with rows of [c, c+w) not yet
        // chosen (one exists for every empty step) and flag the matrix
        int sing = 0;
        for (int jj = 0; jj < w; ++jj) {
            if (__builtin_amdgcn_readlane(chosen, jj) >= 0) continue;
            if (sing == 0) sing = c + jj + 1;
            bool fr = lane < w;
            for (int j2 = 0; j2 < w; ++j2)
                if (__builtin_amdgcn_readlane(chosen, j2) == c + lane) fr = false;
            const unsigned long long fb = __ballot(fr);
            const int rowp = c + __ffsll(static_cast<long long>(fb)) - 1;
            if (lane == jj) chosen = rowp;
        }
        // interchange record: slot s < w is position c+s, slot w+i is position
        // chosen_i when it lies below the panel rows; content = original row
        // now at the slot's position
        const int prs = __shfl(chosen, (lane >= w && lane < 2 * w) ? lane - w : 0);
        int pos = -1;
        if (lane < w) pos = c + lane;
        else if (lane < 2 * w) pos = (prs >= c + w) ? prs : -1;
        int content = pos;
        int mypiv = 0;
        for (int jj = 0; jj < w; ++jj) {
            const int pr = __builtin_amdgcn_readlane(chosen, jj);
            const unsigned long long hb = __ballot(pos >= 0 && content == pr);
            const int bs = __ffsll(static_cast<long long>(hb)) - 1;
            const int pp = __builtin_amdgcn_readlane(pos, bs);
            const int va = __builtin_amdgcn_readlane(content, jj);
            const int vb = __builtin_amdgcn_readlane(content, bs);
            if (lane == jj) { content = vb; mypiv = pp; }
            else if (lane == bs) content = va;
        }
        const bool mv = pos >= 0 && content != pos;
        const unsigned long long bal = __ballot(mv);
        // T (chosen rows x side range) and the moved rows' old values go
        // straight to LDS (and T to the sibling's staging)
        {
            const int pr = (lane < w) ? chosen : -1;
            if (lane < GJW && xw > 0) {
#pragma unroll
                for (int j = 0; j < GJW; ++j) {
                    const double tv = (pr >= 0 && j < xw) ? M[(int64_t)pr * ld + xa0 + j] : 0.0;
                    s_T[lane][j] = tv;
                    stg_T(S)[lane * GJW + j] = tv;
                }
            }
            if (mv) {
                const int at = __popcll(bal & ((1ull << lane) - 1ull));
                s_dst[at] = pos;
                s_src[at] = content;
#pragma unroll
                for (int j = 0; j < GJW; ++j) {
                    s_mv[0][at][j] = (j < w) ? M[(int64_t)content * ld + c + j] : 0.0;
                    s_mv[1][at][j] = (j < xw) ? M[(int64_t)content * ld + xa0 + j] : 0.0;
                }
            }
        }
        double a[GJW];
        {
            const int pr = (lane < w) ? chosen : -1;
#pragma unroll
            for (int j = 0; j < GJW; ++j)
                a[j] = (pr >= 0 && j < w) ? M[(int64_t)pr * ld + c + j] : ((lane == j) ? 1.0 : 0.0);
        }
#pragma unroll
        for (int k = 0; k < GJW; ++k) {
            double pk[GJW];
#pragma unroll
            for (int j = 0; j < GJW; ++j) pk[j] = lane_read(a[j], k);
            const double pinv = 1.0 / pk[k];
            if (lane == k) {
#pragma unroll
                for (int j = 0; j < GJW; ++j) a[j] = (j == k) ? pinv : pk[j] * pinv;
            } else {
                const double f = a[k] * pinv;
#pragma unroll
                for (int j = 0; j < GJW; ++j) a[j] = (j == k) ? -f : fma(-f, pk[j], a[j]);
            }
        }
        if (lane < GJW) {
#pragma unroll
            for (int j = 0; j < GJW; ++j) s_inv[lane][j] = a[j];
        }
        if (lane == 0) {
            s_cnt = __popcll(bal);
            if (sing != 0 && info[b] == 0) info[b] = sing;
        }
        if (lane < w) piv[b * n + c + lane] = mypiv;
    }
    __syncthreads();
    const int cnt = s_cnt;
    // ---- phase 3: panel rows and side range ----
#pragma unroll 1
    for (int q = 0; q < R; ++q) {
        asm volatile("" ::: "memory");   // keeps the LDS operands from being hoisted out of the loop
        const int r0 = q * WG_THREADS, r = r0 + t;
        int e_src = -1;
        for (int e = 0; e < cnt; ++e)
            if (s_dst[e] == r) e_src = e;
        const bool inK = (r >= c && r < c + w);
        double g[GJW], o[GJW];
        {
            double a[GJW];
            tile_load(r0, c, w, g);
            to_row(g, a);
            if (e_src >= 0) {
#pragma unroll
                for (int j = 0; j < GJW; ++j) a[j] = s_mv[0][e_src][j];
            }
#pragma unroll
            for (int j = 0; j < GJW; ++j) o[j] = 0.0;
            if (inK) {
#pragma unroll
                for (int j = 0; j < GJW; ++j) o[j] = s_inv[r - c][j];
            } else {
#pragma unroll
                for (int kk = 0; kk < GJW; ++kk) {
#pragma unroll
                    for (int j = 0; j < GJW; ++j) o[j] = fma(-a[kk], s_inv[kk][j], o[j]);
                }
            }
        }
        from_row(o, g);
        tile_store(r0, c, w, g);
        if (post) {
            double x[GJW];
            tile_load(r0, xa0, xw, g);
            to_row(g, x);
            if (inK) {
#pragma unroll
                for (int j = 0; j < GJW; ++j) x[j] = 0.0;
            } else if (e_src >= 0) {
#pragma unroll
                for (int j = 0; j < GJW; ++j) x[j] = s_mv[1][e_src][j];
            }
#pragma unroll
            for (int kk = 0; kk < GJW; ++kk) {
                if (kk < w) {
#pragma unroll
                    for (int j = 0; j < GJW; ++j) x[j] = fma(o[kk], s_T[kk][j], x[j]);
                }
            }
            from_row(x, g);
            tile_store(r0, xa0, xw, g);
        }
    }
    // ---- phase 4: row moves in the other columns ----
    if (cnt > 0) {
        for (int jc = t; jc < n; jc += WG_THREADS) {
            if ((jc >= c && jc < c + w) || (post && jc >= xa0 && jc < xa0 + xw)) continue;
            // all sources of this column are read before any destination is
            // written (the two sets overlap); 2w <= 32 values per column
            double xm[2 * GJW];
#pragma unroll
            for (int e = 0; e < 2 * GJW; ++e) xm[e] = (e < cnt) ? M[(int64_t)s_src[e] * ld + jc] : 0.0;
#pragma unroll
            for (int e = 0; e < 2 * GJW; ++e)
                if (e < cnt) M[(int64_t)s_dst[e] * ld + jc] = xm[e];
        }
    }
}

namespace {
std::atomic<int> g_gj_pivoting{[] {
    const char* e = std::getenv("SEHF_GJ_LEAF");
    return (e && std::string(e) == "partial") ? 1 : 0;
}()};
}  // namespace
void set_gj_pivoting(int mode) {
    if (mode < 0 || mode > 1) throw std::runtime_error("Gauss-Jordan pivoting mode must be 0 or 1");
    g_gj_pivoting.store(mode);
}
int gj_pivoting() { return g_gj_pivoting.load(); }

size_t gj_tbuf_elems(int n) { return static_cast<size_t>(n) * n; }

size_t gj_nom_elems(int n) { return static_cast<size_t>((n + LEAF_BLK - 1) / LEAF_BLK) * GJW; }
size_t gj_stg_elems() { return STG_D; }
size_t gj_stgi_elems() { return STG_I; }

// One leaf: pivots [c, c+w).  pre_w > 0: apply the left sibling's block
// [pre_c, pre_c+pre_w) to this panel first (T staged by that sibling's
// merge).  The merge stages chosen rows x [xa0, xa0+xw); post = 1 applies
// this block to those columns, post = 0 leaves them to the sibling.
static void gj_leaf(double* mat, int n, int64_t ld, int64_t stride, int batch, int c, int w, int pre_c,
                    int pre_w, int xa0, int xw, int post, const GjWork& W, hipStream_t s) {
    double fl = 2.0 * n * static_cast<double>(w) * w;
    if (pre_w > 0) fl += 2.0 * n * static_cast<double>(pre_w) * w;
    if (post) fl += 2.0 * n * static_cast<double>(w) * xw;
    ProfScope ps(s, fl * batch, 1);
    // Partial pivoting (orders up to 2 * WG_THREADS): the one-workgroup leaf,
    // 1.3x the inverse time of the tournament leaf at order 1566, batch 48.
    if (g_gj_pivoting.load(std::memory_order_relaxed) == 1 && n <= 2 * WG_THREADS) {
        if (n <= WG_THREADS)
            hipLaunchKernelGGL(k_leaf_wg<1>, dim3(batch), dim3(WG_THREADS), 0, s, mat, ld, stride, n, c, w, pre_c,
                               pre_w, xa0, xw, post, W.stg, W.piv, W.info);
        else
            hipLaunchKernelGGL(k_leaf_wg<2>, dim3(batch), dim3(WG_THREADS), 0, s, mat, ld, stride, n, c, w, pre_c,
                               pre_w, xa0, xw, post, W.stg, W.piv, W.info);
        return;
    }
    const int nblk = (n + LEAF_BLK - 1) / LEAF_BLK;
    const int nc = (nblk * GJW + 63) / 64;
#define SEHF_NOMINATE(NC)                                                                                   \
    hipLaunchKernelGGL((k_leaf_nominate<NC>), dim3(nblk, batch), dim3(LEAF_BLK), 0, s, mat, ld, stride, n, c, \
                       w, pre_c, pre_w, xa0, xw, W.stg, W.stgi, W.nom, nblk, W.cnt, W.piv, W.info)
    if (nc <= 1) SEHF_NOMINATE(1);
    else if (nc <= 2) SEHF_NOMINATE(2);
    else if (nc <= 4) SEHF_NOMINATE(4);
    else if (nc <= 8) SEHF_NOMINATE(8);
    else if (nc <= 16) SEHF_NOMINATE(16);
    else throw std::runtime_error("Gauss-Jordan leaf: matrix order above 16384");
#undef SEHF_NOMINATE
    hipLaunchKernelGGL(k_leaf_apply, dim3(nblk, batch), dim3(LEAF_BLK), 0, s, mat, ld, stride, n, c, w, xa0, xw,
                       post, W.stg, W.stgi);
}

// Block step of pivot block K = [k0, k0+nk) applied to columns J = [j0, j0+nj):
//   rows O:  X(O, J) += X(O, K) T ;  rows K: X(K, J) = X(K, K) T ;  T = X(K, J)_old
static void gj_update(double* mat, int n, int64_t ld, int64_t stride, int batch, int k0, int nk,
                      int j0, int nj, const GjWork& W, hipStream_t s) {
    const int64_t ts = static_cast<int64_t>(gj_tbuf_elems(n));
    const int64_t ne = static_cast<int64_t>(nk) * nj;
    dim3 cg(static_cast<unsigned>((ne + 255) / 256), batch);
    hipLaunchKernelGGL(k_copy_block, cg, dim3(256), 0, s, mat, ld, stride, k0, nk, j0, nj, W.tbuf, ts);
    Gemm g;
    g.M = nj; g.N = n; g.K = nk;
    g.A = W.tbuf; g.lda = nj; g.sA = ts; g.tA = 0;
    g.B = mat + k0; g.ldb = ld; g.sB = stride; g.tB = 0;
    g.C = mat + j0; g.ldc = ld; g.sC = stride;
    g.alpha = 1.0; g.beta = 1.0; g.zc0 = k0; g.zc1 = k0 + nk;
    g.batch = batch;
    gemm(g, s);
}

static void gj_rec(double* mat, int n, int64_t ld, int64_t stride, int batch, int c, int w,
                   const GjWork& W, hipStream_t s) {
    if (w <= GJW) { gj_leaf(mat, n, ld, stride, batch, c, w, 0, 0, 0, 0, 0, W, s); return; }
    int w1 = (w / 2 / GJW) * GJW;
    if (w1 < GJW) w1 = GJW;
    if (w - w1 <= GJW) {
        // two leaves: the block steps between them are fused into the leaves
        gj_leaf(mat, n, ld, stride, batch, c, w1, 0, 0, c + w1, w - w1, 0, W, s);
        gj_leaf(mat, n, ld, stride, batch, c + w1, w - w1, c, w1, c, w1, 1, W, s);
        return;
    }
    gj_rec(mat, n, ld, stride, batch, c, w1, W, s);
    gj_update(mat, n, ld, stride, batch, c, w1, c + w1, w - w1, W, s);
    gj_rec(mat, n, ld, stride, batch, c + w1, w - w1, W, s);
    gj_update(mat, n, ld, stride, batch, c + w1, w - w1, c, w1, W, s);
}

void gj_inverse(double* mat, double* out, int n, int64_t ld, int64_t stride, int batch,
                bool symmetrize, const GjWork& W, hipStream_t s) {
    if (n <= 0 || batch <= 0) return;
    gj_rec(mat, n, ld, stride, batch, 0, n, W, s);
    const int staged = (n <= 8192) ? 1 : 0;
    hipLaunchKernelGGL(k_gj_perm, dim3(batch), dim3(256), (staged ? 2 : 1) * sizeof(int) * n, s, W.piv, W.perm, n,
                       staged);
    const int nt = (n + PT - 1) / PT;
    if (symmetrize) {
        hipLaunchKernelGGL(k_gj_final_sym, dim3(nt * (nt + 1) / 2, 1, batch), dim3(256), 0, s, mat, out, ld, stride, n,
                           W.perm);
    } else {
        hipLaunchKernelGGL(k_gj_final, dim3(nt, nt, batch), dim3(256), 0, s, mat, out, ld, stride, n, W.perm, 0);
    }
}

// ---------------------------------------------------------------------------
// Elementwise helpers.
// ---------------------------------------------------------------------------
__global__ void k_fill(double* p, size_t n, double v) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_int(int* p, size_t n, int v) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
void fill(double* p, size_t n, double v, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
}
void fill_int(int* p, size_t n, int v, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill_int, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
}
__global__ void k_axpbyi(double* Y, int64_t ldy, int64_t sY, const double* X, int64_t ldx, int64_t sX,
                         double a, const double* Z, int64_t ldz, int64_t sZ, double c, double d, int n,
                         int m) {
    const int64_t b = blockIdx.z;
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n || j >= m) return;
    double v = a * X[b * sX + i + (int64_t)j * ldx];
    if (Z) v += c * Z[b * sZ + i + (int64_t)j * ldz];
    if (i == j) v += d;
    Y[b * sY + i + (int64_t)j * ldy] = v;
}
void axpbyi(double* Y, int64_t ldy, int64_t sY, const double* X, int64_t ldx, int64_t sX, double a,
            const double* Z, int64_t ldz, int64_t sZ, double c, double d, int n, int m, int batch,
            hipStream_t s) {
    dim3 g((n + 255) / 256, m, batch);
    hipLaunchKernelGGL(k_axpbyi, g, dim3(256), 0, s, Y, ldy, sY, X, ldx, sX, a, Z, ldz, sZ, c, d, n, m);
}

__global__ void k_shifted_copy(double* Y, int64_t ldy, int64_t sY, const double* M0, int64_t ldm,
                               const double* shift, int n) {
    const int64_t b = blockIdx.z;
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n) return;
    double v = M0[i + (int64_t)j * ldm];
    if (i == j) v += shift[b];
    Y[b * sY + i + (int64_t)j * ldy] = v;
}
void shifted_copy(double* Y, int64_t ldy, int64_t sY, const double* M0, int64_t ldm,
                  const double* shift, int n, int batch, hipStream_t s) {
    dim3 g((n + 255) / 256, n, batch);
    hipLaunchKernelGGL(k_shifted_copy, g, dim3(256), 0, s, Y, ldy, sY, M0, ldm, shift, n);
}

__global__ void k_transpose_scaled(double* Y, int64_t ldy, int64_t sY, const double* X, int64_t ldx,
                                   int64_t sX, double a, int r, int c) {
    __shared__ double tile[32][33];
    const int64_t b = blockIdx.z;
    const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int yy = ty; yy < 32; yy += 8) {
        const int i = i0 + tx, j = j0 + yy;
        if (i < r && j < c) tile[yy][tx] = X[b * sX + i + (int64_t)j * ldx];
    }
    __syncthreads();
    for (int yy = ty; yy < 32; yy += 8) {
        const int j = j0 + tx, i = i0 + yy;   // Y(j, i)
        if (i < r && j < c) Y[b * sY + j + (int64_t)i * ldy] = a * tile[tx][yy];
    }
}
void transpose_scaled(double* Y, int64_t ldy, int64_t sY, const double* X, int64_t ldx,
                      int64_t sX, double a, int r, int c, int batch, hipStream_t s) {
    dim3 g((r + 31) / 32, (c + 31) / 32, batch);
    hipLaunchKernelGGL(k_transpose_scaled, g, dim3(256), 0, s, Y, ldy, sY, X, ldx, sX, a, r, c);
}

__global__ void k_col_scale(double* Y, int64_t ldy, const double* X, int64_t ldx, const double* d,
                            int r, int c) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i < r && j < c) Y[i + (int64_t)j * ldy] = X[i + (int64_t)j * ldx] * d[j];
}
void col_scale(double* Y, int64_t ldy, const double* X, int64_t ldx, const double* d, int r, int c,
               hipStream_t s) {
    dim3 g((r + 255) / 256, c);
    hipLaunchKernelGGL(k_col_scale, g, dim3(256), 0, s, Y, ldy, X, ldx, d, r, c);
}

__device__ inline void atomic_max_pos(double* addr, double v) {
    // v >= 0: IEEE order of non-negative doubles matches their bit order
    atomicMax(reinterpret_cast<unsigned long long*>(addr), __double_as_longlong(v));
}
__global__ void k_max_asym(const double* X, int64_t ld, int64_t sX, int n, double* out) {
    const int64_t b = blockIdx.z;
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    double v = 0.0;
    if (i < n && j < n) v = fabs(X[b * sX + i + (int64_t)j * ld] - X[b * sX + j + (int64_t)i * ld]);
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    if ((threadIdx.x & 63) == 0 && v > 0.0) atomic_max_pos(out, v);
}
void max_asym(const double* X, int64_t ld, int64_t sX, int n, int batch, double* out, hipStream_t s) {
    dim3 g((n + 255) / 256, n, batch);
    hipLaunchKernelGGL(k_max_asym, g, dim3(256), 0, s, X, ld, sX, n, out);
}

}  // namespace sehf
