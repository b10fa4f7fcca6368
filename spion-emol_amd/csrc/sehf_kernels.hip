// sehf_kernels.hip -- gfx950 FP64 kernels of the static-exchange solver.
//
//  * k_dgemm       : batched FP64 GEMM on v_mfma_f64_16x16x4f64, 64x64 output
//                    tile per 256-thread workgroup (4 waves x 32x32), K staged
//                    through LDS in slabs of 16.
//  * Gauss-Jordan  : in-place inversion with partial pivoting, recursive over
//                    column blocks.  Leaves (<=16 columns) are factored by one
//                    1024-thread workgroup per matrix with the panel held in
//                    registers; each pivot swaps the two full rows at once so
//                    no deferred row interchanges exist.  Interior nodes are
//                    one MFMA GEMM over all rows (rank-w block update).
//
// Gauss-Jordan works on the row-major view of the column-major buffer
// (element (i,j) at base[i*ld + j], i.e. on M^T).  Row pivoting of M^T is
// column pivoting of M, and the row-major result of inv(M^T) is inv(M) in
// column-major order, so callers pass and receive ordinary column-major data.
#include "sehf_kernels.h"

#include <climits>
#include <cstdlib>
#include <stdexcept>
#include <vector>

namespace sehf {

typedef double d4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// GEMM
// ---------------------------------------------------------------------------
constexpr int GBM = 64, GBN = 64, GBK = 16;

struct GemmK {
    int M, N, K;
    const double* A; int64_t lda, sA; int tA;
    const double* B; int64_t ldb, sB; int tB;
    double* C; int64_t ldc, sC;
    double alpha, beta;
    int zc0, zc1;
};

__global__ __launch_bounds__(256) void k_dgemm(GemmK p) {
    const int64_t b = blockIdx.z;
    const double* __restrict__ A = p.A + b * p.sA;
    const double* __restrict__ B = p.B + b * p.sB;
    double* __restrict__ C = p.C + b * p.sC;
    const int m0 = blockIdx.x * GBM, n0 = blockIdx.y * GBN;
    __shared__ double As[GBK][GBM + 2];   // [k][i] of op(A)
    __shared__ double Bs[GBK][GBN + 2];   // [k][j] of op(B)
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int wm = (wv & 1) * 32, wn = (wv >> 1) * 32;
    d4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
    double ra[4], rb[4];
    // Element assignment of the 64x16 slabs: 4 loads per thread, coalesced
    // along the contiguous dimension of the stored operand.
    auto a_ik = [&](int r, int& i, int& k) {
        if (!p.tA) { i = t & 63; k = (t >> 6) + 4 * r; } else { k = t & 15; i = (t >> 4) + 16 * r; }
    };
    auto b_kj = [&](int r, int& k, int& j) {
        if (!p.tB) { k = t & 15; j = (t >> 4) + 16 * r; } else { j = t & 63; k = (t >> 6) + 4 * r; }
    };
    auto load = [&](int k0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int i, k; a_ik(r, i, k);
            const int gi = m0 + i, gk = k0 + k;
            double v = 0.0;
            if (gi < p.M && gk < p.K) v = p.tA ? A[gk + (int64_t)gi * p.lda] : A[gi + (int64_t)gk * p.lda];
            ra[r] = v;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
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
        for (int r = 0; r < 4; ++r) { int i, k; a_ik(r, i, k); As[k][i] = ra[r]; }
#pragma unroll
        for (int r = 0; r < 4; ++r) { int k, j; b_kj(r, k, j); Bs[k][j] = rb[r]; }
        __syncthreads();
        if (k0 + GBK < p.K) load(k0 + GBK);
#pragma unroll
        for (int kk = 0; kk < GBK; kk += 4) {
            const int kq = kk + (lane >> 4);
            double av[2], bv[2];
#pragma unroll
            for (int bi = 0; bi < 2; ++bi) av[bi] = As[kq][wm + bi * 16 + (lane & 15)];
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) bv[bj] = Bs[kq][wn + bj * 16 + (lane & 15)];
            // MFMA computes the transposed 16x16 block: its A operand is the
            // op(B) column slab, its B operand the op(A) row slab.  Result
            // register v of lane l holds D(4v + l/16, l%16), i.e.
            // C(row l%16, col 4v + l/16), so stores run along C's contiguous
            // row index.
#pragma unroll
            for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                for (int bj = 0; bj < 2; ++bj)
                    acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[bj], av[bi], acc[bi][bj], 0, 0, 0);
        }
    }
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = m0 + wm + bi * 16 + (lane & 15);
                const int col = n0 + wn + bj * 16 + 4 * v + (lane >> 4);
                if (row < p.M && col < p.N) {
                    double* c = C + row + (int64_t)col * p.ldc;
                    const double be = (col >= p.zc0 && col < p.zc1) ? 0.0 : p.beta;
                    double val = p.alpha * acc[bi][bj][v];
                    if (be != 0.0) val += be * (*c);
                    *c = val;
                }
            }
}

// ---------------------------------------------------------------------------
// Optional per-launch timing (HIP events on the launch stream).
// ---------------------------------------------------------------------------
namespace {
struct ProfRec { hipEvent_t a, b; double flops; int cls; };
bool g_prof = false;
int g_every = 1;
long long g_count[2] = {0, 0};
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
    g_prof = true;
    g_every = every > 0 ? every : 1;
    g_count[0] = g_count[1] = 0;
    g_recs.clear();
}

void prof_end(double* out) {
    g_prof = false;
    for (int i = 0; i < 6; ++i) out[i] = 0.0;
    for (auto& r : g_recs) {
        (void)hipEventSynchronize(r.b);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, r.a, r.b);
        out[3 * r.cls + 0] += ms;
        out[3 * r.cls + 1] += r.flops;
        out[3 * r.cls + 2] += 1.0;
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
        if (on) on = (g_count[cls]++ % g_every) == 0;
        if (on) { r.a = take_event(); r.b = take_event(); r.flops = flops; r.cls = cls; (void)hipEventRecord(r.a, s); }
    }
    ~ProfScope() {
        if (on) { (void)hipEventRecord(r.b, s); g_recs.push_back(r); }
    }
};

void gemm(const Gemm& g, hipStream_t s) {
    if (g.M <= 0 || g.N <= 0 || g.batch <= 0) return;
    GemmK p{g.M, g.N, g.K, g.A, g.lda, g.sA, g.tA, g.B, g.ldb, g.sB, g.tB, g.C, g.ldc, g.sC,
            g.alpha, g.beta, g.zc0, g.zc1};
    dim3 grid((g.M + GBM - 1) / GBM, (g.N + GBN - 1) / GBN, g.batch);
    ProfScope ps(s, 2.0 * g.M * g.N * static_cast<double>(g.K) * g.batch, 0);
    hipLaunchKernelGGL(k_dgemm, grid, dim3(256), 0, s, p);
}

// ---------------------------------------------------------------------------
// Gauss-Jordan leaf: pivots c .. c+w-1 (w <= 16) on columns [c, c+w) of the
// row-major view, one workgroup per matrix, thread t owning rows t, t+T, ...
// (R rows) of the panel in registers.
//
// Per pivot column: a max-|.| search (wave shuffles, then one LDS exchange
// between waves), one LDS broadcast of the pivot row and of row k, and the
// rank-1 elimination in registers -- two barriers and no global memory
// traffic.  The row interchanges of the columns outside the panel are
// applied once at the end as one composite permutation of the <= 2w rows
// involved (loads of all source rows, then stores).
//
// Optional fused sibling updates (the recursion's block steps for a node of
// two leaves):
//   pre  (pre_w > 0): before factoring, apply pivot block Kp = [pre_c,
//        pre_c+pre_w) to this panel:  X(O,J) += X(O,Kp) T,  X(Kp,J) =
//        X(Kp,Kp) T,  T = X(Kp,J)_old.
//   post (post_w > 0): after factoring and the row interchanges, apply this
//        pivot block K = [c, c+w) to columns Jq = [post_c, post_c+post_w).
// ---------------------------------------------------------------------------
constexpr int GJW = 16;

template <int R, int T>
__global__ __launch_bounds__(T) void k_gj_leaf(double* mat, int64_t ld, int64_t stride, int n, int c,
                                               int w, int pre_c, int pre_w, int post_c, int post_w,
                                               int* piv, int* info) {
    constexpr int NW = T / 64;
    const int64_t b = blockIdx.x;
    double* __restrict__ M = mat + b * stride;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    __shared__ double s_best[2][NW];
    __shared__ int s_idx[2][NW];
    __shared__ double s_rowp[2][GJW], s_rowk[2][GJW];
    __shared__ double s_T[GJW][GJW];
    __shared__ int s_piv[GJW];
    __shared__ int s_src[2 * GJW], s_dst[2 * GJW];
    __shared__ int s_cnt, s_sing;
    if (t == 0) s_sing = 0;

    double v[R][GJW];
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int r = t + q * T;
#pragma unroll
        for (int j = 0; j < GJW; ++j) v[q][j] = (r < n && j < w) ? M[(int64_t)r * ld + c + j] : 0.0;
    }

    // ---- fused pre-update from the left sibling's pivot block
    if (pre_w > 0) {
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int r = t + q * T;
            if (r >= pre_c && r < pre_c + pre_w) {
#pragma unroll
                for (int j = 0; j < GJW; ++j) s_T[r - pre_c][j] = v[q][j];
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int r = t + q * T;
            if (r < n) {
                const double* xr = M + (int64_t)r * ld + pre_c;
                double acc[GJW];
                const bool inK = (r >= pre_c && r < pre_c + pre_w);
#pragma unroll
                for (int j = 0; j < GJW; ++j) acc[j] = inK ? 0.0 : v[q][j];
#pragma unroll 2
                for (int kk = 0; kk < pre_w; ++kk) {
                    {
                        const double x = xr[kk];
#pragma unroll
                        for (int j = 0; j < GJW; ++j) acc[j] = fma(x, s_T[kk][j], acc[j]);
                    }
                }
#pragma unroll
                for (int j = 0; j < GJW; ++j) v[q][j] = (j < w) ? acc[j] : 0.0;
            }
        }
    }

    // ---- elimination.  A rolled loop over the pivot columns (the unrolled
    // form is straight-line code run once per launch, i.e. bound by
    // instruction fetch); column jj of the register panel is read with
    // selects.
#pragma unroll 1
    for (int jj = 0; jj < w; ++jj) {
        const int k = c + jj, buf = jj & 1;
        double colv[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            double f = v[q][0];
#pragma unroll
            for (int j = 1; j < GJW; ++j) f = (j == jj) ? v[q][j] : f;
            colv[q] = f;
        }
        double best = -1.0;
        int bidx = INT_MAX;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int r = t + q * T;
            const double a = fabs(colv[q]);
            if (r >= k && r < n && a > best) { best = a; bidx = r; }
        }
        const double mine = best;
        // wave max, then the lowest lane holding it
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) best = fmax(best, __shfl_xor(best, off));
        {
            const unsigned long long bal = __ballot(mine == best);
            const int src = bal ? __ffsll(static_cast<long long>(bal)) - 1 : 0;
            bidx = __shfl(bidx, src);
        }
        if (lane == 0) { s_best[buf][wv] = best; s_idx[buf][wv] = bidx; }
        __syncthreads();
        // every wave reduces the NW wave results itself
        double bb = (lane < NW) ? s_best[buf][lane] : -2.0;
        int bi = (lane < NW) ? s_idx[buf][lane] : INT_MAX;
        double mx = bb;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            if (off < NW) mx = fmax(mx, __shfl_xor(mx, off));
        }
        mx = __shfl(mx, 0);
        {
            const unsigned long long bal = __ballot(lane < NW && bb == mx);
            const int src = __ffsll(static_cast<long long>(bal)) - 1;
            bi = __shfl(bi, src);
        }
        // no finite candidate (a column of NaN): keep row k, flag singular
        const int p = (bi == INT_MAX) ? k : bi;
        // publish rows p and k of the panel
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int r = t + q * T;
            if (r == p) {
#pragma unroll
                for (int j = 0; j < GJW; ++j) s_rowp[buf][j] = v[q][j];
            }
            if (r == k) {
#pragma unroll
                for (int j = 0; j < GJW; ++j) s_rowk[buf][j] = v[q][j];
            }
        }
        if (t == 0) {
            s_piv[jj] = p;
            if (!(mx > 0.0) && s_sing == 0) s_sing = k + 1;
        }
        __syncthreads();
        const double pinv = 1.0 / s_rowp[buf][jj];
        double prow[GJW];
#pragma unroll
        for (int j = 0; j < GJW; ++j) prow[j] = s_rowp[buf][j] * pinv;   // scaled pivot row
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int r = t + q * T;
            if (r == k) {
#pragma unroll
                for (int j = 0; j < GJW; ++j) v[q][j] = (j == jj) ? pinv : prow[j];
            } else {
                double f = colv[q];
                if (r == p) {
#pragma unroll
                    for (int j = 0; j < GJW; ++j) v[q][j] = s_rowk[buf][j];
                    f = s_rowk[buf][jj];
                }
#pragma unroll
                for (int j = 0; j < GJW; ++j) v[q][j] = (j == jj) ? -f * pinv : fma(-f, prow[j], v[q][j]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int r = t + q * T;
        if (r < n) {
#pragma unroll
            for (int j = 0; j < GJW; ++j)
                if (j < w) M[(int64_t)r * ld + c + j] = v[q][j];
        }
    }

    // ---- composite row interchange for the columns outside the panel.
    // Slots: s < w is row c+s; w+i is pivot row p_i when p_i lies outside
    // the panel rows and is its first occurrence.  Wave 0 follows the w
    // transpositions (slot jj <-> slot of p_jj) on lane-held slot contents.
    if (wv == 0) {
        int bslot = 0;
        int myrow = -1;
        if (lane < w) {
            const int p = s_piv[lane];
            if (p < c + w) bslot = p - c;
            else {
                int f = lane;
                for (int j2 = 0; j2 < lane; ++j2)
                    if (s_piv[j2] == p) { f = j2; break; }
                bslot = w + f;
            }
            myrow = c + lane;
        } else if (lane < 2 * w) {
            myrow = s_piv[lane - w];
        }
        int cur = lane;
        for (int jj = 0; jj < w; ++jj) {
            const int bs = __builtin_amdgcn_readlane(bslot, jj);
            const int va = __builtin_amdgcn_readlane(cur, jj);
            const int vb = __builtin_amdgcn_readlane(cur, bs);
            if (lane == jj) cur = vb;
            else if (lane == bs) cur = va;
        }
        const int srcrow = __shfl(myrow, cur < 64 ? cur : 0);
        const bool mv = lane < 2 * w && cur != lane;
        const unsigned long long bal = __ballot(mv);
        if (mv) {
            const int pos = __popcll(bal & ((1ull << lane) - 1ull));
            s_dst[pos] = myrow;
            s_src[pos] = srcrow;
        }
        if (lane == 0) {
            s_cnt = __popcll(bal);
            if (s_sing != 0 && info[b] == 0) info[b] = s_sing;
        }
        if (lane < w) piv[b * n + c + lane] = s_piv[lane];
    }
    __syncthreads();
    const int cnt = s_cnt;
    if (cnt > 0) {
        for (int j = t; j < n; j += T) {
            if (j >= c && j < c + w) continue;
            double x[2 * GJW];
#pragma unroll
            for (int e = 0; e < 2 * GJW; ++e)
                if (e < cnt) x[e] = M[(int64_t)s_src[e] * ld + j];
#pragma unroll
            for (int e = 0; e < 2 * GJW; ++e)
                if (e < cnt) M[(int64_t)s_dst[e] * ld + j] = x[e];
        }
    }

    // ---- fused post-update of columns Jq with this pivot block
    if (post_w > 0) {
        __syncthreads();
        if (t < GJW * GJW) {
            const int kk = t / GJW, j = t % GJW;
            s_T[kk][j] = (kk < w && j < post_w) ? M[(int64_t)(c + kk) * ld + post_c + j] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int r = t + q * T;
            if (r < n) {
                double* xr = M + (int64_t)r * ld + post_c;
                const double* xk = M + (int64_t)r * ld + c;   // X(r, K), final (re-read: frees v for the swaps)
                const bool inK = (r >= c && r < c + w);
                double acc[GJW];
#pragma unroll
                for (int j = 0; j < GJW; ++j) acc[j] = (!inK && j < post_w) ? xr[j] : 0.0;
#pragma unroll 2
                for (int kk = 0; kk < w; ++kk) {
                    {
                        const double x = xk[kk];
#pragma unroll
                        for (int j = 0; j < GJW; ++j) acc[j] = fma(x, s_T[kk][j], acc[j]);
                    }
                }
#pragma unroll
                for (int j = 0; j < GJW; ++j)
                    if (j < post_w) xr[j] = acc[j];
            }
        }
    }
}

// Copy the block rows K (nk rows from k0) x columns J (nj from j0) of the
// row-major view into T (row-major nk x nj).
__global__ void k_copy_block(const double* mat, int64_t ld, int64_t stride, int k0, int nk, int j0,
                             int nj, double* tb, int64_t tstride) {
    const int64_t b = blockIdx.z;
    const double* M = mat + b * stride;
    double* T = tb + b * tstride;
    const int jj = blockIdx.x * blockDim.x + threadIdx.x;
    const int kk = blockIdx.y;
    if (jj < nj) T[(int64_t)kk * nj + jj] = M[(int64_t)(k0 + kk) * ld + j0 + jj];
}

// Column-interchange composition: perm = S_{n-1} ... S_0 applied to identity.
__global__ void k_gj_perm(const int* piv, int* perm, int n) {
    extern __shared__ int sp[];
    const int64_t b = blockIdx.x;
    const int* pv = piv + b * n;
    int* pm = perm + b * n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) sp[i] = i;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = n - 1; k >= 0; --k) {
            const int q = pv[k];
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

size_t gj_tbuf_elems(int n) { return static_cast<size_t>(n) * n; }

// Leaf launch: T threads, R = rows per thread (power of two covering n).
template <int T>
static void leaf_launch(int R, double* mat, int64_t ld, int64_t stride, int n, int c, int w, int pre_c,
                        int pre_w, int post_c, int post_w, const GjWork& W, int batch, hipStream_t s) {
#define SEHF_LEAF(RR)                                                                                   \
    hipLaunchKernelGGL((k_gj_leaf<RR, T>), dim3(batch), dim3(T), 0, s, mat, ld, stride, n, c, w, pre_c, \
                       pre_w, post_c, post_w, W.piv, W.info)
    switch (R) {
    case 1: SEHF_LEAF(1); break;
    case 2: SEHF_LEAF(2); break;
    case 4: SEHF_LEAF(4); break;
    case 8: SEHF_LEAF(8); break;
    default: SEHF_LEAF(16); break;
    }
#undef SEHF_LEAF
}

static int g_leaf_threads = [] {   // 0: size-based choice; SEHF_GJ_LEAF_THREADS overrides
    const char* e = std::getenv("SEHF_GJ_LEAF_THREADS");
    return e ? std::atoi(e) : 0;
}();

void gj_set_leaf_threads(int t) { g_leaf_threads = t; }

static void gj_leaf(double* mat, int n, int64_t ld, int64_t stride, int batch, int c, int w, int pre_c,
                    int pre_w, int post_c, int post_w, const GjWork& W, hipStream_t s) {
    double fl = 2.0 * n * static_cast<double>(w) * w;
    if (pre_w > 0) fl += 2.0 * n * static_cast<double>(pre_w) * w;
    if (post_w > 0) fl += 2.0 * n * static_cast<double>(w) * post_w;
    ProfScope ps(s, fl * batch, 1);
    int T = g_leaf_threads;
    // registers: about 37 VGPRs per panel row held, 256 at most without spills
    if (T != 256 && T != 512 && T != 1024) T = (n <= 256) ? 256 : (n <= 512) ? 512 : (n <= 1024) ? 1024 : (n <= 2048) ? 512 : 1024;
    int R = 1;
    while (R * T < n) R *= 2;
    if (R > 16) {
        T = 1024;
        R = 1;
        while (R * T < n) R *= 2;
        if (R > 16) throw std::runtime_error("Gauss-Jordan leaf: matrix order above 16384");
    }
    if (T == 256) leaf_launch<256>(R, mat, ld, stride, n, c, w, pre_c, pre_w, post_c, post_w, W, batch, s);
    else if (T == 512) leaf_launch<512>(R, mat, ld, stride, n, c, w, pre_c, pre_w, post_c, post_w, W, batch, s);
    else leaf_launch<1024>(R, mat, ld, stride, n, c, w, pre_c, pre_w, post_c, post_w, W, batch, s);
}

// Block step of pivot block K = [k0, k0+nk) applied to columns J = [j0, j0+nj):
//   rows O:  X(O, J) += X(O, K) T ;  rows K: X(K, J) = X(K, K) T ;  T = X(K, J)_old
static void gj_update(double* mat, int n, int64_t ld, int64_t stride, int batch, int k0, int nk,
                      int j0, int nj, const GjWork& W, hipStream_t s) {
    const int64_t ts = static_cast<int64_t>(gj_tbuf_elems(n));
    dim3 cg((nj + 255) / 256, nk, batch);
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
    if (w <= GJW) { gj_leaf(mat, n, ld, stride, batch, c, w, 0, 0, 0, 0, W, s); return; }
    int w1 = (w / 2 / GJW) * GJW;
    if (w1 < GJW) w1 = GJW;
    if (w - w1 <= GJW) {
        // two leaves: the right one applies both block steps itself
        gj_leaf(mat, n, ld, stride, batch, c, w1, 0, 0, 0, 0, W, s);
        gj_leaf(mat, n, ld, stride, batch, c + w1, w - w1, c, w1, c, w1, W, s);
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
    hipLaunchKernelGGL(k_gj_perm, dim3(batch), dim3(256), sizeof(int) * n, s, W.piv, W.perm, n);
    dim3 fg((n + PT - 1) / PT, (n + PT - 1) / PT, batch);
    hipLaunchKernelGGL(k_gj_final, fg, dim3(256), 0, s, mat, out, ld, stride, n, W.perm, symmetrize ? 1 : 0);
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
