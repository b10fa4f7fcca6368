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
// row-major view.  Thread t owns rows t, t+T, ... (R rows).
// ---------------------------------------------------------------------------
constexpr int GJW = 16;

template <int R>
__global__ __launch_bounds__(1024) void k_gj_panel(double* mat, int64_t ld, int64_t stride, int n,
                                                   int c, int w, int* piv, int* info) {
    const int64_t b = blockIdx.x;
    double* __restrict__ M = mat + b * stride;
    int* pv = piv + b * n;
    const int T = blockDim.x, t = threadIdx.x, lane = t & 63, wv = t >> 6, nw = T >> 6;
    __shared__ double s_wbest[16];
    __shared__ int s_widx[16];
    __shared__ int s_p;
    __shared__ double s_rowk[GJW], s_rowp[GJW], s_srow[GJW];
    double v[R][GJW];
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int r = t + q * T;
#pragma unroll
        for (int j = 0; j < GJW; ++j) v[q][j] = (r < n && j < w) ? M[(int64_t)r * ld + c + j] : 0.0;
    }
    for (int jj = 0; jj < GJW; ++jj) {
        if (jj < w) {   // w is uniform across the workgroup
        const int k = c + jj;
        // ---- pivot search: max |v(r, jj)| over rows r >= k, first index on ties
        double best = -1.0;
        int bidx = INT_MAX;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int r = t + q * T;
            if (r >= k && r < n) {
                double colv = 0.0;
#pragma unroll
                for (int j = 0; j < GJW; ++j) if (j == jj) colv = v[q][j];
                const double a = fabs(colv);
                if (a > best) { best = a; bidx = r; }
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ob = __shfl_xor(best, off);
            const int oi = __shfl_xor(bidx, off);
            if (ob > best || (ob == best && oi < bidx)) { best = ob; bidx = oi; }
        }
        if (lane == 0) { s_wbest[wv] = best; s_widx[wv] = bidx; }
        __syncthreads();
        if (t == 0) {
            double bb = s_wbest[0]; int bi = s_widx[0];
            for (int x = 1; x < nw; ++x)
                if (s_wbest[x] > bb || (s_wbest[x] == bb && s_widx[x] < bi)) { bb = s_wbest[x]; bi = s_widx[x]; }
            s_p = bi;
            pv[k] = bi;
            if (bb == 0.0 && info[b] == 0) info[b] = k + 1;
        }
        __syncthreads();
        const int p = s_p;
        // ---- exchange rows k and p (panel part through LDS)
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int r = t + q * T;
            if (r == k) {
#pragma unroll
                for (int j = 0; j < GJW; ++j) s_rowk[j] = v[q][j];
            }
            if (r == p) {
#pragma unroll
                for (int j = 0; j < GJW; ++j) s_rowp[j] = v[q][j];
            }
        }
        __syncthreads();
        // full-row swap outside the panel; each column stays with one thread
        if (p != k) {
            double* rk = M + (int64_t)k * ld;
            double* rp = M + (int64_t)p * ld;
            for (int j = t; j < n; j += T) {
                if (j >= c && j < c + w) continue;
                const double x = rk[j];
                rk[j] = rp[j];
                rp[j] = x;
            }
        }
        // ---- elimination: pivot row is old row p
        const double pinv = 1.0 / s_rowp[jj];
        __syncthreads();
        if (t < GJW) s_srow[t] = s_rowp[t] * pinv;   // scaled pivot row
        __syncthreads();
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int r = t + q * T;
            if (r >= n) continue;
            if (r == p && p != k) {
#pragma unroll
                for (int j = 0; j < GJW; ++j) v[q][j] = s_rowk[j];
            }
            if (r == k) {
#pragma unroll
                for (int j = 0; j < GJW; ++j) v[q][j] = (j == jj) ? pinv : s_srow[j];
            } else {
                double f = 0.0;
#pragma unroll
                for (int j = 0; j < GJW; ++j) if (j == jj) f = v[q][j];
#pragma unroll
                for (int j = 0; j < GJW; ++j) v[q][j] = (j == jj) ? -f * pinv : v[q][j] - f * s_srow[j];
            }
        }
        __syncthreads();
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

static void gj_panel(double* mat, int n, int64_t ld, int64_t stride, int batch, int c, int w,
                     const GjWork& W, hipStream_t s) {
    ProfScope ps(s, 2.0 * n * static_cast<double>(w) * w * batch, 1);
    if (n <= 256) hipLaunchKernelGGL(k_gj_panel<1>, dim3(batch), dim3(256), 0, s, mat, ld, stride, n, c, w, W.piv, W.info);
    else if (n <= 512) hipLaunchKernelGGL(k_gj_panel<1>, dim3(batch), dim3(512), 0, s, mat, ld, stride, n, c, w, W.piv, W.info);
    else if (n <= 1024) hipLaunchKernelGGL(k_gj_panel<1>, dim3(batch), dim3(1024), 0, s, mat, ld, stride, n, c, w, W.piv, W.info);
    else if (n <= 2048) hipLaunchKernelGGL(k_gj_panel<2>, dim3(batch), dim3(1024), 0, s, mat, ld, stride, n, c, w, W.piv, W.info);
    else if (n <= 4096) hipLaunchKernelGGL(k_gj_panel<4>, dim3(batch), dim3(1024), 0, s, mat, ld, stride, n, c, w, W.piv, W.info);
    else hipLaunchKernelGGL(k_gj_panel<12>, dim3(batch), dim3(1024), 0, s, mat, ld, stride, n, c, w, W.piv, W.info);
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
    if (w <= GJW) { gj_panel(mat, n, ld, stride, batch, c, w, W, s); return; }
    int w1 = (w / 2 / GJW) * GJW;
    if (w1 < GJW) w1 = GJW;
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
