// sehf_spd.hip -- batched inverse of symmetric positive definite matrices.
//
// The Schur complement S_n = A_n - B_n D_n^{-1} B_n^T of the Numerov matrix is
// symmetric positive definite: below n = 60 the stability shift makes its
// smallest eigenvalue at least W_min + 0.01 (SchurInverter.cpp:30-59), and
// beyond, A_n = I + h^2/6 (E - V_n) stays within a few percent of I while
// B D^{-1} B^T is O(h^4).  Its inverse therefore needs no pivoting: recursive
// 2 x 2 block elimination
//
//   [A  B ]^{-1}   [A^{-1} + X S^{-1} X^T   -X S^{-1}]
//   [B^T C]      = [-S^{-1} X^T              S^{-1}  ],  X = A^{-1} B, S = C - B^T X,
//
// costs 4/3 n^3 flop of deep GEMMs (against 2 n^3 for Gauss-Jordan with
// pivoting) and ends in small diagonal blocks inverted in LDS.  Each level
// reuses the lower-left block for X^T, so the only workspace is one c x a
// product per matrix (the Gauss-Jordan block-row buffer is large enough).
// A leaf pivot that is not positive (S not positive definite) sets
// info = -(1 + index) for the caller's check; the result is then meaningless.
#include <hip/hip_runtime.h>

#include "sehf_kernels.h"

namespace sehf {
namespace {

constexpr int SPD_LEAF = 64;
constexpr int SPD_LD = SPD_LEAF + 1;   // LDS leading dimension (odd: no bank aliasing along rows)

// In-place Gauss-Jordan inverse without pivoting of the m x m diagonal block at
// (off, off), m <= 64, in LDS; one workgroup of 256 threads per matrix.  Thread
// t owns column t & 63 and rows (t >> 6) + 4 r.  Step k with pivot p:
//   M_ij -= M_ik (M_kj / p) (i, j != k),  M_kj /= p,  M_ik = -M_ik / p,  M_kk = 1 / p.
__global__ __launch_bounds__(256) void k_spd_leaf(double* M, int64_t ld, int64_t stride, int off, int m,
                                                 int* info) {
    __shared__ double L[SPD_LEAF * SPD_LD];
    double* Mb = M + static_cast<int64_t>(blockIdx.x) * stride + off + static_cast<int64_t>(off) * ld;
    const int t = threadIdx.x;
    for (int e = t; e < m * m; e += 256) {
        const int i = e % m, j = e / m;
        L[i + j * SPD_LD] = Mb[i + static_cast<int64_t>(j) * ld];
    }
    __syncthreads();
    const int j = t & 63, i0 = t >> 6;
    int bad = 0;
    for (int k = 0; k < m; ++k) {
        const double p = L[k + k * SPD_LD];
        const double mkj = (j < m) ? L[k + j * SPD_LD] : 0.0;
        double mik[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int i = i0 + 4 * r;
            mik[r] = (i < m) ? L[i + k * SPD_LD] : 0.0;
        }
        __syncthreads();
        if (!(p > 0.0) && bad == 0) bad = k + 1;
        const double ip = 1.0 / p;
        if (j < m) {
            const double f = mkj * ip;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int i = i0 + 4 * r;
                if (i < m) {
                    double v;
                    if (i == k) v = (j == k) ? ip : f;
                    else if (j == k) v = -mik[r] * ip;
                    else v = L[i + j * SPD_LD] - mik[r] * f;
                    L[i + j * SPD_LD] = v;
                }
            }
        }
        __syncthreads();
    }
    if (bad != 0 && t == 0) atomicCAS(&info[blockIdx.x], 0, -(off + bad));
    for (int e = t; e < m * m; e += 256) {
        const int i = e % m, jj = e / m;
        Mb[i + static_cast<int64_t>(jj) * ld] = L[i + jj * SPD_LD];
    }
}

// M <- 1/2 (M + M^T) in place (n x n blocks, leading dim ld).
// M <- (M + M^T) / 2 over 32 x 32 tile pairs: the workgroup of pair (P, Q),
// Q <= P, reads the lower tile (P, Q) and its mirror (Q, P) coalesced through
// LDS and writes m(i, j) = m(j, i) = (m(i, j) + m(j, i)) / 2 for i > j.
constexpr int SYM_T = 32;
__global__ __launch_bounds__(256) void k_spd_sym(double* M, int64_t ld, int64_t stride, int n) {
    double* m = M + static_cast<int64_t>(blockIdx.z) * stride;
    const int t = blockIdx.x;   // t = P (P + 1) / 2 + Q
    int P = static_cast<int>((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((P + 1) * (P + 2) / 2 <= t) ++P;
    while (P * (P + 1) / 2 > t) --P;
    const int Q = t - P * (P + 1) / 2;
    const int p0 = P * SYM_T, q0 = Q * SYM_T;
    __shared__ double Lt[SYM_T][SYM_T + 1];   // Lt[jj][ii] = m(p0 + ii, q0 + jj)
    __shared__ double Ut[SYM_T][SYM_T + 1];   // Ut[jj][ii] = m(q0 + ii, p0 + jj)
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int yy = ty; yy < SYM_T; yy += 8) {
        if (p0 + tx < n && q0 + yy < n) Lt[yy][tx] = m[(p0 + tx) + static_cast<int64_t>(q0 + yy) * ld];
        if (q0 + tx < n && p0 + yy < n) Ut[yy][tx] = m[(q0 + tx) + static_cast<int64_t>(p0 + yy) * ld];
    }
    __syncthreads();
    for (int yy = ty; yy < SYM_T; yy += 8) {
        {   // lower element (i, j) = (p0 + tx, q0 + yy)
            const int i = p0 + tx, j = q0 + yy;
            if (i < n && j < n && i > j) m[i + static_cast<int64_t>(j) * ld] = 0.5 * (Lt[yy][tx] + Ut[tx][yy]);
        }
        {   // upper element (j, i) = (q0 + tx, p0 + yy), mirror of (p0 + yy, q0 + tx)
            const int i = p0 + yy, j = q0 + tx;
            if (i < n && j < n && i > j) m[j + static_cast<int64_t>(i) * ld] = 0.5 * (Lt[tx][yy] + Ut[yy][tx]);
        }
    }
}

void spd_rec(double* M, int64_t ld, int64_t stride, int off, int m, int batch, double* T, int* info,
             hipStream_t s) {
    if (m <= SPD_LEAF) {
        hipLaunchKernelGGL(k_spd_leaf, dim3(batch), dim3(256), 0, s, M, ld, stride, off, m, info);
        return;
    }
    const int a = m / 2, c = m - a;
    double* A = M + off + static_cast<int64_t>(off) * ld;
    double* UR = M + off + static_cast<int64_t>(off + a) * ld;
    double* LL = M + (off + a) + static_cast<int64_t>(off) * ld;
    double* C = M + (off + a) + static_cast<int64_t>(off + a) * ld;
    const int64_t sT = static_cast<int64_t>(c) * a;   // T: c x a, ld c
    spd_rec(M, ld, stride, off, a, batch, T, info, s);   // A <- A^{-1}
    {
        Gemm g;   // T = B^T A^{-1} = X^T
        g.M = c; g.N = a; g.K = a;
        g.A = LL; g.lda = ld; g.sA = stride; g.B = A; g.ldb = ld; g.sB = stride;
        g.C = T; g.ldc = c; g.sC = sT; g.batch = batch;
        gemm(g, s);
    }
    {
        Gemm g;   // C <- S = C - B^T X
        g.M = c; g.N = c; g.K = a;
        g.A = LL; g.lda = ld; g.sA = stride; g.B = T; g.ldb = c; g.sB = sT; g.tB = 1;
        g.C = C; g.ldc = ld; g.sC = stride; g.alpha = -1.0; g.beta = 1.0; g.batch = batch;
        gemm(g, s);
    }
    axpbyi(LL, ld, stride, T, c, sT, 1.0, nullptr, 0, 0, 0.0, 0.0, c, a, batch, s);   // LL <- X^T
    spd_rec(M, ld, stride, off + a, c, batch, T, info, s);                           // C <- S^{-1}
    {
        Gemm g;   // UR = -X S^{-1}
        g.M = a; g.N = c; g.K = c;
        g.A = LL; g.lda = ld; g.sA = stride; g.tA = 1; g.B = C; g.ldb = ld; g.sB = stride;
        g.C = UR; g.ldc = ld; g.sC = stride; g.alpha = -1.0; g.batch = batch;
        gemm(g, s);
    }
    {
        Gemm g;   // A <- A^{-1} + X S^{-1} X^T = A^{-1} - UR X^T
        g.M = a; g.N = a; g.K = c;
        g.A = UR; g.lda = ld; g.sA = stride; g.B = LL; g.ldb = ld; g.sB = stride;
        g.C = A; g.ldc = ld; g.sC = stride; g.alpha = -1.0; g.beta = 1.0; g.batch = batch;
        gemm(g, s);
    }
    transpose_scaled(LL, ld, stride, UR, ld, stride, 1.0, a, c, batch, s);   // LL <- UR^T
}

}  // namespace

void spd_inverse(double* mat, int n, int64_t ld, int64_t stride, int batch, bool symmetrize, const GjWork& w,
                 hipStream_t s) {
    if (n <= 0 || batch <= 0) return;
    spd_rec(mat, ld, stride, 0, n, batch, w.tbuf, w.info, s);
    if (symmetrize) {
        const int nt = (n + SYM_T - 1) / SYM_T;
        hipLaunchKernelGGL(k_spd_sym, dim3(nt * (nt + 1) / 2, 1, batch), dim3(256), 0, s, mat, ld, stride, n);
    }
}

}  // namespace sehf
