// sehf_kernels.h -- launch wrappers for the gfx950 FP64 kernels.
//
// All matrices are FP64.  Unless stated otherwise they are column-major with
// an explicit leading dimension, and "batched" means `batch` matrices at a
// fixed element stride.  Every wrapper enqueues on the given stream and
// returns without synchronising.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sehf {

// C[b] = alpha * alpha_b[b] * op(A[b]) * op(B[b]) + beta * C[b]   (column-major)
// alpha_b (device, one factor per batch entry) is optional; without it the
// factor is 1.  Columns of C in [zc0, zc1) take beta = 0 (pivot-block rows of
// the Gauss-Jordan update, see gj_inverse).  A stride of 0 shares an operand
// across the batch.
struct Gemm {
    int M = 0, N = 0, K = 0;
    const double* A = nullptr; int64_t lda = 0, sA = 0; int tA = 0;
    const double* B = nullptr; int64_t ldb = 0, sB = 0; int tB = 0;
    double* C = nullptr; int64_t ldc = 0, sC = 0;
    double alpha = 1.0, beta = 0.0;
    const double* alpha_b = nullptr;
    int zc0 = 0, zc1 = 0;
    int batch = 1;
};
void gemm(const Gemm& g, hipStream_t s);

// GEMM tile policy (process-wide): 0 = by shape (128 x 128 tiles over 8 waves
// where M > 64 and K >= 256, 64 x 128 over 8 waves where M > 64 and K < 256,
// 16 x 128 or 32 x 128 over 4 waves where M <= 16 or 32, else 64 x 64 over 4
// waves), 1 = 64 x 64 for every product, 2 = 128 x 128 wherever M > 64, else
// as 0.  With 128 x 128 tiles, a remainder of at
// most 64 rows past the last full tile goes to a thin-tile launch.  The
// autotuner probes the policies; C is bitwise the same under each.
void set_gemm_tiles(int policy);
int gemm_tiles();

// In-place batched inverse by blocked Gauss-Jordan elimination with partial
// pivoting.  `mat` holds `batch` column-major n x n matrices (leading dim ld,
// element stride `stride`).  The inverse is written to `out` (may NOT alias
// mat; same ld/stride), symmetrised as 1/2 (X + X^T) when `symmetrize`.
// Workspace: piv  = batch * n ints, perm = batch * n ints,
//            tbuf = batch * n * n doubles (block-row copies),
//            info = batch ints: a zero pivot at index k sets 1 + k if still 0
//            (sticky: the caller clears it, so one check covers many calls).
struct GjWork {
    int* piv = nullptr;
    int* perm = nullptr;
    double* tbuf = nullptr;
    int* info = nullptr;
    int* nom = nullptr;      // leaf nominees (gj_nom_elems per matrix)
    double* stg = nullptr;   // leaf staging (gj_stg_elems per matrix)
    int* stgi = nullptr;     // leaf move list (gj_stgi_elems per matrix)
    int* cnt = nullptr;      // leaf completion counters, one per matrix, zero between leaves
};
void gj_inverse(double* mat, double* out, int n, int64_t ld, int64_t stride, int batch,
                bool symmetrize, const GjWork& w, hipStream_t s);
// Pivot order of gj_inverse (process-wide): 0 tournament pivoting, 1 partial
// pivoting for n <= 2048 (one workgroup per matrix).  Initialised from
// $SEHF_GJ_LEAF ("partial" selects 1).
void set_gj_pivoting(int mode);
int gj_pivoting();
size_t gj_tbuf_elems(int n);   // doubles per matrix needed in tbuf
size_t gj_nom_elems(int n);    // ints per matrix: leaf nominees
size_t gj_stg_elems();         // doubles per matrix: leaf staging
size_t gj_stgi_elems();        // ints per matrix: leaf move list

// In-place batched inverse of symmetric positive definite matrices by
// recursive 2 x 2 block elimination without pivoting (4/3 n^3 flop of GEMMs,
// 64 x 64 diagonal leaves in LDS), symmetrised as 1/2 (X + X^T) when
// `symmetrize`.  Uses w.tbuf (at least (n/2 + 1)^2 doubles per matrix) and
// w.info: a non-positive leaf pivot at index k sets info = -(1 + k) if still 0.
void spd_inverse(double* mat, int n, int64_t ld, int64_t stride, int batch, bool symmetrize, const GjWork& w,
                 hipStream_t s);

// Y[b] = a * X[b] + c * Z[b] + d * I   (n x m column-major blocks; Z may be null).
void axpbyi(double* Y, int64_t ldy, int64_t sY, const double* X, int64_t ldx, int64_t sX, double a,
            const double* Z, int64_t ldz, int64_t sZ, double c, double d, int n, int m, int batch,
            hipStream_t s);

// Y[b] = M0 + shift[b] * I   (shared n x n source M0 with leading dim ldm).
void shifted_copy(double* Y, int64_t ldy, int64_t sY, const double* M0, int64_t ldm,
                  const double* shift, int n, int batch, hipStream_t s);

// Transposed copy: Y[b](j, i) = a * X[b](i, j) for an r x c block X.
void transpose_scaled(double* Y, int64_t ldy, int64_t sY, const double* X, int64_t ldx,
                      int64_t sX, double a, int r, int c, int batch, hipStream_t s);

// Column scale: Y(i, j) = X(i, j) * d[j]  (r x c, single matrix).
void col_scale(double* Y, int64_t ldy, const double* X, int64_t ldx, const double* d, int r, int c,
               hipStream_t s);

// Fill n entries with a value (graph-capturable replacement for memset).
void fill(double* p, size_t n, double v, hipStream_t s);
void fill_int(int* p, size_t n, int v, hipStream_t s);

// max_{b,i,j} |X[b](i,j) - X[b](j,i)| into *out (device double), n x n blocks.
void max_asym(const double* X, int64_t ld, int64_t sX, int n, int batch, double* out, hipStream_t s);

// Per-launch timing of k_dgemm and the Gauss-Jordan leaf kernel with HIP
// events on the launch stream, for every `every`-th launch of each class.
// prof_end waits for the recorded events and writes out[16] = five triples
// {ms, flop, launches}: all GEMMs, leaves, GEMMs with K < 64, 64 <= K < 256,
// K >= 256 (sampling counts per bucket); out[15] = the event-pair cost of an
// empty launch on the sampled stream (median of 32), already subtracted from
// every sampled duration.
void prof_begin(int every);
void prof_end(double* out);

}  // namespace sehf
