// FP64 products of the batched Davidson subspace on gfx950, per instance b of
// a batch whose vectors are padded to P pairs but hold only rows[b] real ones:
//
//   NN  Y_b (P x c)  = A_b (P x m) S_b (m x c)      [or Y_b -= A_b S_b]
//   TN  C_b (a x c)  = A_b^T (a x P) B_b (P x c)
//
// Rows at and past rows[b] are neither read nor (TN, NN subtract) written;
// NN assignment writes zeros there so the padding stays zero.  The padded
// reads are what made a library batched GEMM move ~3x the bytes for the
// uneven batch (pair counts 100..7920).
#include "common.h"

#include <algorithm>

namespace mbpt {
namespace {

constexpr int kNnCols = 16;   // output columns per pass (registers)

// one thread per row p, S_b staged in LDS, kNnCols accumulators per pass
__global__ __launch_bounds__(256) void k_bnn(int P, int m, int c, const int *__restrict__ rows,
                                             const double *__restrict__ A, long sA, int lda,
                                             const double *__restrict__ S, long sS, int lds,
                                             double *__restrict__ Y, long sY, int ldy,
                                             int subtract) {
  extern __shared__ double ss[];   // m x c
  const int b = blockIdx.y, p = blockIdx.x * 256 + threadIdx.x;
  const int nr = rows[b];
  const double *Ab = A + b * sA, *Sb = S + b * sS;
  double *Yb = Y + b * sY;
  if ((int)(blockIdx.x * 256) >= nr) {   // a block of padding only
    if (!subtract && p < P)
      for (int j = 0; j < c; ++j) Yb[(long)p * ldy + j] = 0.0;
    return;
  }
  for (int e = threadIdx.x; e < m * c; e += 256) ss[e] = Sb[(long)(e / c) * lds + e % c];
  __syncthreads();
  if (p >= P) return;
  if (p >= nr) {
    if (!subtract)
      for (int j = 0; j < c; ++j) Yb[(long)p * ldy + j] = 0.0;
    return;
  }
  const double *ar = Ab + (long)p * lda;
  for (int j0 = 0; j0 < c; j0 += kNnCols) {
    double acc[kNnCols];
#pragma unroll
    for (int j = 0; j < kNnCols; ++j) acc[j] = 0.0;
    const int nj = min(kNnCols, c - j0);
    for (int i = 0; i < m; ++i) {
      const double a = ar[i];
      const double *si = ss + i * c + j0;
#pragma unroll
      for (int j = 0; j < kNnCols; ++j)
        if (j < nj) acc[j] += a * si[j];
    }
    double *yr = Yb + (long)p * ldy + j0;
#pragma unroll
    for (int j = 0; j < kNnCols; ++j)
      if (j < nj) yr[j] = subtract ? yr[j] - acc[j] : acc[j];
  }
}

struct LdAtB {  // A(m = i, k = r) = A[r][i] for r < K
  const double *p;
  int ld, rows, K;
  __device__ double operator()(int m, int k) const {
    return (m < rows && k < K) ? p[(long)k * ld + m] : 0.0;
  }
};
struct LdBrB {  // B(k = r, n = j) = B[r][j]
  const double *p;
  int ld, K, cols;
  __device__ double operator()(int k, int n) const {
    return (k < K && n < cols) ? p[(long)k * ld + n] : 0.0;
  }
};

constexpr int kTnChunk = 512;   // pair rows per partial product

// grid (tiles x chunks, B): the chunk's partial product as a 64x64 MFMA tile
__global__ __launch_bounds__(256) void k_btn_partial(int a, int c, int nch,
                                                     const int *__restrict__ rows,
                                                     const double *__restrict__ A, long sA,
                                                     int lda, const double *__restrict__ Bm,
                                                     long sB, int ldb, double *__restrict__ W,
                                                     int nta, int ntb) {
  const int b = blockIdx.y;
  const int t = blockIdx.x % (nta * ntb), q = blockIdx.x / (nta * ntb);
  const int tb = t % ntb, ta = t / ntb;
  const int r0 = q * kTnChunk, rc = min(kTnChunk, rows[b] - r0);
  if (rc <= 0) return;
  __shared__ TileSmem<double> sm;
  MF<double>::acc acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) zero(acc[i][j]);
  gemm64<double, false, false>(LdAtB{A + b * sA + (long)r0 * lda, lda, a, rc},
                               LdBrB{Bm + b * sB + (long)r0 * ldb, ldb, rc, c}, ta * kTile,
                               tb * kTile, rc, acc, sm);
  double *out = W + ((long)b * nch + q) * a * c;
  for (int fm = 0; fm < 2; ++fm)
    for (int fn = 0; fn < 2; ++fn) {
      const int j = tile_col(tb * kTile, fn);
      for (int r = 0; r < 4; ++r) {
        const int i = tile_row<double>(ta * kTile, fm, r);
        if (i < a && j < c) out[(long)i * c + j] = acc[fm][fn][r];
      }
    }
}

__global__ void k_btn_sum(int B, int a, int c, int nch, const int *__restrict__ rows,
                          const double *__restrict__ W, double *__restrict__ C, long sC,
                          int ldc) {
  const long n = (long)B * a * c;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x) {
    const int b = (int)(e / ((long)a * c)), ij = (int)(e % ((long)a * c));
    const int nq = (rows[b] + kTnChunk - 1) / kTnChunk;
    double s = 0.0;
    for (int q = 0; q < nq; ++q) s += W[((long)b * nch + q) * a * c + ij];
    C[b * sC + (long)(ij / c) * ldc + ij % c] = s;
  }
}

}  // namespace
}  // namespace mbpt

using namespace mbpt;

extern "C" {

int mbpt_bnn_f64(int B, int P, int m, int c, const int *rows_dev, const double *A, long sA,
                 int lda, const double *S, long sS, int lds, double *Y, long sY, int ldy,
                 int subtract, void *stream) {
  MBPT_REQUIRE(B >= 0 && P >= 0 && m >= 0 && c >= 0, "mbpt_bnn_f64: negative size");
  if (B == 0 || P == 0 || c == 0) return MBPT_OK;
  MBPT_REQUIRE(rows_dev && A && S && Y && lda >= m && lds >= c && ldy >= c,
               "mbpt_bnn_f64: bad pointer or leading dimension");
  MBPT_REQUIRE((long)m * c * 8 <= 64 * 1024, "mbpt_bnn_f64: m c exceeds the 64 KiB LDS stage");
  static bool attr = false;
  if (!attr) {
    MBPT_HIP_TRY(hipFuncSetAttribute((const void *)k_bnn,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, 64 * 1024));
    attr = true;
  }
  dim3 grid((P + 255) / 256, B);
  k_bnn<<<grid, 256, sizeof(double) * (size_t)std::max(m * c, 1), (hipStream_t)stream>>>(
      P, m, c, rows_dev, A, sA, lda, S, sS, lds, Y, sY, ldy, subtract);
  return launch_status("k_bnn");
}

long mbpt_btn_f64_work(int B, int P, int a, int c) {
  if (B <= 0 || P <= 0 || a <= 0 || c <= 0) return 0;
  return (long)B * ((P + kTnChunk - 1) / kTnChunk) * a * c;
}

int mbpt_btn_f64(int B, int P, int a, int c, const int *rows_dev, const double *A, long sA,
                 int lda, const double *Bm, long sB, int ldb, double *C, long sC, int ldc,
                 double *work, long work_elems, void *stream) {
  MBPT_REQUIRE(B >= 0 && P >= 0 && a >= 0 && c >= 0, "mbpt_btn_f64: negative size");
  if (B == 0 || a == 0 || c == 0) return MBPT_OK;
  MBPT_REQUIRE(rows_dev && A && Bm && C && work && lda >= a && ldb >= c && ldc >= c,
               "mbpt_btn_f64: bad pointer or leading dimension");
  const int nch = std::max(1, (P + kTnChunk - 1) / kTnChunk);
  MBPT_REQUIRE(work_elems >= (long)B * nch * a * c, "mbpt_btn_f64: work buffer too small");
  hipStream_t st = (hipStream_t)stream;
  const int nta = (a + kTile - 1) / kTile, ntb = (c + kTile - 1) / kTile;
  dim3 grid(nta * ntb * nch, B);
  k_btn_partial<<<grid, 256, 0, st>>>(a, c, nch, rows_dev, A, sA, lda, Bm, sB, ldb, work, nta,
                                      ntb);
  int rc = launch_status("k_btn_partial");
  if (rc) return rc;
  const long n = (long)B * a * c;
  k_btn_sum<<<(int)std::min<long>((n + 255) / 256, 8192L), 256, 0, st>>>(B, a, c, nch, rows_dev,
                                                                        work, C, sC, ldc);
  return launch_status("k_btn_sum");
}

}  // extern "C"
