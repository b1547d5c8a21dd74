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

// NN on v_mfma_f64_16x16x4f64: 64 rows per block, 16 per wave, the output's
// columns in groups of up to four 16-wide tiles.  Per 4-deep K step a lane
// supplies A[row lr][k lq] (its row's 4 consecutive values) and S[k lq][col
// lr] of each tile; S is shared by every block of the instance, so its reads
// come from the caches.
__global__ __launch_bounds__(256) void k_bnn(int P, int m, int c, const int *__restrict__ rows,
                                             const double *__restrict__ A, long sA, int lda,
                                             const double *__restrict__ S, long sS, int lds,
                                             double *__restrict__ Y, long sY, int ldy,
                                             int subtract) {
  const int b = blockIdx.y, r0 = blockIdx.x * 64, tid = threadIdx.x;
  const int nr = rows[b];
  double *Yb = Y + b * sY;
  if (r0 >= nr) {   // a block of padding only
    if (!subtract)
      for (int e = tid; e < 64 * c; e += 256) {
        const int p = r0 + e / c;
        if (p < P) Yb[(long)p * ldy + e % c] = 0.0;
      }
    return;
  }
  const double *Ab = A + b * sA, *Sb = S + b * sS;
  const int lane = tid & 63, wave = tid >> 6, lr = lane & 15, lq = lane >> 4;
  const int ra = r0 + 16 * wave + lr;          // the row this lane feeds to A
  const bool rok = ra < nr;
  for (int c0 = 0; c0 < c; c0 += 64) {
    const int nct = min(4, (c - c0 + 15) / 16);
    f64x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) zero(acc[t]);
    for (int k = 0; k < m; k += 4) {
      const int ka = k + lq;
      const double a = (rok && ka < m) ? Ab[(long)ra * lda + ka] : 0.0;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        if (t >= nct) break;   // uniform
        const int col = c0 + 16 * t + lr;
        const double sv = (ka < m && col < c) ? Sb[(long)ka * lds + col] : 0.0;
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, sv, acc[t], 0, 0, 0);
      }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (t >= nct) break;
      const int col = c0 + 16 * t + lr;
      if (col >= c) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int p = r0 + 16 * wave + MF<double>::orow(lq, r);
        if (p >= P) continue;
        double *y = Yb + (long)p * ldy + col;
        if (p < nr) *y = subtract ? *y - acc[t][r] : acc[t][r];
        else if (!subtract) *y = 0.0;
      }
    }
  }
}

constexpr int kTnChunk = 256;   // pair rows per partial product

// TN partial over one 256-row chunk of one instance on v_mfma_f64_16x16x4f64:
// the output's 16x16 tiles (a, c <= 64: at most 4 x 4) are dealt round-robin
// to the 4 waves; per 4 pair rows a lane loads its A element (row k = l/16,
// output row l%16 of the tile) and its B element straight from global memory
// (the 4 x 16 patches are 128-byte row pieces), no LDS.
__global__ __launch_bounds__(256) void k_btn_partial(int a, int c, int nch,
                                                     const int *__restrict__ rows,
                                                     const double *__restrict__ A, long sA,
                                                     int lda, const double *__restrict__ Bm,
                                                     long sB, int ldb, double *__restrict__ W) {
  const int b = blockIdx.y, q = blockIdx.x;
  const int nr = rows[b], p0 = q * kTnChunk;
  if (p0 >= nr) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lq = lane >> 4;
  const int nti = (a + 15) / 16, ntj = (c + 15) / 16, nt = nti * ntj;
  const double *Ab = A + b * sA, *Bb = Bm + b * sB;
  const int pend = min(nr, p0 + kTnChunk);
  f64x4 acc[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) zero(acc[u]);
  for (int p = p0; p < pend; p += 4) {
    const int pr = p + lq;
    const bool rok = pr < pend;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int t = wave + 4 * u;
      if (t >= nt) break;
      const int i = 16 * (t / ntj) + lr, j = 16 * (t % ntj) + lr;
      const double x = (rok && i < a) ? Ab[(long)pr * lda + i] : 0.0;
      const double y = (rok && j < c) ? Bb[(long)pr * ldb + j] : 0.0;
      acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, acc[u], 0, 0, 0);
    }
  }
  double *out = W + ((long)b * nch + q) * a * c;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int t = wave + 4 * u;
    if (t >= nt) break;
    const int j = 16 * (t % ntj) + lr;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = 16 * (t / ntj) + MF<double>::orow(lq, r);
      if (i < a && j < c) out[(long)i * c + j] = acc[u][r];
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
  MBPT_REQUIRE(m <= 64, "mbpt_bnn_f64: m > 64 is not supported");
  MBPT_REQUIRE(rows_dev && A && S && Y && lda >= m && lds >= c && ldy >= c,
               "mbpt_bnn_f64: bad pointer or leading dimension");
  dim3 grid((P + 63) / 64, B);
  k_bnn<<<grid, 256, 0, (hipStream_t)stream>>>(P, m, c, rows_dev, A, sA, lda, S, sS, lds, Y,
                                               sY, ldy, subtract);
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
  MBPT_REQUIRE(a <= 64 && c <= 64, "mbpt_btn_f64: a and c up to 64");
  MBPT_REQUIRE(rows_dev && A && Bm && C && work && lda >= a && ldb >= c && ldc >= c,
               "mbpt_btn_f64: bad pointer or leading dimension");
  const int nch = std::max(1, (P + kTnChunk - 1) / kTnChunk);
  MBPT_REQUIRE(work_elems >= (long)B * nch * a * c, "mbpt_btn_f64: work buffer too small");
  hipStream_t st = (hipStream_t)stream;
  k_btn_partial<<<dim3(nch, B), 256, 0, st>>>(a, c, nch, rows_dev, A, sA, lda, Bm, sB, ldb,
                                              work);
  int rc = launch_status("k_btn_partial");
  if (rc) return rc;
  const long n = (long)B * a * c;
  k_btn_sum<<<(int)std::min<long>((n + 255) / 256, 8192L), 256, 0, st>>>(B, a, c, nch, rows_dev,
                                                                        work, C, sC, ldc);
  return launch_status("k_btn_sum");
}

}  // extern "C"
