// Batched FP64 symmetric eigensolver for the small Rayleigh-Ritz matrices of
// a batch of Davidson runs (m <= 64), one workgroup per matrix.
//
// Cyclic Jacobi with the round-robin (circle) ordering: a sweep is mp - 1
// rounds of mp / 2 disjoint rotations (mp = m rounded up to even; the pad row
// and column are zero and never rotate).  Each round computes its rotations
// from the 2x2 blocks, then applies all of them to the rows, then to the
// columns of A and of the eigenvector matrix V, both held in LDS.  Sweeps
// stop when every off-diagonal entry is below 1e-15 max(sqrt(|a_pp a_qq|),
// scale): relative to its own diagonal pair, so that decoupled blocks of very
// different size (such as slots shifted out of the way by the caller) do not
// loosen each other's tolerance, and absolute below the caller's scale, so
// that near-zero eigenvalues do not demand more than working precision (Golub & Van Loan, "Matrix Computations", sections 8.5.2-8.5.5:
// the 2x2 symmetric Schur decomposition and the parallel ordering; Demmel &
// Veselic, SIAM J. Matrix Anal. Appl. 13 (1992) 1204, for the criterion).  The input is
// symmetrised on load, (A + A^T) / 2, so the Rayleigh matrix of a
// lower-precision operator, which is not exactly symmetric, is accepted as is.
// Output: eigenvalues ascending and the matching eigenvectors as columns.
#include "common.h"

#include <algorithm>

namespace mbpt {
namespace {

constexpr int kMaxM = 64;
constexpr int kLD = kMaxM + 1;     // odd row stride: column walks hit distinct banks
constexpr size_t kSyevSmem = 2 * sizeof(double) * kMaxM * kLD;

__device__ __forceinline__ double block_max(double x, double *red) {
  for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_down(x, o, 64));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = x;
  __syncthreads();
  double s = 0;
  for (int q = 0; q < (int)(blockDim.x >> 6); ++q) s = fmax(s, red[q]);
  return s;
}

__global__ __launch_bounds__(256) void k_syev_jacobi(int m, const double *__restrict__ A,
                                                     int lda, long sA, double *__restrict__ w,
                                                     long sw, double *__restrict__ V, int ldv,
                                                     long sV, int max_sweeps, double scale2) {
  extern __shared__ double smem[];
  double *a = smem, *v = smem + kMaxM * kLD;
  __shared__ double cc[kMaxM / 2], ss[kMaxM / 2], red[4], ev[kMaxM];
  __shared__ int pp[kMaxM / 2], pq[kMaxM / 2], rank[kMaxM];
  const int tid = threadIdx.x, b = blockIdx.x;
  const int mp = (m + 1) & ~1, half = mp / 2;
  const double *Ab = A + b * sA;
  for (int e = tid; e < mp * mp; e += 256) {
    const int i = e / mp, j = e % mp;
    const double x = (i < m && j < m) ? 0.5 * (Ab[(long)i * lda + j] + Ab[(long)j * lda + i]) : 0.0;
    a[i * kLD + j] = x;
    v[i * kLD + j] = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int sweep = 0; sweep < max_sweeps && mp > 1; ++sweep) {
    double off = 0;   // max over i != j of a_ij^2 / |a_ii a_jj|
    for (int e = tid; e < mp * mp; e += 256) {
      const int i = e / mp, j = e % mp;
      const double x = a[i * kLD + j];
      if (i != j && x != 0.0)
        off = fmax(off, x * x / fmax(fabs(a[i * kLD + i] * a[j * kLD + j]), scale2));
    }
    off = block_max(off, red);
    if (!(off > 1e-30)) break;
    for (int r = 0; r < mp - 1; ++r) {
      if (tid < half) {
        int p, q;
        if (tid == 0) { p = r; q = mp - 1; }
        else { p = (r + tid) % (mp - 1); q = (r - tid + mp - 1) % (mp - 1); }
        const double apq = a[p * kLD + q];
        double c = 1.0, s = 0.0;
        if (apq != 0.0) {
          const double tau = (a[q * kLD + q] - a[p * kLD + p]) / (2.0 * apq);
          const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
          c = 1.0 / sqrt(1.0 + t * t);
          s = t * c;
        }
        pp[tid] = p; pq[tid] = q; cc[tid] = c; ss[tid] = s;
      }
      __syncthreads();
      for (int e = tid; e < half * mp; e += 256) {   // rows: A <- J^T A
        const int k = e / mp, j = e % mp, p = pp[k], q = pq[k];
        const double c = cc[k], s = ss[k];
        const double x = a[p * kLD + j], y = a[q * kLD + j];
        a[p * kLD + j] = c * x - s * y;
        a[q * kLD + j] = s * x + c * y;
      }
      __syncthreads();
      for (int e = tid; e < half * mp; e += 256) {   // columns: A <- A J, V <- V J
        const int k = e / mp, i = e % mp, p = pp[k], q = pq[k];
        const double c = cc[k], s = ss[k];
        double x = a[i * kLD + p], y = a[i * kLD + q];
        a[i * kLD + p] = c * x - s * y;
        a[i * kLD + q] = s * x + c * y;
        x = v[i * kLD + p]; y = v[i * kLD + q];
        v[i * kLD + p] = c * x - s * y;
        v[i * kLD + q] = s * x + c * y;
      }
      __syncthreads();
    }
  }
  // ascending order by rank (ties broken by index); the pad column is last
  if (tid < mp) ev[tid] = a[tid * kLD + tid];
  __syncthreads();
  if (tid < m) {
    const double x = ev[tid];
    int rk = 0;
    for (int j = 0; j < m; ++j) rk += (ev[j] < x) || (ev[j] == x && j < tid);
    rank[tid] = rk;
    w[b * sw + rk] = x;
  }
  __syncthreads();
  double *Vb = V + b * sV;
  for (int e = tid; e < m * m; e += 256) {
    const int i = e / m, j = e % m;
    Vb[(long)i * ldv + rank[j]] = v[i * kLD + j];
  }
}

}  // namespace
}  // namespace mbpt

using namespace mbpt;

extern "C" {

int mbpt_syev_batched_f64(int nbatch, int m, const double *A_dev, int lda, long strideA,
                          double *w_dev, long strideW, double *V_dev, int ldv, long strideV,
                          int max_sweeps, double scale, void *stream) {
  MBPT_REQUIRE(nbatch >= 0 && m >= 1 && m <= kMaxM,
               "mbpt_syev_batched_f64: need 1 <= m <= 64 and nbatch >= 0");
  MBPT_REQUIRE(lda >= m && ldv >= m && strideW >= m && max_sweeps >= 1,
               "mbpt_syev_batched_f64: bad leading dimension or stride");
  MBPT_REQUIRE(A_dev && w_dev && V_dev, "mbpt_syev_batched_f64: NULL pointer");
  MBPT_REQUIRE(scale >= 0.0, "mbpt_syev_batched_f64: scale must be non-negative");
  if (nbatch == 0) return MBPT_OK;
  static bool attr_set = false;
  if (!attr_set) {
    MBPT_HIP_TRY(hipFuncSetAttribute((const void *)k_syev_jacobi,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)kSyevSmem));
    attr_set = true;
  }
  k_syev_jacobi<<<nbatch, 256, kSyevSmem, (hipStream_t)stream>>>(
      m, A_dev, lda, strideA, w_dev, strideW, V_dev, ldv, strideV, max_sweeps,
      std::max(scale * scale, 1e-300));
  return launch_status("k_syev_jacobi");
}

}  // extern "C"
