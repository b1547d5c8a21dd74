// FP32 MFMA GEMM on K-major operands, gfx950 (v_mfma_f32_16x16x4f32):
//
//   C[m][n] = sum_k A[k][m] B[k][n],  A (K x lda), B (K x ldb) row-major
//
// for the skinny products of the sigma-vector (head z X_v^T and tail T' X_v,
// K = n_vir or M).  The operands are stored K-major and zero-padded (K to a
// multiple of 16, the m and n extents to a multiple of 128), so the main loop
// has no bounds checks and every global load is a coalesced float4: the
// same 128 x 128 tile scheme as the chi0 kernel (8 waves in 2 x 4, 64 x 32
// each, 16-deep K stages in double-buffered LDS with a 144-float row stride).
// The K range can be split into chunks (blockIdx -> (tile m, tile n, chunk));
// the epilogue functor receives (chunk, row, col, value).
#pragma once

#include "common.h"

namespace mbpt {

constexpr int kKmT = 128;     // tile edge
constexpr int kKmBK = 16;     // K per stage
constexpr int kKmLD = 144;    // LDS row stride (floats)

__device__ __forceinline__ void km_gemm(const float *__restrict__ A, int lda,
                                        const float *__restrict__ B, int ldb, int m0,
                                        int n0, int K, f32x4 (&acc)[4][2],
                                        float (*sa)[kKmBK][kKmLD],
                                        float (*sb)[kKmBK][kKmLD]) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 2, wn = wave & 3;
  const int lr = lane & 15, lq = lane >> 4;
  const int srow = tid >> 5, scol = (tid & 31) * 4;
  const float *ga = A + (long)srow * lda + m0 + scol;
  const float *gb = B + (long)srow * ldb + n0 + scol;
  if (K <= 0) return;
  float4 ra = *(const float4 *)ga, rb = *(const float4 *)gb;
  int buf = 0;
  *(float4 *)&sa[0][srow][scol] = ra;
  *(float4 *)&sb[0][srow][scol] = rb;
  __syncthreads();
  for (int k0 = 0; k0 < K; k0 += kKmBK) {
    const bool more = k0 + kKmBK < K;
    if (more) {
      ra = *(const float4 *)(ga + (long)(k0 + kKmBK) * lda);
      rb = *(const float4 *)(gb + (long)(k0 + kKmBK) * ldb);
    }
#pragma unroll
    for (int kk = 0; kk < kKmBK / 4; ++kk) {
      float a[4], b[2];
#pragma unroll
      for (int fm = 0; fm < 4; ++fm) a[fm] = sa[buf][4 * kk + lq][64 * wm + 16 * fm + lr];
#pragma unroll
      for (int fn = 0; fn < 2; ++fn) b[fn] = sb[buf][4 * kk + lq][32 * wn + 16 * fn + lr];
#pragma unroll
      for (int fm = 0; fm < 4; ++fm)
#pragma unroll
        for (int fn = 0; fn < 2; ++fn)
          acc[fm][fn] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[fm], b[fn], acc[fm][fn], 0, 0, 0);
    }
    if (more) {
      *(float4 *)&sa[buf ^ 1][srow][scol] = ra;
      *(float4 *)&sb[buf ^ 1][srow][scol] = rb;
    }
    __syncthreads();
    buf ^= 1;
  }
}

// grid: ntm x ntn tiles x nq K-chunks of kchunk (a multiple of 16) over K
template <class Epi>
__global__ __launch_bounds__(512) void k_gemm_km(int K, int kchunk, const float *__restrict__ A,
                                                 int lda, const float *__restrict__ B, int ldb,
                                                 int ntm, int ntn, int nq, Epi epi, int nblk) {
  const int lid = xcd_logical(blockIdx.x, nblk);
  if (lid >= ntm * ntn * nq) return;
  const int tn = lid % ntn, tm = (lid / ntn) % ntm, q = lid / (ntn * ntm);
  const int k0 = q * kchunk, kc = min(kchunk, K - k0);
  __shared__ __attribute__((aligned(16))) float sa[2][kKmBK][kKmLD];
  __shared__ __attribute__((aligned(16))) float sb[2][kKmBK][kKmLD];
  f32x4 acc[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) zero(acc[i][j]);
  km_gemm(A + (long)k0 * lda, lda, B + (long)k0 * ldb, ldb, tm * kKmT, tn * kKmT, kc, acc,
          sa, sb);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 2, wn = wave & 3, lr = lane & 15, lq = lane >> 4;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        epi(q, tm * kKmT + 64 * wm + 16 * i + 4 * lq + r, tn * kKmT + 32 * wn + 16 * j + lr,
            acc[i][j][r]);
}

// C row-major (ldc) for row < rows and col < cols, chunk q at C + q rows ldc
struct EpiStore {
  float *C;
  int rows, cols, ldc;
  __device__ void operator()(int q, int row, int col, float v) const {
    if (row < rows && col < cols) C[((long)q * rows + row) * ldc + col] = v;
  }
};

// out[a][m] = in[m][a] for the padded K-major copy (m < rows, a < cols read; the
// padding up to pr x pc written as zero), 64 x 64 tiles through LDS
template <typename TI>
__global__ __launch_bounds__(256) void k_transpose_pad(int rows, int cols, const TI *__restrict__ in,
                                                       int ldi, float *__restrict__ out, int ldo,
                                                       int pr, int pc) {
  __shared__ float t[64][65];
  const int m0 = blockIdx.x * 64, a0 = blockIdx.y * 64;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int r = e >> 6, c = e & 63, m = m0 + r, a = a0 + c;
    t[r][c] = (m < rows && a < cols) ? (float)in[(long)m * ldi + a] : 0.f;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int r = e >> 6, c = e & 63, a = a0 + r, m = m0 + c;
    if (a < pc && m < pr) out[(long)a * ldo + m] = t[c][r];
  }
}

}  // namespace mbpt
