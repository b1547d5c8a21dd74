// Split-BF16 GEMM with both operands K-contiguous, gfx950
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation):
//
//   C[m][n] = sum_k A[m][k] B[n][k]
//
// for the head (z X_v^T: A = z, B = X_v, K = n_vir) and tail (T' X_v: A = T',
// B = X_v^T, K = M) of the split-BF16 sigma-vector.  Each operand is held as
// two BF16 arrays, hi = bf16(x) and lo = bf16(x - hi), row-major with K
// contiguous and zero padding (rows to 128, K to 32); every product is
// hi*hi + hi*lo + lo*hi.  With K contiguous, the 8 consecutive K values a
// lane supplies to the MFMA (A[l%16][8(l/16)+s], B[8(l/16)+s][l%16]) are one
// 16-byte piece of a row, so the operands need no transpose.
//
// Workgroup tile 128 x 128, 4 waves in 2 x 2 (64 x 64 each, 4 x 4 MFMA tiles).
// K advances 32 at a time through double-buffered LDS filled by direct
// global-to-LDS 16-byte loads (4 lanes per 64-byte row piece); the 16-byte
// pieces of a row are XOR-swizzled by (row / 4) % 4 so that the 16 rows a
// fragment read touches fall in distinct banks.  K may be split in chunks
// (blockIdx -> (tile m, tile n, chunk)); the epilogue functor receives
// (chunk, row, col, value).
#pragma once

#include "common.h"

namespace mbpt {

typedef __bf16 nt3_bf16x8 __attribute__((ext_vector_type(8)));
typedef __attribute__((address_space(3))) void nt3_lds;

constexpr int kNt3T = 128;   // tile edge
constexpr int kNt3K = 32;    // K per stage

__device__ __forceinline__ int nt3_swz(int row) { return (row >> 2) & 3; }

// grid: ntm x ntn tiles x nq K chunks of kchunk (a multiple of 32) over K
template <class Epi>
__global__ __launch_bounds__(256) void k_gemm_nt3(int K, int kchunk, const __bf16 *__restrict__ Ah,
                                                  const __bf16 *__restrict__ Al, int lda,
                                                  const __bf16 *__restrict__ Bh,
                                                  const __bf16 *__restrict__ Bl, int ldb, int ntm,
                                                  int ntn, int nq, Epi epi, int nblk) {
  // [buffer][array: A hi, A lo, B hi, B lo][128 rows x 32 k]
  __shared__ __attribute__((aligned(16))) __bf16 sm[2][4][kNt3T * kNt3K];
  const int lid = xcd_logical(blockIdx.x, nblk);
  if (lid >= ntm * ntn * nq) return;
  const int tn = lid % ntn, tm = (lid / ntn) % ntm, q = lid / (ntn * ntm);
  const int k0 = q * kchunk, kend = min(K, k0 + kchunk);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave >> 1, wn = wave & 1, lr = lane & 15, lq = lane >> 4;
  const long m0 = (long)tm * kNt3T, n0 = (long)tn * kNt3T;

  // stage fill: 32 instructions of 1 KB (16 rows x 64 B), 8 per wave; lane L
  // of instruction (array, 16-row block) writes LDS piece L = (row L/4, slot
  // L%4) with global piece (L%4) ^ swz(row)
  auto fetch = [&](int kk, int buf) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int j = wave * 8 + i, arr = j >> 3, rb = j & 7;
      const int row = rb * 16 + (lane >> 2);
      const int piece = (lane & 3) ^ nt3_swz(row);
      const __bf16 *src = arr == 0 ? Ah : arr == 1 ? Al : arr == 2 ? Bh : Bl;
      const long ld = arr < 2 ? lda : ldb;
      const long grow = (arr < 2 ? m0 : n0) + row;
      __builtin_amdgcn_global_load_lds((const void *)(src + grow * ld + kk + 8 * piece),
                                       (nt3_lds *)&sm[buf][arr][rb * 16 * kNt3K], 16, 0, 0);
    }
  };
  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) zero(acc[i][j]);
  if (k0 < kend) {
    fetch(k0, 0);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    int cur = 0;
    for (int kk = k0; kk < kend; kk += kNt3K) {
      if (kk + kNt3K < kend) fetch(kk + kNt3K, cur ^ 1);   // buffer read before the last barrier
      nt3_bf16x8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int ra = 64 * wm + 16 * i + lr, rb = 64 * wn + 16 * i + lr;
        const int pa = (lq ^ nt3_swz(ra)) * 8, pb = (lq ^ nt3_swz(rb)) * 8;
        ah[i] = *(const nt3_bf16x8 *)&sm[cur][0][ra * kNt3K + pa];
        al[i] = *(const nt3_bf16x8 *)&sm[cur][1][ra * kNt3K + pa];
        bh[i] = *(const nt3_bf16x8 *)&sm[cur][2][rb * kNt3K + pb];
        bl[i] = *(const nt3_bf16x8 *)&sm[cur][3][rb * kNt3K + pb];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        }
      __builtin_amdgcn_s_waitcnt(0);   // this wave's stage loads landed
      __syncthreads();                 // ... and every other wave's
      cur ^= 1;
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        epi(q, (int)(m0 + 64 * wm + 16 * i + 4 * lq + r), (int)(n0 + 64 * wn + 16 * j + lr),
            acc[i][j][r]);
}

// hi / lo BF16 rows of a (rows x cols) matrix, zero-padded to (pr x pc):
// out[r][c] for c < pc, r < pr; transposed (TR) reads in[c][r]
template <typename TI, bool TR>
__global__ void k_split_rows_bf16(int rows, int cols, const TI *__restrict__ in, int ldi, int pr,
                                  int pc, __bf16 *__restrict__ hi, __bf16 *__restrict__ lo) {
  const long n = (long)pr * pc;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x) {
    const int r = (int)(e / pc), c = (int)(e % pc);
    float x = 0.f;
    if (r < rows && c < cols) x = (float)(TR ? in[(long)c * ldi + r] : in[(long)r * ldi + c]);
    const __bf16 h = (__bf16)x;
    hi[e] = h;
    lo[e] = (__bf16)(x - (float)h);
  }
}

}  // namespace mbpt
