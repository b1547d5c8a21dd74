// Split-BF16 exchange kernel of the sigma-vector, gfx950
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation).
//
//   T[b][i][nu] = sum_mu X_o[mu][i] W[mu][nu] sum_j X_o[mu][j] zXv[b][j][nu]
//
// Every FP32 operand x is carried as hi = bf16(x), lo = bf16(x - hi) and every
// product as hi*hi + hi*lo + lo*hi (the lo*lo term, ~2^-32 relative, is
// dropped): three BF16 MFMAs per FP32 product.
//
// v_mfma_f32_16x16x32_bf16 lane layout (measured on gfx950): lane l supplies
// A[l%16][8(l/16)+s] and B[8(l/16)+s][l%16] in slot s = 0..7 of its bf16x8
// operands; the 16x16 FP32 result leaves D[4(l/16)+r][l%16] in element r.
//
// Per wave: 16 grid columns nu.  Per 32 grid rows mu (two 16-row P tiles):
//   P_t (16 mu x 16 nu) = X_o[mu, :] zXv[b][:, nu]          K = NP, 3 MFMAs/32
//   S = P o W                                               FP32, in registers
//   T[:, nu] += X_o[mu, :]^T S                              K = 32 mu, 3 MFMAs
// The two P tiles' result elements (rows 4q+r and 16+4q+r of lane q = l/16)
// are the eight K slots of lane l in the second product: slot s holds row
// pi(q, s) = 4q + s (s < 4) or 16 + 4q + (s - 4) (s >= 4), and the X_o^T
// operand is read in the same order, so the product needs no shuffle.
#pragma once

#include "common.h"

namespace mbpt {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mma_bf16(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// acc += (ah + al)(bh + bl) without the al*bl term
__device__ __forceinline__ f32x4 mma3(bf16x8 ah, bf16x8 al, bf16x8 bh, bf16x8 bl,
                                      f32x4 c) {
  c = mma_bf16(al, bh, c);
  c = mma_bf16(ah, bl, c);
  return mma_bf16(ah, bh, c);
}

__device__ __forceinline__ void split8(const float (&x)[8], bf16x8 &hi, bf16x8 &lo) {
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const __bf16 h = (__bf16)x[s];
    hi[s] = h;
    lo[s] = (__bf16)(x[s] - (float)h);
  }
}

// X_o split once per handle: A-side [mu][NP] and transposed [NP][M] copies
__global__ void k_split_bf16_xo(int M, int no, int np, const float *__restrict__ Xo,
                                __bf16 *__restrict__ xh, __bf16 *__restrict__ xl,
                                __bf16 *__restrict__ xth, __bf16 *__restrict__ xtl) {
  const long n = (long)M * np;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x) {
    const int mu = (int)(e / np), j = (int)(e % np);
    const float x = j < no ? Xo[(long)mu * no + j] : 0.f;
    const __bf16 h = (__bf16)x;
    const __bf16 l = (__bf16)(x - (float)h);
    xh[e] = h;
    xl[e] = l;
    xth[(long)j * M + mu] = h;
    xtl[(long)j * M + mu] = l;
  }
}

template <int NP>
__global__ __launch_bounds__(256) void k_exchange_bf16(
    int M, int no, int k, const __bf16 *__restrict__ xh, const __bf16 *__restrict__ xl,
    const __bf16 *__restrict__ xth, const __bf16 *__restrict__ xtl,
    const float *__restrict__ W, const float *__restrict__ zXv,
    float *__restrict__ Tpart, int nsplit, int rows_per_split, int nvb, int nblk) {
  const int lid = xcd_logical(blockIdx.x, nblk);
  if (lid >= nvb * k * nsplit) return;
  const int b = lid % k, t = lid / k, vb = t % nvb, s = t / nvb;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lq = lane >> 4;
  const int col = vb * 64 + wave * 16 + lr;
  const bool colok = col < M;
  constexpr int NKB = NP / 32, NI = NP / 16, ROWS = 64;
  constexpr int LDA = NP + 8;     // bf16 per staged X_o row (16 B pad)
  constexpr int LDT = ROWS + 4;   // bf16 per staged X_o^T row (8 B pad)
  __shared__ __bf16 sah[ROWS * LDA], sal[ROWS * LDA];
  __shared__ __bf16 sth[NP * LDT], stl[NP * LDT];

  // zXv[b][:, col] as the B operand of the P products, split once
  bf16x8 bzh[NKB], bzl[NKB];
#pragma unroll
  for (int kb = 0; kb < NKB; ++kb) {
    float v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int j = 32 * kb + 8 * lq + q;
      v[q] = (j < no && colok) ? zXv[((long)b * no + j) * M + col] : 0.f;
    }
    split8(v, bzh[kb], bzl[kb]);
  }
  f32x4 acc[NI];
#pragma unroll
  for (int it = 0; it < NI; ++it) zero(acc[it]);

  const int mu_beg = s * rows_per_split;
  const int mu_end = min(M, mu_beg + rows_per_split);
  for (int mg = mu_beg; mg < mu_end; mg += ROWS) {
    __syncthreads();
    // stage 64 rows of X_o (hi, lo) in both layouts; rows past mu_end are 0
    for (int e = tid; e < ROWS * NP; e += 256) {
      const int r = e / NP, j = e % NP, mu = mg + r;
      const bool ok = mu < mu_end;
      sah[r * LDA + j] = ok ? xh[(long)mu * NP + j] : (__bf16)0.f;
      sal[r * LDA + j] = ok ? xl[(long)mu * NP + j] : (__bf16)0.f;
    }
    for (int e = tid; e < NP * ROWS; e += 256) {
      const int j = e / ROWS, r = e % ROWS, mu = mg + r;
      const bool ok = mu < mu_end;
      sth[j * LDT + r] = ok ? xth[(long)j * M + mu] : (__bf16)0.f;
      stl[j * LDT + r] = ok ? xtl[(long)j * M + mu] : (__bf16)0.f;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < ROWS / 32; ++c) {
      const int r0 = 32 * c;
      if (mg + r0 >= mu_end) break;
      float sv[8];
#pragma unroll
      for (int tt = 0; tt < 2; ++tt) {
        f32x4 P;
        zero(P);
        const int row = r0 + 16 * tt + lr;
#pragma unroll
        for (int kb = 0; kb < NKB; ++kb) {
          const bf16x8 ah = *(const bf16x8 *)&sah[row * LDA + 32 * kb + 8 * lq];
          const bf16x8 al = *(const bf16x8 *)&sal[row * LDA + 32 * kb + 8 * lq];
          P = mma3(ah, al, bzh[kb], bzl[kb], P);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int mu = mg + r0 + 16 * tt + 4 * lq + r;
          const float w = (mu < mu_end && colok) ? W[(long)mu * M + col] : 0.f;
          sv[4 * tt + r] = P[r] * w;
        }
      }
      bf16x8 sh, sl;
      split8(sv, sh, sl);
#pragma unroll
      for (int it = 0; it < NI; ++it) {
        const int row = (16 * it + lr) * LDT + r0 + 4 * lq;
        bf16x8 ah, al;
        const bf16x4 h0 = *(const bf16x4 *)&sth[row];
        const bf16x4 h1 = *(const bf16x4 *)&sth[row + 16];
        const bf16x4 l0 = *(const bf16x4 *)&stl[row];
        const bf16x4 l1 = *(const bf16x4 *)&stl[row + 16];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          ah[q] = h0[q]; ah[4 + q] = h1[q];
          al[q] = l0[q]; al[4 + q] = l1[q];
        }
        acc[it] = mma3(ah, al, sh, sl, acc[it]);
      }
    }
  }
  if (colok) {
#pragma unroll
    for (int it = 0; it < NI; ++it)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        Tpart[(((long)s * k + b) * NP + 16 * it + 4 * lq + r) * M + col] = acc[it][r];
  }
}

}  // namespace mbpt
