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
// Per 16 grid columns nu and 32 grid rows mu (two 16-row P tiles):
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

// ---- X_o in MFMA fragment order -------------------------------------------------
// The two products read X_o as A operands in two shapes: X_o[mu][j] for P
// (K = j) and X_o^T[i][mu] for T (K = mu, in the slot order pi above).  Both
// are packed once per handle, split into hi and lo, so that every operand a
// lane needs is one 16-byte slot and a wave's fragment is 1 KB contiguous.
// Per 32-row chunk of the grid, in 16-byte slots:
//   P part  [tt < 2][kb < NP/32][hl < 2][lane]  rows 16 tt + l%16, j = 32 kb + 8 (l/16) + s
//   T part  [it < NP/16][hl < 2][lane]          rows pi(l/16, s),  i = 16 it + l%16
// rows >= M and columns >= nocc are zero.
__host__ __device__ constexpr int bf16_chunk_slots(int np) { return 64 * (4 * (np / 32) + 2 * (np / 16)); }

__global__ void k_pack_bf16_xo(int M, int no, int np, const float *__restrict__ Xo,
                               bf16x8 *__restrict__ frag) {
  const int nkb = np / 32, spc = bf16_chunk_slots(np), toff = 64 * 4 * nkb;
  const long n = (long)((M + 31) / 32) * spc;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x) {
    const int chunk = (int)(e / spc), rem = (int)(e % spc);
    int lane, hl, rowbase, jbase, mode;
    if (rem < toff) {
      lane = rem & 63;
      const int g = rem >> 6;
      hl = g & 1;
      const int kb = (g >> 1) % nkb, tt = (g >> 1) / nkb;
      rowbase = 32 * chunk + 16 * tt + (lane & 15);
      jbase = 32 * kb + 8 * (lane >> 4);
      mode = 0;
    } else {
      const int r2 = rem - toff;
      lane = r2 & 63;
      const int g = r2 >> 6;
      hl = g & 1;
      rowbase = 32 * chunk;
      jbase = 16 * (g >> 1) + (lane & 15);
      mode = 1;
    }
    bf16x8 out;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int q = lane >> 4;
      const int row = mode == 0 ? rowbase : rowbase + (s < 4 ? 4 * q + s : 16 + 4 * q + (s - 4));
      const int j = mode == 0 ? jbase + s : jbase;
      const float x = (row < M && j < no) ? Xo[(long)row * no + j] : 0.f;
      const __bf16 h = (__bf16)x;
      out[s] = hl == 0 ? h : (__bf16)(x - (float)h);
    }
    frag[e] = out;
  }
}

// ---- the exchange ----------------------------------------------------------------
typedef __attribute__((address_space(3))) void lds_void;
// Workgroup = 4 waves x 32 grid columns (two 16-column tiles per wave, so each
// X_o fragment read from LDS feeds two MFMA chains), one vector b, one split
// of the grid rows.  The packed X_o chunks stream through a double-buffered
// LDS stage with direct global-to-LDS 16-byte loads (global_load_lds_dwordx4:
// no staging registers, so two waves fit per SIMD), and the next chunk's W
// values are loaded into registers while the current chunk computes.
template <int NP>
__global__ __launch_bounds__(256) void k_exchange_bf16(
    int M, int no, int k, const bf16x8 *__restrict__ frag, const float *__restrict__ W,
    const float *__restrict__ zXv, float *__restrict__ Tpart, int nsplit,
    int rows_per_split, int nvb, int nblk) {
  constexpr int NKB = NP / 32, NI = NP / 16, C = 2;
  constexpr int SPC = bf16_chunk_slots(NP), TOFF = 64 * 4 * NKB, LPT = SPC / 256;
  __shared__ bf16x8 buf[2][SPC];
  const int lid = xcd_logical(blockIdx.x, nblk);
  if (lid >= nvb * k * nsplit) return;
  const int b = lid % k, t = lid / k, vb = t % nvb, s = t / nvb;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lq = lane >> 4;
  const int cbase = vb * 128 + wave * 32;

  // zXv[b][:, col] as the B operand of the P products, split once
  bf16x8 bzh[C][NKB], bzl[C][NKB];
#pragma unroll
  for (int cc = 0; cc < C; ++cc) {
    const int col = cbase + 16 * cc + lr;
#pragma unroll
    for (int kb = 0; kb < NKB; ++kb) {
      float v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int j = 32 * kb + 8 * lq + q;
        v[q] = (j < no && col < M) ? zXv[((long)b * no + j) * M + col] : 0.f;
      }
      split8(v, bzh[cc][kb], bzl[cc][kb]);
    }
  }
  f32x4 acc[NI][C];
#pragma unroll
  for (int it = 0; it < NI; ++it)
#pragma unroll
    for (int cc = 0; cc < C; ++cc) zero(acc[it][cc]);

  const int c_beg = (s * rows_per_split) / 32;
  const int c_end = min((M + 31) / 32, ((s + 1) * rows_per_split) / 32);
  auto load_w = [&](int c, float (&w)[2][C][4]) {
#pragma unroll
    for (int tt = 0; tt < 2; ++tt)
#pragma unroll
      for (int cc = 0; cc < C; ++cc) {
        const int col = cbase + 16 * cc + lr;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int mu = 32 * c + 16 * tt + 4 * lq + r;
          w[tt][cc][r] = (mu < M && col < M) ? W[(long)mu * M + col] : 0.f;
        }
      }
  };
  // chunk c into buf[n]: slot 256 q + tid from thread tid, each wave's
  // instruction filling 1 KB of LDS at base + 16 lane
  auto fetch = [&](int c, int n) {
    const bf16x8 *src = frag + (long)c * SPC;
#pragma unroll
    for (int q = 0; q < LPT; ++q)
      __builtin_amdgcn_global_load_lds((const void *)(src + 256 * q + tid),
                                       (lds_void *)&buf[n][256 * q + 64 * wave], 16, 0, 0);
  };
  if (c_beg < c_end) fetch(c_beg, 0);
  float w[2][C][4];
  load_w(c_beg, w);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  int cur = 0;
  for (int c = c_beg; c < c_end; ++c) {
    const bool more = c + 1 < c_end;
    float wn[2][C][4];
    if (more) {
      fetch(c + 1, cur ^ 1);   // buf[cur ^ 1] was last read before the previous barrier
      load_w(c + 1, wn);
    }
    const bf16x8 *fb = buf[cur];
    f32x4 P[2][C];
#pragma unroll
    for (int tt = 0; tt < 2; ++tt)
#pragma unroll
      for (int cc = 0; cc < C; ++cc) zero(P[tt][cc]);
#pragma unroll
    for (int tt = 0; tt < 2; ++tt)
#pragma unroll
      for (int kb = 0; kb < NKB; ++kb) {
        const bf16x8 ah = fb[((tt * NKB + kb) * 2 + 0) * 64 + lane];
        const bf16x8 al = fb[((tt * NKB + kb) * 2 + 1) * 64 + lane];
#pragma unroll
        for (int cc = 0; cc < C; ++cc)
          P[tt][cc] = mma3(ah, al, bzh[cc][kb], bzl[cc][kb], P[tt][cc]);
      }
    bf16x8 sh[C], sl[C];
#pragma unroll
    for (int cc = 0; cc < C; ++cc) {
      float sv[8];
#pragma unroll
      for (int tt = 0; tt < 2; ++tt)
#pragma unroll
        for (int r = 0; r < 4; ++r) sv[4 * tt + r] = P[tt][cc][r] * w[tt][cc][r];
      split8(sv, sh[cc], sl[cc]);
    }
#pragma unroll
    for (int it = 0; it < NI; ++it) {
      const bf16x8 ah = fb[TOFF + (it * 2 + 0) * 64 + lane];
      const bf16x8 al = fb[TOFF + (it * 2 + 1) * 64 + lane];
#pragma unroll
      for (int cc = 0; cc < C; ++cc) acc[it][cc] = mma3(ah, al, sh[cc], sl[cc], acc[it][cc]);
    }
    if (more) {
#pragma unroll
      for (int tt = 0; tt < 2; ++tt)
#pragma unroll
        for (int cc = 0; cc < C; ++cc)
#pragma unroll
          for (int r = 0; r < 4; ++r) w[tt][cc][r] = wn[tt][cc][r];
    }
    __builtin_amdgcn_s_waitcnt(0);   // this wave's LDS loads (and W loads) landed
    __syncthreads();                 // ... and every other wave's
    cur ^= 1;
  }
#pragma unroll
  for (int cc = 0; cc < C; ++cc) {
    const int col = cbase + 16 * cc + lr;
    if (col >= M) continue;
#pragma unroll
    for (int it = 0; it < NI; ++it)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        Tpart[(((long)s * k + b) * NP + 16 * it + 4 * lq + r) * M + col] = acc[it][cc][r];
  }
}

}  // namespace mbpt
