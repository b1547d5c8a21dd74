// FP32 space-time GW self-energy branch sums, gfx950 (v_mfma_f32_16x16x4f32).
//
//   S^<[t][s] = sum_PQ X_s[P][s] (Zt(t) o Go(t))_PQ X_s[Q][s]
//   S^>[t][s] = sum_PQ X_s[P][s] (Zt(t) o Gv(t))_PQ X_s[Q][s]
//
// with Go(t) = Yo Yo^T, Gv(t) = Yv Yv^T from the chi0 prep pass (the
// exponentials split symmetrically over the two operands), Zt(t) the
// correlation part of the screened interaction on the grid at time t, and X_s
// the collocation columns of the requested states.  (The reference's
// Sigma^> carries a minus sign; the reduction applies it.)
//
// One workgroup of 8 waves (2 x 4, 64 x 32 each) owns one 128 x 128
// upper-triangular tile of the (P, Q) plane at one time point.  It builds the
// two Green's function tiles exactly as the chi0 kernel does (K staged 16 deep
// through double-buffered LDS), multiplies both by the Zt tile in registers,
// and contracts them with the states, 16 at a time:
//
//   U[s][Q] = sum_P X_s[P][s] H[P][Q]      (MFMA: the accumulator of H is the
//                                            B operand as it stands -- element
//                                            r of lane l is row 4(l/16)+r, the
//                                            K order of the next product)
//   v[s]    = sum_Q U[s][Q] X_s[Q][s]       (a 16-lane shuffle reduction)
//
// An off-diagonal tile stands for itself and its mirror (H symmetric), so it
// counts twice.  The eight waves' sums meet in LDS and one partial per
// (time, tile, branch, state) is written; a second pass sums the tiles in a
// fixed order into FP64.
#pragma once

#include "common.h"
#include "chi0_f32.h"

namespace mbpt {

// part[((t * ntiles + tile) * 2 + branch) * nsp + s]
//
// WAVES: the waves per SIMD the register allocation is held to.  At 4 (128
// VGPRs, two workgroups per CU) the Green's function products run ~10%
// faster, but ten values spill and are reloaded once per 16-state group, so
// the host picks 4 for a few state groups and the unconstrained 143-VGPR
// build (3 waves, one workgroup per CU) for many.
template <int WAVES>
__global__ __launch_bounds__(512, WAVES) void k_gw_sigma_f32(int M, int Mp, int ko, int kv, int nt,
                                                      int ntau, const float *__restrict__ Yt,
                                                      const float *__restrict__ Zt,
                                                      const float *__restrict__ Xs, int nsp,
                                                      float *__restrict__ part, int nblk) {
  const int lid = xcd_logical(blockIdx.x, nblk);
  const int ntiles = nt * (nt + 1) / 2;
  if (lid >= ntiles * ntau) return;
  const int t = lid / ntiles, tl = lid % ntiles;
  int ti, tj;
  tri_decode2(tl, nt, ti, tj);
  const int m0 = ti * kCT, n0 = tj * kCT;
  __shared__ __attribute__((aligned(16))) float sa[2][kCBK][kCLD];
  __shared__ __attribute__((aligned(16))) float sb[2][kCBK][kCLD];
  __shared__ float red[8][2][16];
  const int K = ko + kv;
  const float *Y = Yt + (long)t * K * Mp;
  f32x4 go[4][2], gv[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) { zero(go[i][j]); zero(gv[i][j]); }
  chi0_gemm(Y, Mp, m0, n0, ko, go, sa, sb);
  chi0_gemm(Y + (long)ko * Mp, Mp, m0, n0, kv, gv, sa, sb);

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 2, wn = wave & 3, lr = lane & 15, lq = lane >> 4;
  // H = Zt o G, in place; Zt read once per element, zero outside M x M
  const float *Z = Zt + (long)t * M * M;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + 64 * wm + 16 * i + 4 * lq + r;
        const int col = n0 + 32 * wn + 16 * j + lr;
        const float z = (row < M && col < M) ? Z[(long)row * M + col] : 0.f;
        go[i][j][r] *= z;
        gv[i][j][r] *= z;
      }
  const float twice = ti == tj ? 1.f : 2.f;
  for (int s0 = 0; s0 < nsp; s0 += 16) {
    f32x4 ul[2], ug[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) { zero(ul[j]); zero(ug[j]); }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        // A[state = lr][k = lq]: X_s at row 16i + 4lq + r of this wave's rows
        const float a = Xs[(long)(m0 + 64 * wm + 16 * i + 4 * lq + r) * nsp + s0 + lr];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          ul[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, go[i][j][r], ul[j], 0, 0, 0);
          ug[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, gv[i][j][r], ug[j], 0, 0, 0);
        }
      }
    // U[s = 4lq + r][Q = 16j + lr] times X_s[Q][s], summed over Q
    float vl[4], vg[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      vl[r] = 0.f;
      vg[r] = 0.f;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const float x = Xs[(long)(n0 + 32 * wn + 16 * j + lr) * nsp + s0 + 4 * lq + r];
        vl[r] += ul[j][r] * x;
        vg[r] += ug[j][r] * x;
      }
#pragma unroll
      for (int off = 8; off >= 1; off >>= 1) {
        vl[r] += __shfl_xor(vl[r], off, 16);
        vg[r] += __shfl_xor(vg[r], off, 16);
      }
    }
    if (lr == 0) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        red[wave][0][4 * lq + r] = vl[r];
        red[wave][1][4 * lq + r] = vg[r];
      }
    }
    __syncthreads();
    if (threadIdx.x < 32) {
      const int b = threadIdx.x >> 4, s = threadIdx.x & 15;
      float v = 0.f;
#pragma unroll
      for (int w = 0; w < 8; ++w) v += red[w][b][s];
      part[(((long)t * ntiles + tl) * 2 + b) * nsp + s0 + s] = twice * v;
    }
    __syncthreads();   // red is rewritten by the next state group
  }
}

// out[(b * ntau + t) * ns + s] = (b ? -1 : 1) sum_tiles part[...]: one
// workgroup per (branch, time, 16-state group); thread (q, s) sums the tiles
// q, q + 16, ... of state s (16 consecutive states: one 64-byte read per
// tile), then the 16 partial sums of a state meet in a fixed-order tree.
__global__ __launch_bounds__(256) void k_gw_sigma_reduce(int ntau, int ntiles, int ns, int nsp,
                                                         const float *__restrict__ part,
                                                         double *__restrict__ out) {
  const int ngrp = nsp / 16;
  const int g = blockIdx.x % ngrp, t = (blockIdx.x / ngrp) % ntau, b = blockIdx.x / (ngrp * ntau);
  const int s = threadIdx.x & 15, q = threadIdx.x >> 4;
  double v = 0.0;
  for (int tl = q; tl < ntiles; tl += 16)
    v += part[(((long)t * ntiles + tl) * 2 + b) * nsp + 16 * g + s];
  __shared__ double red[16][17];
  red[q][s] = v;
  __syncthreads();
  if (q == 0) {
    double tot = 0.0;
#pragma unroll
    for (int i = 0; i < 16; ++i) tot += red[i][s];
    const int st = 16 * g + s;
    if (st < ns) out[((long)b * ntau + t) * ns + st] = b ? -tot : tot;
  }
}

// Xs[mu][j] (Mp x nsp): collocation column cols[j] of the K-major XT, zero
// past M and past the ns states
__global__ void k_gw_pack_states(int M, int Mp, int ns, int nsp, const int *__restrict__ cols,
                                 const float *__restrict__ XT, float *__restrict__ Xs) {
  const long n = (long)Mp * nsp;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x) {
    const int mu = (int)(e / nsp), j = (int)(e % nsp);
    Xs[e] = (mu < M && j < ns) ? XT[(long)cols[j] * M + mu] : 0.f;
  }
}

}  // namespace mbpt
