// FP32 space-time chi0 kernels, gfx950 (v_mfma_f32_16x16x4f32).
//
//   out += sum_t w_t (-2) Go(t) o Gv(t),  Go(t) = Yo(t) Yo(t)^T,  Gv(t) = Yv(t) Yv(t)^T
//
// with the exponentials split symmetrically, Yo(t) = X_o e^{e_o tau_t / 2} and
// Yv(t) = X_v e^{-e_v tau_t / 2}, so both GEMM operands of a Green's function
// tile are rows of one matrix.  A prep pass writes Y for every time point in
// a K-major, zero-padded layout Yt[t][c][mu] (c over the padded occupied then
// virtual columns, mu over M padded to the tile), so the main loop needs no
// bounds checks and every global load is a coalesced float4.
//
// Main kernel: one 128 x 128 upper-triangular output tile and one chunk of the
// time points per workgroup of 8 waves (2 x 4, 64 x 32 each: 4 x 2 MFMA
// fragments per Green's function and per running sum).  K is staged 16 deep
// through double-buffered LDS (row stride 144 floats: the two 16-lane halves of
// a read land in disjoint banks).  Per time point the occupied and the virtual
// products are accumulated in registers, multiplied, and folded into the
// tile's running sum, written once per chunk to a partial buffer; a last pass
// sums the chunks into the FP64 output and mirrors the lower triangle.
#pragma once

#include "common.h"

namespace mbpt {

constexpr int kCT = 128;      // output tile edge
constexpr int kCBK = 16;      // K per stage
constexpr int kCLD = 144;     // LDS row stride (floats)

struct TauW {
  int n;
  double tau_half[MBPT_CHI0_MAX_TAU];
  float w[MBPT_CHI0_MAX_TAU];
};

// Yt[t][c][mu] (c < ko occupied, padded; then virtual, padded to kv) from the
// K-major collocation XT[c][mu] (same column layout, M columns): one block per
// (t, c) row, the exponential computed once per row.
__global__ __launch_bounds__(256) void k_chi0_prep(int M, int Mp, int no, int nv, int ko,
                                                   int kv, const float *__restrict__ XT,
                                                   const double *__restrict__ eo,
                                                   const double *__restrict__ ev, TauW tw,
                                                   float *__restrict__ Yt) {
  const int K = ko + kv;
  const int t = blockIdx.x / K, c = blockIdx.x % K;
  float f = 0.f;
  if (c < no)
    f = (float)exp(eo[c] * tw.tau_half[t]);
  else if (c >= ko && c - ko < nv)
    f = (float)exp(-ev[c - ko] * tw.tau_half[t]);
  const float *x = XT + (long)c * M;
  float *y = Yt + ((long)t * K + c) * Mp;
  for (int mu = threadIdx.x; mu < Mp; mu += 256) y[mu] = (mu < M && f != 0.f) ? x[mu] * f : 0.f;
}

// XT[c][mu]: the collocation K-major with the chi0 column padding
__global__ void k_chi0_xt(int M, int no, int nv, int ko, int kv,
                          const double *__restrict__ X, float *__restrict__ XT) {
  const long n = (long)(ko + kv) * M;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x) {
    const int c = (int)(e / M), mu = (int)(e % M);
    float v = 0.f;
    if (c < no) v = (float)X[(long)mu * (no + nv) + c];
    else if (c >= ko && c - ko < nv) v = (float)X[(long)mu * (no + nv) + no + (c - ko)];
    XT[e] = v;
  }
}

__device__ __forceinline__ void tri_decode2(int lid, int nt, int &ti, int &tj) {
  int r = 0, base = 0;
  while (lid >= base + (nt - r)) { base += nt - r; ++r; }
  ti = r;
  tj = r + (lid - base);
}

// acc[fm][fn] += A[m-rows of this wave] B[n-rows]^T over K rows of Yt
// (Yt rows are K-major: Y^T), staged through double-buffered LDS.
__device__ __forceinline__ void chi0_gemm(const float *__restrict__ Yk, int Mp, int m0,
                                          int n0, int K, f32x4 (&acc)[4][2],
                                          float (*sa)[kCBK][kCLD],
                                          float (*sb)[kCBK][kCLD]) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 2, wn = wave & 3;      // 2 x 4 waves: 64 x 32 each
  const int lr = lane & 15, lq = lane >> 4;
  // 512 threads stage 16 x 128 of A and of B: one float4 each per operand
  const int srow = tid >> 5, scol = (tid & 31) * 4;
  const float *ga = Yk + (long)srow * Mp + m0 + scol;
  const float *gb = Yk + (long)srow * Mp + n0 + scol;
  float4 ra = *(const float4 *)ga, rb = *(const float4 *)gb;
  int buf = 0;
  *(float4 *)&sa[0][srow][scol] = ra;
  *(float4 *)&sb[0][srow][scol] = rb;
  __syncthreads();
  for (int k0 = 0; k0 < K; k0 += kCBK) {
    const bool more = k0 + kCBK < K;
    if (more) {
      ra = *(const float4 *)(ga + (long)(k0 + kCBK) * Mp);
      rb = *(const float4 *)(gb + (long)(k0 + kCBK) * Mp);
    }
#pragma unroll
    for (int kk = 0; kk < kCBK / 4; ++kk) {
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

// Held to 128 VGPRs (4 waves per SIMD, two workgroups per CU, 5% faster than
// the 130-VGPR allocation's one workgroup); the one value that spills is
// stored before the loops and reloaded after them.
__global__ __launch_bounds__(512, 4) void k_chi0_f32(int Mp, int ko, int kv, int nt,
                                                  int nchunk, int tpc,
                                                  const float *__restrict__ Yt, TauW tw,
                                                  float *__restrict__ part, int nblk) {
  const int lid = xcd_logical(blockIdx.x, nblk);
  const int ntiles = nt * (nt + 1) / 2;
  if (lid >= ntiles * nchunk) return;
  const int chunk = lid / ntiles, tl = lid % ntiles;
  int ti, tj;
  tri_decode2(tl, nt, ti, tj);
  const int m0 = ti * kCT, n0 = tj * kCT;
  __shared__ __attribute__((aligned(16))) float sa[2][kCBK][kCLD];
  __shared__ __attribute__((aligned(16))) float sb[2][kCBK][kCLD];
  f32x4 chi[4][2], go[4][2], gv[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) zero(chi[i][j]);
  const int K = ko + kv;
  const int t0 = chunk * tpc, t1 = min(tw.n, t0 + tpc);
  for (int t = t0; t < t1; ++t) {
    const float *Y = Yt + (long)t * K * Mp;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) { zero(go[i][j]); zero(gv[i][j]); }
    chi0_gemm(Y, Mp, m0, n0, ko, go, sa, sb);
    chi0_gemm(Y + (long)ko * Mp, Mp, m0, n0, kv, gv, sa, sb);
    const float c = -2.f * tw.w[t];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) chi[i][j][r] += c * go[i][j][r] * gv[i][j][r];
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 2, wn = wave & 3, lr = lane & 15, lq = lane >> 4;
  float *o = part + (long)chunk * Mp * Mp;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + 64 * wm + 16 * i + 4 * lq + r;
        const int col = n0 + 32 * wn + 16 * j + lr;
        o[(long)row * Mp + col] = chi[i][j][r];
      }
}

// out = [acc ? out : 0] + sum_chunks part, one pair of mirrored 64 x 64 output
// tiles per block: the upper tile's partials are read once, and the sum is
// written to the upper tile and, transposed in LDS, to the lower one, so
// loads and stores stay coalesced and the result is exactly symmetric.
__global__ __launch_bounds__(256) void k_chi0_reduce(int M, int Mp, int nchunk, int nto,
                                                     const float *__restrict__ part,
                                                     double *__restrict__ out,
                                                     int accumulate) {
  // block pair p -> 64-row blocks (bi, bj), bi <= bj: the partials hold every
  // such block (it lies on or above the 128-tile diagonal); the sum is
  // written there and, mirrored through LDS, at (bj, bi).  A thread sums
  // four consecutive columns (16-byte loads, Mp a multiple of 128), the
  // chunk loop unrolled so that several planes' loads are in flight, and
  // stores pairs of doubles where M is even and out 16-byte aligned.
  const int p = blockIdx.x;
  int bi = 0, base = 0;
  while (p >= base + (nto - bi)) { base += nto - bi; ++bi; }
  const int bj = bi + (p - base);
  const int sr0 = bi * 64, sc0 = bj * 64;
  const long plane = (long)Mp * Mp;
  __shared__ float t[64][65];
  for (int e = threadIdx.x; e < 64 * 16; e += 256) {
    const int r = e >> 4, c = 4 * (e & 15);
    const float *src = part + (long)(sr0 + r) * Mp + sc0 + c;
    f32x4 s = {0.f, 0.f, 0.f, 0.f};
    int q = 0;
    for (; q + 4 <= nchunk; q += 4) {
      const f32x4 a0 = *(const f32x4 *)(src + q * plane);
      const f32x4 a1 = *(const f32x4 *)(src + (q + 1) * plane);
      const f32x4 a2 = *(const f32x4 *)(src + (q + 2) * plane);
      const f32x4 a3 = *(const f32x4 *)(src + (q + 3) * plane);
      s += a0;
      s += a1;
      s += a2;
      s += a3;
    }
    for (; q < nchunk; ++q) s += *(const f32x4 *)(src + q * plane);
#pragma unroll
    for (int u = 0; u < 4; ++u) t[r][c + u] = s[u];
  }
  __syncthreads();
  const bool pairs = (M & 1) == 0 && ((unsigned long)out & 15) == 0;
  for (int e = threadIdx.x; e < 64 * 32; e += 256) {
    const int r = e >> 5, c = 2 * (e & 31);
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      if (side == 1 && bi == bj) break;
      const int row = side ? sc0 + r : sr0 + r, col = side ? sr0 + c : sc0 + c;
      const double v0 = side ? (double)t[c][r] : (double)t[r][c];
      const double v1 = side ? (double)t[c + 1][r] : (double)t[r][c + 1];
      if (row >= M) continue;
      double *o = out + (long)row * M + col;
      if (pairs && col + 1 < M) {
        double2 v = {v0, v1};
        if (accumulate) {
          const double2 w = *(const double2 *)o;
          v.x += w.x;
          v.y += w.y;
        }
        *(double2 *)o = v;
      } else {
        if (col < M) o[0] = accumulate ? o[0] + v0 : v0;
        if (col + 1 < M) o[1] = accumulate ? o[1] + v1 : v1;
      }
    }
  }
}

}  // namespace mbpt
