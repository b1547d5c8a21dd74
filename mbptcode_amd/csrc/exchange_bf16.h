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
//   P_t (16 mu x 16 nu) = X_o[mu, :] zXv[b][:, nu]          K = n_o
//   S = P o W                                               FP32, in registers
//   T[:, nu] += X_o[mu, :]^T S                              K = 32 mu, 3 MFMAs
// The two P tiles' result elements (rows 4q+r and 16+4q+r of lane q = l/16)
// are the eight K slots of lane l in the second product: slot s holds row
// pi(q, s) = 4q + s (s < 4) or 16 + 4q + (s - 4) (s >= 4), and the X_o^T
// operand is read in the same order, so the product needs no shuffle.
//
// Padding of n_o.  The P products run over K blocks of 32 occupied indices.
// The last block holds r = n_o - 32 (NKB - 1) real ones; when 3r <= 64 its
// three split terms are laid side by side along K (hi.hi over the first r
// slots, lo.hi over the next r, hi.lo over the next r), so the block costs
// ceil(3r / 32) MFMAs instead of three.  The T product computes only the
// ceil(n_o / 16) row tiles of T that hold real rows.
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

// ---- shape of the packed X_o --------------------------------------------------
// np: n_o padded to 32.  tail: MFMAs of the last K block of the P products
// (1 or 2 with the split terms side by side along K, 3 = the plain hi/lo
// pair).  nie: row tiles of T computed.
struct Bf16Shape {
  int np, nkb, nie, tail, fp, spp, spc;
};

__host__ __device__ inline Bf16Shape bf16_shape(int no) {
  Bf16Shape s;
  s.np = (no + 31) / 32 * 32;
  s.nkb = s.np / 32;
  const int r = no - 32 * (s.nkb - 1);
  s.tail = 3 * r <= 32 ? 1 : (3 * r <= 64 ? 2 : 3);
  s.nie = (no + 15) / 16;
  s.fp = 2 * (s.nkb - 1) + (s.tail == 3 ? 2 : s.tail);   // P fragments per 16-row tile
  s.spp = 128 * s.fp;                                     // P slots per chunk (two tiles)
  s.spc = s.spp + 128 * s.nie;                            // + T slots (hi, lo per tile)
  return s;
}

// ---- X_o in MFMA fragment order -------------------------------------------------
// Per 32-row chunk of the grid, in 16-byte slots (one bf16x8 per lane):
//   P part [tt < 2][f < fp][lane]   rows 16 tt + l%16; fragments 2kb + hl for
//                                   the full K blocks (j = 32 kb + 8 (l/16) + s),
//                                   then the last block (hi, lo, or the
//                                   side-by-side split terms)
//   T part [it < nie][hl < 2][lane]  rows pi(l/16, s), i = 16 it + l%16
// rows >= M and indices >= n_o are zero.
__global__ void k_pack_bf16_xo(int M, int no, const float *__restrict__ Xo,
                               bf16x8 *__restrict__ frag) {
  const Bf16Shape sh = bf16_shape(no);
  const int j0 = 32 * (sh.nkb - 1), r = no - j0;
  const long n = (long)((M + 31) / 32) * sh.spc;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x) {
    const int chunk = (int)(e / sh.spc), rem = (int)(e % sh.spc);
    const int lane = rem & 63, q = lane >> 4, g = rem >> 6;
    bf16x8 out;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      int row, j, lo;
      if (rem < sh.spp) {
        const int tt = g / sh.fp, f = g % sh.fp;
        row = 32 * chunk + 16 * tt + (lane & 15);
        if (f < 2 * (sh.nkb - 1) || sh.tail == 3) {
          j = 32 * (f >> 1) + 8 * q + s;
          lo = f & 1;
        } else {
          const int kap = 32 * (f - 2 * (sh.nkb - 1)) + 8 * q + s;
          const int term = kap < r ? 0 : (kap < 2 * r ? 1 : (kap < 3 * r ? 2 : 3));
          j = term == 3 ? no : j0 + kap - term * r;
          lo = term == 1;
        }
      } else {
        const int g2 = (rem - sh.spp) >> 6;
        row = 32 * chunk + (s < 4 ? 4 * q + s : 16 + 4 * q + (s - 4));
        j = 16 * (g2 >> 1) + (lane & 15);
        lo = g2 & 1;
      }
      const float x = (row < M && j < no) ? Xo[(long)row * no + j] : 0.f;
      const __bf16 h = (__bf16)x;
      out[s] = lo ? (__bf16)(x - (float)h) : h;
    }
    frag[e] = out;
  }
}

// ---- W in 32 x 64 tiles ---------------------------------------------------------------
// nvb >= ceil(M / 64) tiles per 32-row chunk (an even count, so that a
// 128-column workgroup's two tiles always exist; the extra one is zero).
// wtile[(c * nvb + vb) * 512 + 256 tt + 64 w + 16 q + r] holds, as one 16-byte
// slot, W[32 c + 16 tt + 4 q + 0..3][64 vb + 16 w + r]: the four W values lane
// 16 q + r of wave w multiplies its P tile tt by (P's result rows 4 q + 0..3),
// zero outside the M x M matrix.  A tile is 8 KB in the order the exchange
// reads it from LDS, so it streams in with the X_o chunk's own loads.
__global__ void k_pack_w_tiles(int M, int nvb, const float *__restrict__ W,
                               f32x4 *__restrict__ wt) {
  const long n = (long)((M + 31) / 32) * nvb * 512;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x) {
    const long tile = e / 512;
    const int slot = (int)(e % 512);
    const int c = (int)(tile / nvb), vb = (int)(tile % nvb);
    const int tt = slot >> 8, w = (slot >> 6) & 3, q = (slot >> 4) & 3, r = slot & 15;
    const int row0 = 32 * c + 16 * tt + 4 * q, col = 64 * vb + 16 * w + r;
    f32x4 v;
#pragma unroll
    for (int rr = 0; rr < 4; ++rr)
      v[rr] = (row0 + rr < M && col < M) ? W[(long)(row0 + rr) * M + col] : 0.f;
    wt[e] = v;
  }
}

// ---- the exchange ----------------------------------------------------------------
typedef __attribute__((address_space(3))) void lds_void;
// Workgroup = 4 waves; a wave computes NV vectors x NC 16-column tiles (NV = 2,
// NC = 1 shares every W value and every X_o fragment between two vectors; NV =
// 1, NC = 2 serves a single vector).  The packed X_o chunks stream through a
// double-buffered LDS stage with direct global-to-LDS 16-byte loads
// (global_load_lds_dwordx4: no staging registers, two waves fit per SIMD).
// WT: the chunk's W tile (k_pack_w_tiles) streams into the same stage and each
// lane reads its eight W values from LDS, so the only wait on global memory is
// the one at the chunk's end.  Without WT (NC = 2) the next chunk's W values
// are loaded into registers while the current chunk computes, at clamped
// indices: a row >= M meets a zero row of the packed X_o (P = 0 there) and a
// column >= M a zero column of zXv.
// NW waves per workgroup (4, or 8 with W tiles: 128 columns, one workgroup
// per CU, so the X_o stage is filled once per CU and chunk instead of twice).
template <int NP, int NIE, int TAIL, int NV, int NC, bool WT, int NW = 4>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(2))) void k_exchange_bf16(
    int M, int no, int k, const bf16x8 *__restrict__ frag, const float *__restrict__ W,
    const f32x4 *__restrict__ wtile, const float *__restrict__ zXv, float *__restrict__ Tpart,
    int nsplit, int rows_per_split, int nvb, int nblk, int nvb64) {
  constexpr int NKB = NP / 32, NFB = TAIL == 3 ? NKB : NKB - 1, NT = NV * NC;
  constexpr int TP = TAIL == 3 ? 2 : TAIL, FP = 2 * (NKB - 1) + TP;
  constexpr int NTH = 64 * NW;                    // threads
  constexpr int SPP = 128 * FP, SPC = SPP + 128 * NIE, COLS = 16 * NC * NW;
  static_assert(!WT || NC == 1, "a W tile holds 64 columns");
  static_assert(NW == 4 || (NW == 8 && WT), "eight waves only with W tiles");
  constexpr int NWT = NW / 4;                     // W tiles per stage
  constexpr int SPW = WT ? 512 * NWT : 0;         // W tile slots per stage
  __shared__ bf16x8 buf[2][SPC + SPW];
  const int kp = (k + NV - 1) / NV;
  const int lid = xcd_logical(blockIdx.x, nblk);
  if (lid >= nvb * kp * nsplit) return;
  const int bp = lid % kp, t = lid / kp, vb = t % nvb, s = t / nvb;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lr = lane & 15, lq = lane >> 4;
  int colv[NC];
  unsigned woff[NC];
#pragma unroll
  for (int ct = 0; ct < NC; ++ct) {
    colv[ct] = vb * COLS + wave * 16 * NC + 16 * ct + lr;
    woff[ct] = (unsigned)(4 * lq * M + min(colv[ct], M - 1));
  }

  // zXv[b][:, col] as the B operands of the P products, split once
  bf16x8 bzh[NT][NFB > 0 ? NFB : 1], bzl[NT][NFB > 0 ? NFB : 1], bzt[NT][TAIL == 3 ? 1 : TAIL];
  const int j0 = 32 * (NKB - 1), r = no - j0;
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int ct = 0; ct < NC; ++ct) {
      const int b = NV * bp + v, col = colv[ct], u = v * NC + ct;
      const bool ok = b < k && col < M;
      const float *zc = zXv + (long)b * no * M + col;
#pragma unroll
      for (int kb = 0; kb < NFB; ++kb) {
        float x[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int j = 32 * kb + 8 * lq + q;
          x[q] = (ok && j < no) ? zc[(long)j * M] : 0.f;
        }
        split8(x, bzh[u][kb], bzl[u][kb]);
      }
      if (TAIL < 3) {
#pragma unroll
        for (int tf = 0; tf < (TAIL == 3 ? 1 : TAIL); ++tf)
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const int kap = 32 * tf + 8 * lq + q;
            const int term = kap < r ? 0 : (kap < 2 * r ? 1 : (kap < 3 * r ? 2 : 3));
            const int j = j0 + kap - term * r;
            const float x = (ok && term < 3) ? zc[(long)j * M] : 0.f;
            const __bf16 h = (__bf16)x;
            bzt[u][tf][q] = term == 2 ? (__bf16)(x - (float)h) : (term == 3 ? (__bf16)0.f : h);
          }
      }
    }
  f32x4 acc[NIE][NT];
#pragma unroll
  for (int it = 0; it < NIE; ++it)
#pragma unroll
    for (int u = 0; u < NT; ++u) zero(acc[it][u]);

  const int c_beg = (s * rows_per_split) / 32;
  const int c_end = min((M + 31) / 32, ((s + 1) * rows_per_split) / 32);
  auto load_w = [&](int c, float (&w)[NC][2][4]) {
    if (32 * c + 32 <= M) {   // whole chunk inside the grid: uniform row bases
#pragma unroll
      for (int tt = 0; tt < 2; ++tt)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          const float *base = W + (long)(32 * c + 16 * tt + rr) * M;
#pragma unroll
          for (int ct = 0; ct < NC; ++ct) w[ct][tt][rr] = base[woff[ct]];
        }
    } else {
#pragma unroll
      for (int ct = 0; ct < NC; ++ct)
#pragma unroll
        for (int tt = 0; tt < 2; ++tt)
#pragma unroll
          for (int rr = 0; rr < 4; ++rr)
            w[ct][tt][rr] = W[(long)min(32 * c + 16 * tt + 4 * lq + rr, M - 1) * M +
                              min(colv[ct], M - 1)];
    }
  };
  // chunk c into buf[n]: slot NTH q + tid from thread tid, each wave's
  // instruction filling 1 KB of LDS at base + 16 lane
  auto fetch = [&](int c, int n) {
    const bf16x8 *src = frag + (long)c * SPC;
#pragma unroll
    for (int q = 0; q < (SPC + NTH - 1) / NTH; ++q)
      if (SPC % NTH == 0 || NTH * q + 64 * wave < SPC)
        __builtin_amdgcn_global_load_lds((const void *)(src + NTH * q + tid),
                                         (lds_void *)&buf[n][NTH * q + 64 * wave], 16, 0, 0);
    if (WT) {   // the workgroup's NWT consecutive tiles
      const f32x4 *wsrc = wtile + ((long)c * nvb64 + NWT * vb) * 512;
#pragma unroll
      for (int q = 0; q < SPW / NTH; ++q)
        __builtin_amdgcn_global_load_lds((const void *)(wsrc + NTH * q + tid),
                                         (lds_void *)&buf[n][SPC + NTH * q + 64 * wave], 16, 0,
                                         0);
    }
  };
  // this lane's W values from stage n's tile: rows 16 tt + 4 lq + rr, column colv[0]
  auto lds_w = [&](int n, float (&w)[NC][2][4]) {
#pragma unroll
    for (int tt = 0; tt < 2; ++tt) {
      const f32x4 v = *(const f32x4 *)&buf[n][SPC + 512 * (wave >> 2) + 256 * tt +
                                              64 * (wave & 3) + lane];
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) w[0][tt][rr] = v[rr];
    }
  };
  if (c_beg < c_end) fetch(c_beg, 0);
  float w[NC][2][4];
  if (!WT) load_w(c_beg, w);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  int cur = 0;
  for (int c = c_beg; c < c_end; ++c) {
    const bool more = c + 1 < c_end;
    float wn[NC][2][4];
    if (more) {
      fetch(c + 1, cur ^ 1);   // buf[cur ^ 1] was last read before the previous barrier
      if (!WT) load_w(c + 1, wn);
    }
    const bf16x8 *fb = buf[cur];
    // WT: all of the P products' fragments (and the W values) are read up
    // front, so the chunk waits on LDS once before its P products rather than
    // once per K block; the scheduling barrier keeps the compiler from sinking
    // the reads back to their uses.  (Without WT the one-vector kernel has no
    // registers to spare and reads each fragment where it is used.)
    bf16x8 fa[2][FP];
    auto pf = [&](int tt, int f) { return WT ? fa[tt][f] : fb[(tt * FP + f) * 64 + lane]; };
    if constexpr (WT) {
#pragma unroll
      for (int tt = 0; tt < 2; ++tt)
#pragma unroll
        for (int f = 0; f < FP; ++f) fa[tt][f] = fb[(tt * FP + f) * 64 + lane];
      lds_w(cur, w);
      __builtin_amdgcn_sched_barrier(0);
    }
    f32x4 P[2][NT];
#pragma unroll
    for (int tt = 0; tt < 2; ++tt)
#pragma unroll
      for (int u = 0; u < NT; ++u) zero(P[tt][u]);
#pragma unroll
    for (int tt = 0; tt < 2; ++tt) {
#pragma unroll
      for (int kb = 0; kb < NFB; ++kb)
#pragma unroll
        for (int u = 0; u < NT; ++u)
          P[tt][u] = mma3(pf(tt, 2 * kb), pf(tt, 2 * kb + 1), bzh[u][kb], bzl[u][kb], P[tt][u]);
      if (TAIL < 3) {
#pragma unroll
        for (int tf = 0; tf < (TAIL == 3 ? 1 : TAIL); ++tf)
#pragma unroll
          for (int u = 0; u < NT; ++u)
            P[tt][u] = mma_bf16(pf(tt, 2 * (NKB - 1) + tf), bzt[u][tf], P[tt][u]);
      }
    }
    // WT: the T products' fragments are in flight while S is formed and split
    bf16x8 ft[NIE][2];
    auto tf2 = [&](int it, int hl) { return WT ? ft[it][hl] : fb[SPP + (it * 2 + hl) * 64 + lane]; };
    if constexpr (WT) {
#pragma unroll
      for (int it = 0; it < NIE; ++it) {
        ft[it][0] = fb[SPP + (it * 2 + 0) * 64 + lane];
        ft[it][1] = fb[SPP + (it * 2 + 1) * 64 + lane];
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    bf16x8 sh[NT], sl[NT];
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
      for (int ct = 0; ct < NC; ++ct) {
        float sv[8];
#pragma unroll
        for (int tt = 0; tt < 2; ++tt)
#pragma unroll
          for (int rr = 0; rr < 4; ++rr) sv[4 * tt + rr] = P[tt][v * NC + ct][rr] * w[ct][tt][rr];
        split8(sv, sh[v * NC + ct], sl[v * NC + ct]);
      }
#pragma unroll
    for (int it = 0; it < NIE; ++it)
#pragma unroll
      for (int u = 0; u < NT; ++u) acc[it][u] = mma3(tf2(it, 0), tf2(it, 1), sh[u], sl[u], acc[it][u]);
    if (!WT && more) {
#pragma unroll
      for (int ct = 0; ct < NC; ++ct)
#pragma unroll
        for (int tt = 0; tt < 2; ++tt)
#pragma unroll
          for (int rr = 0; rr < 4; ++rr) w[ct][tt][rr] = wn[ct][tt][rr];
    }
    __builtin_amdgcn_s_waitcnt(0);   // this wave's LDS loads (and W loads) landed
    __syncthreads();                 // ... and every other wave's
    cur ^= 1;
  }
  // rows i >= 16 NIE of T (all beyond n_o) are not written
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const int b = NV * bp + v;
    if (b >= k) continue;
#pragma unroll
    for (int ct = 0; ct < NC; ++ct) {
      const int col = colv[ct];
      if (col >= M) continue;
#pragma unroll
      for (int it = 0; it < NIE; ++it)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr)
          Tpart[(((long)s * k + b) * NP + 16 * it + 4 * lq + rr) * M + col] =
              acc[it][v * NC + ct][rr];
    }
  }
}

// ---- the B block's exchange (full Casida) ---------------------------------------
// U^T[b][i][mu] = sum_nu X_o[nu][i] W[nu][mu] P^T[nu][mu],
// P^T[nu][mu] = sum_j zXv[b][j][nu] X_o[mu][j]   (W symmetric: W o P^T = (W o P)^T).
// The exchange above with the operands of the first product swapped: the
// per-vector zXv rows nu are the streamed A operand, packed per apply in the
// fragment order of X_o's P part (k_pack_bf16_zxv), and a wave's own columns
// mu of X_o are the B operand, split once into registers.  The second
// product is the exchange's own (X_o^T fragments, the same LDS stream).  One
// vector and NC 16-column tiles per wave, so every zXv and X_o^T fragment
// read from LDS serves NC tiles.

// zfrag[(b * nchunk + c) * spp + slot]: the P-part fragments of zXv[b] with
// rows nu = 32 c + 16 tt + l%16 and the K index j laid out as X_o's (hi / lo
// pairs, the last block's split terms side by side)
__global__ void k_pack_bf16_zxv(int M, int no, int k, const float *__restrict__ zXv,
                                bf16x8 *__restrict__ zfrag) {
  const Bf16Shape sh = bf16_shape(no);
  const int j0 = 32 * (sh.nkb - 1), r = no - j0;
  const int nch = (M + 31) / 32;
  const long n = (long)k * nch * sh.spp;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x) {
    const int b = (int)(e / ((long)nch * sh.spp));
    const int rem2 = (int)(e % ((long)nch * sh.spp));
    const int chunk = rem2 / sh.spp, rem = rem2 % sh.spp;
    const int lane = rem & 63, q = lane >> 4, g = rem >> 6;
    const int tt = g / sh.fp, f = g % sh.fp;
    const int row = 32 * chunk + 16 * tt + (lane & 15);
    const float *zb = zXv + (long)b * no * M;
    bf16x8 out;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      int j, lo;
      if (f < 2 * (sh.nkb - 1) || sh.tail == 3) {
        j = 32 * (f >> 1) + 8 * q + s;
        lo = f & 1;
      } else {
        const int kap = 32 * (f - 2 * (sh.nkb - 1)) + 8 * q + s;
        const int term = kap < r ? 0 : (kap < 2 * r ? 1 : (kap < 3 * r ? 2 : 3));
        j = term == 3 ? no : j0 + kap - term * r;
        lo = term == 1;
      }
      const float x = (row < M && j < no) ? zb[(long)j * M + row] : 0.f;
      const __bf16 h = (__bf16)x;
      out[s] = lo ? (__bf16)(x - (float)h) : h;
    }
    zfrag[e] = out;
  }
}

template <int NP, int NIE, int TAIL, int NC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_exchange_b_bf16(
    int M, int no, int k, const bf16x8 *__restrict__ frag, const bf16x8 *__restrict__ zfrag,
    const float *__restrict__ Xo, const float *__restrict__ W, float *__restrict__ Upart,
    int nsplit, int rows_per_split, int nvb, int nblk) {
  constexpr int NKB = NP / 32, NFB = TAIL == 3 ? NKB : NKB - 1;
  constexpr int TP = TAIL == 3 ? 2 : TAIL, FP = 2 * (NKB - 1) + TP;
  constexpr int SPP = 128 * FP, SPC = SPP + 128 * NIE, COLS = 64 * NC;
  __shared__ bf16x8 buf[2][SPC];   // zXv[b]'s P part, then X_o's T part
  const int nch = (M + 31) / 32;
  const int lid = xcd_logical(blockIdx.x, nblk);
  if (lid >= nvb * k * nsplit) return;
  const int b = lid % k, t = lid / k, vb = t % nvb, s = t / nvb;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lr = lane & 15, lq = lane >> 4;
  int colv[NC];
  unsigned woff[NC];
#pragma unroll
  for (int ct = 0; ct < NC; ++ct) {
    colv[ct] = vb * COLS + wave * 16 * NC + 16 * ct + lr;
    woff[ct] = (unsigned)(4 * lq * M + min(colv[ct], M - 1));
  }

  // X_o[mu][:] of this lane's column as the B operands of the P^T products
  bf16x8 bxh[NC][NFB > 0 ? NFB : 1], bxl[NC][NFB > 0 ? NFB : 1], bxt[NC][TAIL == 3 ? 1 : TAIL];
  const int j0 = 32 * (NKB - 1), r = no - j0;
#pragma unroll
  for (int ct = 0; ct < NC; ++ct) {
    const int col = colv[ct];
    const bool ok = col < M;
    const float *xr = Xo + (long)min(col, M - 1) * no;
#pragma unroll
    for (int kb = 0; kb < NFB; ++kb) {
      float x[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int j = 32 * kb + 8 * lq + q;
        x[q] = (ok && j < no) ? xr[j] : 0.f;
      }
      split8(x, bxh[ct][kb], bxl[ct][kb]);
    }
    if (TAIL < 3) {
#pragma unroll
      for (int tf = 0; tf < (TAIL == 3 ? 1 : TAIL); ++tf)
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int kap = 32 * tf + 8 * lq + q;
          const int term = kap < r ? 0 : (kap < 2 * r ? 1 : (kap < 3 * r ? 2 : 3));
          const int j = j0 + kap - term * r;
          const float x = (ok && term < 3) ? xr[j] : 0.f;
          const __bf16 h = (__bf16)x;
          bxt[ct][tf][q] = term == 2 ? (__bf16)(x - (float)h) : (term == 3 ? (__bf16)0.f : h);
        }
    }
  }
  f32x4 acc[NIE][NC];
#pragma unroll
  for (int it = 0; it < NIE; ++it)
#pragma unroll
    for (int u = 0; u < NC; ++u) zero(acc[it][u]);

  const int c_beg = (s * rows_per_split) / 32;
  const int c_end = min(nch, ((s + 1) * rows_per_split) / 32);
  auto load_w = [&](int c, float (&w)[NC][2][4]) {
    if (32 * c + 32 <= M) {
#pragma unroll
      for (int tt = 0; tt < 2; ++tt)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          const float *base = W + (long)(32 * c + 16 * tt + rr) * M;
#pragma unroll
          for (int ct = 0; ct < NC; ++ct) w[ct][tt][rr] = base[woff[ct]];
        }
    } else {
#pragma unroll
      for (int ct = 0; ct < NC; ++ct)
#pragma unroll
        for (int tt = 0; tt < 2; ++tt)
#pragma unroll
          for (int rr = 0; rr < 4; ++rr)
            w[ct][tt][rr] = W[(long)min(32 * c + 16 * tt + 4 * lq + rr, M - 1) * M +
                              min(colv[ct], M - 1)];
    }
  };
  // stage c: 64-slot groups, each wholly in zXv[b]'s P part or X_o's T part
  auto fetch = [&](int c, int n) {
    const bf16x8 *zsrc = zfrag + ((long)b * nch + c) * SPP;
    const bf16x8 *xsrc = frag + (long)c * SPC + SPP;
#pragma unroll
    for (int q = 0; q < (SPC + 255) / 256; ++q) {
      const int g0 = 256 * q + 64 * wave;
      if (SPC % 256 == 0 || g0 < SPC) {
        const bf16x8 *src = g0 < SPP ? zsrc + g0 : xsrc + (g0 - SPP);
        __builtin_amdgcn_global_load_lds((const void *)(src + lane), (lds_void *)&buf[n][g0], 16,
                                         0, 0);
      }
    }
  };
  if (c_beg < c_end) fetch(c_beg, 0);
  float w[NC][2][4];
  load_w(c_beg, w);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  int cur = 0;
  for (int c = c_beg; c < c_end; ++c) {
    const bool more = c + 1 < c_end;
    float wn[NC][2][4];
    if (more) {
      fetch(c + 1, cur ^ 1);
      load_w(c + 1, wn);
    }
    const bf16x8 *fb = buf[cur];
    f32x4 P[2][NC];
#pragma unroll
    for (int tt = 0; tt < 2; ++tt)
#pragma unroll
      for (int u = 0; u < NC; ++u) zero(P[tt][u]);
#pragma unroll
    for (int tt = 0; tt < 2; ++tt) {
#pragma unroll
      for (int kb = 0; kb < NFB; ++kb) {
        const bf16x8 ah = fb[(tt * FP + 2 * kb + 0) * 64 + lane];
        const bf16x8 al = fb[(tt * FP + 2 * kb + 1) * 64 + lane];
#pragma unroll
        for (int u = 0; u < NC; ++u) P[tt][u] = mma3(ah, al, bxh[u][kb], bxl[u][kb], P[tt][u]);
      }
      if (TAIL < 3) {
#pragma unroll
        for (int tf = 0; tf < (TAIL == 3 ? 1 : TAIL); ++tf) {
          const bf16x8 a = fb[(tt * FP + 2 * (NKB - 1) + tf) * 64 + lane];
#pragma unroll
          for (int u = 0; u < NC; ++u) P[tt][u] = mma_bf16(a, bxt[u][tf], P[tt][u]);
        }
      }
    }
    bf16x8 sh[NC], sl[NC];
#pragma unroll
    for (int ct = 0; ct < NC; ++ct) {
      float sv[8];
#pragma unroll
      for (int tt = 0; tt < 2; ++tt)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) sv[4 * tt + rr] = P[tt][ct][rr] * w[ct][tt][rr];
      split8(sv, sh[ct], sl[ct]);
    }
#pragma unroll
    for (int it = 0; it < NIE; ++it) {
      const bf16x8 ah = fb[SPP + (it * 2 + 0) * 64 + lane];
      const bf16x8 al = fb[SPP + (it * 2 + 1) * 64 + lane];
#pragma unroll
      for (int u = 0; u < NC; ++u) acc[it][u] = mma3(ah, al, sh[u], sl[u], acc[it][u]);
    }
    if (more) {
#pragma unroll
      for (int ct = 0; ct < NC; ++ct)
#pragma unroll
        for (int tt = 0; tt < 2; ++tt)
#pragma unroll
          for (int rr = 0; rr < 4; ++rr) w[ct][tt][rr] = wn[ct][tt][rr];
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    cur ^= 1;
  }
#pragma unroll
  for (int ct = 0; ct < NC; ++ct) {
    const int col = colv[ct];
    if (col >= M) continue;
#pragma unroll
    for (int it = 0; it < NIE; ++it)
#pragma unroll
      for (int rr = 0; rr < 4; ++rr)
        Upart[(((long)s * k + b) * NP + 16 * it + 4 * lq + rr) * M + col] = acc[it][ct][rr];
  }
}

}  // namespace mbpt
