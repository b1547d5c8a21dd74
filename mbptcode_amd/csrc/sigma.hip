// Sigma-vector of the Tamm-Dancoff ISDF BSE operator on gfx950.
//
// One application to k trial vectors is four stream-ordered launches:
//
//   head      zXv[b] = z[b] X_v^T                   (k n_o x M), K = n_v, MFMA
//   hartree   p[b]_mu = sum_j X_o[mu,j] zXv[b][j,mu];  u = V p          (k x M)
//   exchange  T[b] = X_o^T (W o (X_o zXv[b]))          (n_o x M), fused:
//             the P tile leaves one MFMA in registers, is multiplied by W in
//             registers and is the B operand of the next MFMA -- P and W o P
//             never touch memory; split over grid rows for parallelism
//   tail      Az[b] = d o z[b] - (sum_s T_s[b] - kappa X_o^T o u[b]) X_v
//
// The reference's serial loop (block_actions.py:563-661) runs these per vector
// and per grid-row tile; here all k vectors go in every launch, so the W tile
// a workgroup reads is shared by the k workgroups of one XCD through its L2.
#include "common.h"
#include "exchange_bf16.h"
#include "gemm_nt3.h"
#include "gemm_km.h"

#include <algorithm>
#include <vector>

namespace mbpt {

// ---- conversions -------------------------------------------------------------
template <typename T>
__global__ void k_split_X(int M, int no, int nv, const double *__restrict__ X,
                          T *__restrict__ Xo, T *__restrict__ Xv,
                          T *__restrict__ XoT) {
  const long n = (long)M * (no + nv);
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x) {
    const int mu = (int)(e / (no + nv)), p = (int)(e % (no + nv));
    const T x = (T)X[e];
    if (p < no) {
      Xo[(long)mu * no + p] = x;
      XoT[(long)p * M + mu] = x;
    } else {
      Xv[(long)mu * nv + (p - no)] = x;
    }
  }
}

template <typename T>
__global__ void k_convert(long n, const double *__restrict__ src, T *__restrict__ dst) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x)
    dst[e] = (T)src[e];
}

// ---- loaders -------------------------------------------------------------------
template <typename T> struct LdZ {  // A(m, k) = z[m][k], FP64 in
  const double *z; int rows, K;
  __device__ T operator()(int m, int k) const {
    return (m < rows && k < K) ? (T)z[(long)m * K + k] : T(0);
  }
};
template <typename T> struct LdBnk {  // B(k, n) = P[n][k]
  const T *p; int ld, K, cols;
  __device__ T operator()(int k, int n) const {
    return (k < K && n < cols) ? p[(long)n * ld + k] : T(0);
  }
};
template <typename T> struct LdBkn {  // B(k, n) = P[k][n]
  const T *p; int ld, K, cols;
  __device__ T operator()(int k, int n) const {
    return (k < K && n < cols) ? p[(long)k * ld + n] : T(0);
  }
};
template <typename T> struct LdArm {  // A(m, k) = P[m][k]
  const T *p; int ld, rows, K;
  __device__ T operator()(int m, int k) const {
    return (m < rows && k < K) ? p[(long)m * ld + k] : T(0);
  }
};
template <typename T> struct LdBsq {  // B(k, n) = P[k][n]^2
  const T *p; int ld, K, cols;
  __device__ T operator()(int k, int n) const {
    if (k >= K || n >= cols) return T(0);
    const T x = p[(long)k * ld + n];
    return x * x;
  }
};
template <typename T> struct LdAsq {  // A(m, k) = P[m][k]^2
  const T *p; int ld, rows, K;
  __device__ T operator()(int m, int k) const {
    if (m >= rows || k >= K) return T(0);
    const T x = p[(long)m * ld + k];
    return x * x;
  }
};
// ---- head: zXv = z X_v^T ---------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_head(int rows, int M, int nv,
                                              const double *__restrict__ z,
                                              const T *__restrict__ Xv,
                                              T *__restrict__ zXv, int ntm,
                                              int ntn, int nblk) {
  const int lid = xcd_logical(blockIdx.x, nblk);
  if (lid >= ntm * ntn) return;
  const int tm = lid / ntn, tn = lid % ntn;
  __shared__ TileSmem<T> sm;
  typename MF<T>::acc acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) zero(acc[i][j]);
  gemm64<T, true, true>(LdZ<T>{z, rows, nv}, LdBnk<T>{Xv, nv, nv, M}, tm * kTile,
                        tn * kTile, nv, acc, sm);
  for (int fm = 0; fm < 2; ++fm)
    for (int fn = 0; fn < 2; ++fn) {
      const int col = tile_col(tn * kTile, fn);
      for (int r = 0; r < 4; ++r) {
        const int row = tile_row<T>(tm * kTile, fm, r);
        if (row < rows && col < M) zXv[(long)row * M + col] = acc[fm][fn][r];
      }
    }
}

// ---- hartree ----------------------------------------------------------------------
template <typename T>
__global__ void k_hartree_p(int M, int no, int k, const T *__restrict__ XoT,
                            const T *__restrict__ zXv, T *__restrict__ p) {
  const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= (long)k * M) return;
  const int b = (int)(idx / M), mu = (int)(idx % M);
  T s = 0;
  for (int j = 0; j < no; ++j)
    s += XoT[(long)j * M + mu] * zXv[((long)b * no + j) * M + mu];
  p[idx] = s;
}

// ---- exchange: T[b] = X_o^T (W o (X_o zXv[b])) ------------------------------------
// Workgroup = (64 grid columns nu, one vector b, one split s of the grid rows mu).
// Wave w owns columns nu0 + 16w .. +16.  Per 16 rows of mu:
//   P (16 mu x 16 nu)   = X_o[mu, :] zXv[b][:, nu]       NOPAD/4 MFMAs, K = n_o
//   S_r = P_r * W[mu_r, nu]                              in registers
//   T[:, nu] += X_o[mu, :]^T S                           NOPAD/16 x 4 MFMAs
// The P result layout (row orow(l/16, r) of lane l) is used directly as the B
// operand of step r of the second product, whose k index is taken to be that
// same row, so no shuffle is needed between the two MFMAs.
//
// X_o comes packed (Xp: rows padded to 32, columns to NOPAD, zeros outside
// M x n_o) and streams through double-buffered 32-row LDS stages by direct
// 16-byte global-to-LDS loads, one instruction per row (NOPAD * size / 16
// lanes), issued a stage ahead; the next 16 rows' W are prefetched into
// registers.  Rows of a stage past M are zero rows of Xp, so P = 0 there and W
// is read at a clamped row.  LD keeps every LDS row 16-byte aligned.
typedef __attribute__((address_space(3))) void lds_ptr_t;
template <typename T, int NOPAD>
__global__ __launch_bounds__(256) void k_exchange(
    int M, int no, int k, const T *__restrict__ Xp, const T *__restrict__ W,
    const T *__restrict__ zXv, T *__restrict__ Tpart, int nsplit,
    int rows_per_split, int nvb, int nblk) {
  constexpr int NS = NOPAD / 4, NI = NOPAD / 16, ROWS = 32;
  constexpr int LD = NOPAD + 16 / (int)sizeof(T), NL = NOPAD * (int)sizeof(T) / 16;
  __shared__ __attribute__((aligned(16))) T xs[2][ROWS * LD];
  const int lid = xcd_logical(blockIdx.x, nblk);
  if (lid >= nvb * k * nsplit) return;
  const int b = lid % k, t = lid / k, vb = t % nvb, s = t / nvb;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lr = lane & 15, lq = lane >> 4;
  const int col = vb * 64 + wave * 16 + lr;
  const bool colok = col < M;
  const int colc = min(col, M - 1);

  T bz[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    const int j = 4 * q + lq;
    bz[q] = (j < no && colok) ? zXv[((long)b * no + j) * M + col] : T(0);
  }
  typename MF<T>::acc acc[NI];
#pragma unroll
  for (int it = 0; it < NI; ++it) zero(acc[it]);

  // rows_per_split is a multiple of 32: a stage never straddles two splits
  const int st_beg = (s * rows_per_split) / ROWS;
  const int st_end = min((M + ROWS - 1) / ROWS, ((s + 1) * rows_per_split) / ROWS);
  auto fetch = [&](int st, int n) {
#pragma unroll
    for (int q = 0; q < ROWS / 4; ++q) {
      const int r = 4 * q + wave;   // wave-uniform row of the stage
      if (lane < NL)
        __builtin_amdgcn_global_load_lds(
            (const void *)((const char *)(Xp + (long)(ROWS * st + r) * NOPAD) + 16 * lane),
            (lds_ptr_t *)&xs[n][r * LD], 16, 0, 0);
    }
  };
  auto load_w = [&](int mu0, T (&w)[4]) {
#pragma unroll
    for (int r = 0; r < 4; ++r) w[r] = W[(long)min(mu0 + MF<T>::orow(lq, r), M - 1) * M + colc];
  };
  // The W values of the 32 rows of stage st + 1 are loaded at the top of
  // stage st and moved into place after the stage's closing wait, so no wait
  // on them falls inside the stage (FP32: 3-4% at n_o = 100, M = 6000).  In
  // FP64 the 16 extra VGPRs cost occupancy, and W is loaded 16 rows ahead.
  if constexpr (sizeof(T) == 4) {
    if (st_beg < st_end) {
      fetch(st_beg, 0);
      T w[2][4], wn[2][4];
      load_w(ROWS * st_beg, w[0]);
      load_w(ROWS * st_beg + 16, w[1]);
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      int cur = 0;
      for (int st = st_beg; st < st_end; ++st) {
        const int sn = min(st + 1, st_end - 1);
        fetch(sn, cur ^ 1);   // the last stage reloads itself into the idle buffer
        load_w(ROWS * sn, wn[0]);
        load_w(ROWS * sn + 16, wn[1]);
        const T *x = xs[cur];
#pragma unroll
        for (int ch = 0; ch < ROWS / 16; ++ch) {
          const int r0 = ch * 16;
          typename MF<T>::acc P;
          zero(P);
#pragma unroll
          for (int q = 0; q < NS; ++q) P = MF<T>::mma(x[(r0 + lr) * LD + 4 * q + lq], bz[q], P);
          T S[4];
#pragma unroll
          for (int r = 0; r < 4; ++r) S[r] = P[r] * w[ch][r];
#pragma unroll
          for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int it = 0; it < NI; ++it)
              acc[it] = MF<T>::mma(x[(r0 + MF<T>::orow(lq, r)) * LD + 16 * it + lr], S[r],
                                   acc[it]);
        }
        __builtin_amdgcn_s_waitcnt(0);   // this wave's stage and W loads landed
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int ch = 0; ch < 2; ++ch)
#pragma unroll
          for (int r = 0; r < 4; ++r) w[ch][r] = wn[ch][r];
        __syncthreads();                 // ... and every other wave's
        cur ^= 1;
      }
    }
  } else {
    if (st_beg < st_end) {
      fetch(st_beg, 0);
      T w[4];
      load_w(ROWS * st_beg, w);
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      int cur = 0;
      for (int st = st_beg; st < st_end; ++st) {
        if (st + 1 < st_end) fetch(st + 1, cur ^ 1);   // xs[cur ^ 1] last read before the barrier
        const T *x = xs[cur];
  #pragma unroll
        for (int ch = 0; ch < ROWS / 16; ++ch) {
          const int r0 = ch * 16;
          T wn[4];
          load_w(ROWS * st + r0 + 16, wn);
          typename MF<T>::acc P;
          zero(P);
  #pragma unroll
          for (int q = 0; q < NS; ++q) P = MF<T>::mma(x[(r0 + lr) * LD + 4 * q + lq], bz[q], P);
          T S[4];
  #pragma unroll
          for (int r = 0; r < 4; ++r) S[r] = P[r] * w[r];
  #pragma unroll
          for (int r = 0; r < 4; ++r)
  #pragma unroll
            for (int it = 0; it < NI; ++it)
              acc[it] = MF<T>::mma(x[(r0 + MF<T>::orow(lq, r)) * LD + 16 * it + lr], S[r], acc[it]);
  #pragma unroll
          for (int r = 0; r < 4; ++r) w[r] = wn[r];
        }
        __builtin_amdgcn_s_waitcnt(0);   // this wave's stage loads (and W loads) landed
        __syncthreads();                 // ... and every other wave's
        cur ^= 1;
      }
    }
  }
  if (colok) {
#pragma unroll
    for (int it = 0; it < NI; ++it)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        Tpart[(((long)s * k + b) * NOPAD + 16 * it + MF<T>::orow(lq, r)) * M + col] =
            acc[it][r];
  }
}

// ---- exchange of the B block (full Casida): U^T[b] = X_o^T (W o P[b]^T) ---------
// With W symmetric, W o P^T = (W o P)^T, and the B block's exchange is U^T X_v
// with U^T[i][mu] = sum_nu X_o[nu][i] W[nu][mu] P^T[nu][mu], P^T = zXv^T X_o^T
// (block_actions.py:632-652 builds the same U as (W o P) X_o).  It is the
// exchange above with the roles of zXv and X_o swapped in the first product:
// a wave owns 16 columns mu, X_o[mu, :] sits in registers as the B operand of
// P^T (16 nu x 16 mu) = zXv[b][:, nu]^T X_o[mu, :]^T, and the rows nu stream
// -- zXv[b][:, nu] staged transposed in LDS, X_o[nu, :] by the same
// double-buffered direct loads as above.  The P^T result element r of lane l
// (row orow(l/16, r)) is again the K slot of the second product.  Partials
// U^T_s are written in the layout of T_s, so the tail is shared.  Stages are
// 16 rows (the A kernel's are 32): with zXv's stage beside X_o's in LDS and
// the next stage's operands in registers, 32-row stages held the FP64 build
// to one wave per SIMD; 16-row ones fit two (10% faster in FP64).
template <typename T, int NOPAD>
__global__ __launch_bounds__(256) void k_exchange_b(
    int M, int no, int k, const T *__restrict__ Xp, const T *__restrict__ W,
    const T *__restrict__ zXv, T *__restrict__ Upart, int nsplit,
    int rows_per_split, int nmb, int nblk) {
  constexpr int NS = NOPAD / 4, NI = NOPAD / 16, ROWS = 16;
  constexpr int LD = NOPAD + 16 / (int)sizeof(T), NL = NOPAD * (int)sizeof(T) / 16;
  constexpr int LDZ = NOPAD + 1;
  __shared__ __attribute__((aligned(16))) T xs[2][ROWS * LD];
  __shared__ T zs[ROWS * LDZ];
  const int lid = xcd_logical(blockIdx.x, nblk);
  if (lid >= nmb * k * nsplit) return;
  const int b = lid % k, t = lid / k, mb = t % nmb, s = t / nmb;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lr = lane & 15, lq = lane >> 4;
  const int col = mb * 64 + wave * 16 + lr;   // this lane's mu
  const bool colok = col < M;
  const int colc = min(col, M - 1);

  T xb[NS];   // X_o[mu][4q + l/16]: B operand of the first product
#pragma unroll
  for (int q = 0; q < NS; ++q) xb[q] = Xp[(long)colc * NOPAD + 4 * q + lq] * (colok ? T(1) : T(0));
  typename MF<T>::acc acc[NI];
#pragma unroll
  for (int it = 0; it < NI; ++it) zero(acc[it]);
  const T *zb = zXv + (long)b * no * M;

  const int st_beg = (s * rows_per_split) / ROWS;
  const int st_end = min((M + ROWS - 1) / ROWS, ((s + 1) * rows_per_split) / ROWS);
  auto fetch = [&](int st, int n) {
#pragma unroll
    for (int q = 0; q < ROWS / 4; ++q) {
      const int r = 4 * q + wave;
      if (lane < NL)
        __builtin_amdgcn_global_load_lds(
            (const void *)((const char *)(Xp + (long)(ROWS * st + r) * NOPAD) + 16 * lane),
            (lds_ptr_t *)&xs[n][r * LD], 16, 0, 0);
    }
  };
  // The next stage's operands are requested at the top of each stage, W
  // first: its values, then the X_o rows (direct to LDS), then zXv[b]'s rows
  // (into registers, stored to LDS after the stage's closing barrier).  The
  // stage itself reads only LDS and registers loaded a stage earlier, so no
  // wait inside it stalls on the loads in flight.
  constexpr int ZPT = ROWS * NOPAD / 256;   // zXv values per thread and stage
  T zr[ZPT];
  T wc[ROWS / 16][4], wn[ROWS / 16][4];
  // zXv[b][j][nu0 + r] for e = tid + 256 i = j ROWS + r: coalesced along nu,
  // zero past M and n_o
  auto zload = [&](int st) {
#pragma unroll
    for (int i = 0; i < ZPT; ++i) {
      const int e = tid + 256 * i, j = e / ROWS, r = e % ROWS, nu = ROWS * st + r;
      zr[i] = (j < no && nu < M) ? zb[(long)j * M + nu] : T(0);
    }
  };
  auto zstore = [&]() {
#pragma unroll
    for (int i = 0; i < ZPT; ++i) {
      const int e = tid + 256 * i;
      zs[(e % ROWS) * LDZ + e / ROWS] = zr[i];
    }
  };
  // W[nu][mu] at the rows of P^T's accumulator (rows past M clamped: P = 0)
  auto wload = [&](int st, T (&w)[ROWS / 16][4]) {
#pragma unroll
    for (int ch = 0; ch < ROWS / 16; ++ch)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int nu = min(ROWS * st + 16 * ch + MF<T>::orow(lq, r), M - 1);
        w[ch][r] = W[(long)nu * M + colc];
      }
  };
  if (st_beg < st_end) {
    wload(st_beg, wc);
    fetch(st_beg, 0);
    zload(st_beg);
    zstore();
    int cur = 0;
    for (int st = st_beg; st < st_end; ++st) {
      __builtin_amdgcn_s_waitcnt(0);   // this wave's X_o stage landed
      __syncthreads();                 // ... every wave's, and zs is written
      const bool more = st + 1 < st_end;
      if (more) {
        wload(st + 1, wn);
        fetch(st + 1, cur ^ 1);
        zload(st + 1);
      }
      const T *x = xs[cur];
#pragma unroll
      for (int ch = 0; ch < ROWS / 16; ++ch) {
        const int r0 = ch * 16;
        typename MF<T>::acc P;
        zero(P);
#pragma unroll
        for (int q = 0; q < NS; ++q) P = MF<T>::mma(zs[(r0 + lr) * LDZ + 4 * q + lq], xb[q], P);
        T S[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) S[r] = P[r] * wc[ch][r];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int it = 0; it < NI; ++it)
            acc[it] = MF<T>::mma(x[(r0 + MF<T>::orow(lq, r)) * LD + 16 * it + lr], S[r], acc[it]);
      }
      __syncthreads();   // every wave is done with zs and xs[cur]
      if (more) {
        zstore();
#pragma unroll
        for (int ch = 0; ch < ROWS / 16; ++ch)
#pragma unroll
          for (int r = 0; r < 4; ++r) wc[ch][r] = wn[ch][r];
      }
      cur ^= 1;
    }
  }
  if (colok) {
#pragma unroll
    for (int it = 0; it < NI; ++it)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        Upart[(((long)s * k + b) * NOPAD + 16 * it + MF<T>::orow(lq, r)) * M + col] =
            acc[it][r];
  }
}

// Xp (Mp x nopad) = X_o zero-padded, the exchange's staged copy
template <typename T>
__global__ void k_pack_xo(int M, int no, int Mp, int nopad, const T *__restrict__ Xo,
                          T *__restrict__ Xp) {
  const long n = (long)Mp * nopad;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x) {
    const int r = (int)(e / nopad), c = (int)(e % nopad);
    Xp[e] = (r < M && c < no) ? Xo[(long)r * no + c] : T(0);
  }
}

// u_q[b][nu] = sum_{mu in chunk q} p[b][mu] V[mu][nu] on 16x16x4 MFMAs (FP32 or
// FP64): a wave owns 64 columns (four 16-column tiles), a workgroup 256; the
// trial vectors go in KT tiles of 16 (16 KT per pass over V, so V is streamed
// once for k <= 64).  p is staged in LDS 64 rows at a time (row-major [mu][b],
// the A operand read straight from it); each lane loads its 16 values of V
// (4 rows x 4 tiles) before the MFMAs that use them.
template <typename T, int KT>
__global__ __launch_bounds__(256) void k_hartree_u(int M, int k, int chunk,
                                                   const T *__restrict__ p,
                                                   const T *__restrict__ V,
                                                   T *__restrict__ u) {
  constexpr int KV = 16 * KT, LD = KV + 1;
  __shared__ T ps[64][LD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lq = lane >> 4;
  const int nu0 = blockIdx.x * 256 + wave * 64, q = blockIdx.y;
  const int mu_beg = q * chunk, mu_end = min(M, mu_beg + chunk);
  for (int v0 = 0; v0 < k; v0 += KV) {
    typename MF<T>::acc acc[KT][4];
#pragma unroll
    for (int kt = 0; kt < KT; ++kt)
#pragma unroll
      for (int ct = 0; ct < 4; ++ct) zero(acc[kt][ct]);
    for (int mg = mu_beg; mg < mu_end; mg += 64) {
      __syncthreads();
      for (int e = tid; e < KV * 64; e += 256) {
        const int v = e >> 6, r = e & 63, mu = mg + r;
        ps[r][v] = (v0 + v < k && mu < mu_end) ? p[(long)(v0 + v) * M + mu] : T(0);
      }
      __syncthreads();
      const int rows = min(64, mu_end - mg);
      for (int s0 = 0; s0 < rows; s0 += 16) {
        T x[4][4];
#pragma unroll
        for (int st = 0; st < 4; ++st) {
          const int mu = mg + s0 + 4 * st + lq;
#pragma unroll
          for (int ct = 0; ct < 4; ++ct) {
            const int nu = nu0 + 16 * ct + lr;
            x[st][ct] = (mu < mu_end && nu < M) ? V[(long)mu * M + nu] : T(0);
          }
        }
#pragma unroll
        for (int st = 0; st < 4; ++st) {
          T a[KT];
#pragma unroll
          for (int kt = 0; kt < KT; ++kt) a[kt] = ps[s0 + 4 * st + lq][16 * kt + lr];
#pragma unroll
          for (int kt = 0; kt < KT; ++kt)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct)
              acc[kt][ct] = MF<T>::mma(a[kt], x[st][ct], acc[kt][ct]);
        }
      }
    }
#pragma unroll
    for (int kt = 0; kt < KT; ++kt)
#pragma unroll
      for (int ct = 0; ct < 4; ++ct) {
        const int nu = nu0 + 16 * ct + lr;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int b = v0 + 16 * kt + MF<T>::orow(lq, r);
          if (b < k && nu < M) u[((long)q * k + b) * M + nu] = acc[kt][ct][r];
        }
      }
  }
}

// p[e] = sum_q p[q n + e], in place (each element owned by one thread)
template <typename T>
__global__ void k_sum_partials(long n, int nq, T *__restrict__ p) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x) {
    T s = p[e];
    for (int q = 1; q < nq; ++q) s += p[(long)q * n + e];
    p[e] = s;
  }
}

// T'^T (K-major, for the tail GEMM): Tct[mu][m] = sum_s Tpart_s[m][mu]
// - kappa X_o^T[i][mu] sum_q u_q[b][mu], m = b n_o + i, with the padding up to
// (Mp, ldt) written as zero; 64 x 64 tiles transposed through LDS.
__global__ __launch_bounds__(256) void k_tail_prep_t(int k, int no, int nopad, int M, int Mp,
                                                     int ldt, int nsplit,
                                                     const float *__restrict__ Tp,
                                                     const float *__restrict__ XoT,
                                                     const float *__restrict__ u, int nqu,
                                                     float kappa, float *__restrict__ Tct) {
  __shared__ float t[64][65];
  const int m0 = blockIdx.x * 64, mu0 = blockIdx.y * 64, rows = k * no;
  const long sstride = (long)k * nopad * M;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int r = e >> 6, c = e & 63, m = m0 + r, mu = mu0 + c;
    float val = 0.f;
    if (m < rows && mu < M) {
      const int b = m / no, i = m - b * no;
      const long off = ((long)b * nopad + i) * M + mu;
      val = Tp[off];
      for (int q = 1; q < nsplit; ++q) val += Tp[off + q * sstride];
      if (u) {
        float ub = u[(long)b * M + mu];
        for (int q = 1; q < nqu; ++q) ub += u[((long)q * k + b) * M + mu];
        val -= kappa * XoT[(long)i * M + mu] * ub;
      }
    }
    t[r][c] = val;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int r = e >> 6, c = e & 63, mu = mu0 + r, m = m0 + c;
    if (mu < Mp && m < ldt) Tct[(long)mu * ldt + m] = t[c][r];
  }
}

// T' = sum_s Tpart_s - kappa X_o^T o u as hi / lo BF16 rows [m][mu] (m = b n_o
// + i), zero-padded to ldr x Mp32: the A operand of the split-BF16 tail GEMM
__global__ void k_tail_prep_b3(int k, int no, int nopad, int M, int Mp32, int ldr, int nsplit,
                               const float *__restrict__ Tp, const float *__restrict__ XoT,
                               const float *__restrict__ u, float kappa,
                               __bf16 *__restrict__ Th, __bf16 *__restrict__ Tl) {
  const int rows = k * no;
  const long sstride = (long)k * nopad * M, n = (long)ldr * Mp32;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x) {
    const int m = (int)(e / Mp32), mu = (int)(e % Mp32);
    float val = 0.f;
    if (m < rows && mu < M) {
      const int b = m / no, i = m - b * no;
      const long off = ((long)b * nopad + i) * M + mu;
      val = Tp[off];
      for (int q = 1; q < nsplit; ++q) val += Tp[off + q * sstride];
      if (u) val -= kappa * XoT[(long)i * M + mu] * u[(long)b * M + mu];
    }
    const __bf16 h = (__bf16)val;
    Th[e] = h;
    Tl[e] = (__bf16)(val - (float)h);
  }
}

// row-padded copy out[r][c] = in[r][c] (zero outside rows x cols) of pr x pc
__global__ void k_pad_copy(int rows, int cols, const float *__restrict__ in, int pr, int pc,
                           float *__restrict__ out) {
  const long n = (long)pr * pc;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x) {
    const int r = (int)(e / pc), c = (int)(e % pc);
    out[e] = (r < rows && c < cols) ? in[(long)r * cols + c] : 0.f;
  }
}

// ---- tail: Az = d o z - T' X_v, T' = sum_s T_s - kappa X_o^T o u -------------------
// prep: the split partials summed and the Hartree term folded in, one coalesced
// pass writing T' as (k nocc, M);  tail: T' X_v with the grid index split in
// nq chunks (enough workgroups for the skinny output);  finish: d o z minus
// the chunk partials, in FP64.
template <typename T>
__global__ void k_tail_prep(int k, int no, int nopad, int M, int nsplit,
                            const T *__restrict__ Tp, const T *__restrict__ XoT,
                            const T *__restrict__ u, int nqu, T kappa,
                            T *__restrict__ Tc) {
  const long n = (long)k * no * M;
  const long sstride = (long)k * nopad * M;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x) {
    const int m = (int)(e / M), mu = (int)(e % M);
    const int b = m / no, i = m - b * no;
    const long off = ((long)b * nopad + i) * M + mu;
    T s = Tp[off];
    for (int q = 1; q < nsplit; ++q) s += Tp[off + q * sstride];
    if (u) {
      T ub = u[(long)b * M + mu];
      for (int q = 1; q < nqu; ++q) ub += u[((long)q * k + b) * M + mu];
      s -= kappa * XoT[(long)i * M + mu] * ub;
    }
    Tc[e] = s;
  }
}

// C_q (rows x cols) = A[:, K-chunk q] B[K-chunk q, :] for q < nq, A row-major
// (lda), B row-major (ldb): the skinny products of the tail (T' X_v, K = M)
// and of the Hartree term (p V, K = M) with the contraction split over enough
// workgroups to fill the chip; the chunks are summed by the consumer.
template <typename T>
__global__ __launch_bounds__(256) void k_gemm_nn_splitk(
    int rows, int cols, int K, int kchunk, const T *__restrict__ A, int lda,
    const T *__restrict__ B, int ldb, T *__restrict__ C, int ntm, int ntn, int nq,
    int nblk) {
  const int lid = xcd_logical(blockIdx.x, nblk);
  if (lid >= ntm * ntn * nq) return;
  const int tn = lid % ntn, tm = (lid / ntn) % ntm, q = lid / (ntn * ntm);
  const int k0 = q * kchunk, kc = min(kchunk, K - k0);
  __shared__ TileSmem<T> sm;
  typename MF<T>::acc acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) zero(acc[i][j]);
  gemm64<T, true, false>(LdArm<T>{A + k0, lda, rows, kc},
                         LdBkn<T>{B + (long)k0 * ldb, ldb, kc, cols}, tm * kTile,
                         tn * kTile, kc, acc, sm);
  T *out = C + (long)q * rows * cols;
  for (int fm = 0; fm < 2; ++fm)
    for (int fn = 0; fn < 2; ++fn) {
      const int c = tile_col(tn * kTile, fn);
      for (int r = 0; r < 4; ++r) {
        const int m = tile_row<T>(tm * kTile, fm, r);
        if (m < rows && c < cols) out[(long)m * cols + c] = acc[fm][fn][r];
      }
    }
}

template <typename T>
__global__ void k_tail_finish(int rows, int no, int nv, int nq, const T *__restrict__ Ep,
                              const double *__restrict__ d, const double *__restrict__ z,
                              double *__restrict__ Az) {
  const long n = (long)rows * nv;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x) {
    const int m = (int)(e / nv), a = (int)(e % nv), i = m % no;
    T s = Ep[e];
    for (int q = 1; q < nq; ++q) s += Ep[(long)q * n + e];
    Az[e] = (d ? d[(long)i * nv + a] * z[e] : 0.0) - (double)s;
  }
}

// ---- screened diagonal: d - (X_o o X_o)^T W (X_v o X_v) -------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_wsq(int M, int nv, const T *__restrict__ W,
                                             const T *__restrict__ Xv,
                                             T *__restrict__ Q, int ntm, int ntn,
                                             int nblk) {
  const int lid = xcd_logical(blockIdx.x, nblk);
  if (lid >= ntm * ntn) return;
  const int tm = lid / ntn, tn = lid % ntn;
  __shared__ TileSmem<T> sm;
  typename MF<T>::acc acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) zero(acc[i][j]);
  gemm64<T, true, false>(LdArm<T>{W, M, M, M}, LdBsq<T>{Xv, nv, M, nv},
                         tm * kTile, tn * kTile, M, acc, sm);
  for (int fm = 0; fm < 2; ++fm)
    for (int fn = 0; fn < 2; ++fn) {
      const int a = tile_col(tn * kTile, fn);
      for (int r = 0; r < 4; ++r) {
        const int mu = tile_row<T>(tm * kTile, fm, r);
        if (mu < M && a < nv) Q[(long)mu * nv + a] = acc[fm][fn][r];
      }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_sdiag(int M, int no, int nv,
                                               const T *__restrict__ XoT,
                                               const T *__restrict__ Q,
                                               const double *__restrict__ d,
                                               double *__restrict__ out, int ntm,
                                               int ntn, int nblk) {
  const int lid = xcd_logical(blockIdx.x, nblk);
  if (lid >= ntm * ntn) return;
  const int tm = lid / ntn, tn = lid % ntn;
  __shared__ TileSmem<T> sm;
  typename MF<T>::acc acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) zero(acc[i][j]);
  gemm64<T, true, false>(LdAsq<T>{XoT, M, no, M}, LdBkn<T>{Q, nv, M, nv},
                         tm * kTile, tn * kTile, M, acc, sm);
  for (int fm = 0; fm < 2; ++fm)
    for (int fn = 0; fn < 2; ++fn) {
      const int a = tile_col(tn * kTile, fn);
      for (int r = 0; r < 4; ++r) {
        const int i = tile_row<T>(tm * kTile, fm, r);
        if (i < no && a < nv)
          out[(long)i * nv + a] = d[(long)i * nv + a] - (double)acc[fm][fn][r];
      }
    }
}

}  // namespace mbpt

// ===================================================================================
// handle
// ===================================================================================
using namespace mbpt;

// per-apply workspace slots of the handle
enum { WS_ZXV, WS_P, WS_U, WS_TPART, WS_TC, WS_EP, WS_ZT, WS_UPART, WS_ZFRAG, WS_N };

struct mbpt_sigma {
  int M, no, nv, nopad, prec;
  double kappa;
  size_t esz;              // bytes per working-precision element
  void *Xo, *Xv, *XoT, *W, *V;
  void *Xp;                // X_o zero-padded to (M rounded to 32) x nopad (FP32 / FP64)
  bf16x8 *frag;            // X_o split and packed in MFMA fragment order (MBPT_BF16X3)
  f32x4 *wtile;            // W in the exchange's 32 x 64 LDS tiles (MBPT_BF16X3)
  // FP32 working precision: K-major, zero-padded copies for the head and tail
  // GEMMs, X_v^T (nvp x Mq) and X_v (Mp x nvq); p pads to 16, q pads to 128
  float *XvT, *XvP;
  int nvp, Mq, Mp, nvq;
  // split-BF16 head and tail operands (hi, lo), K contiguous: X_v (Mq x nvp32)
  // and X_v^T (nvq x Mp32); nvp32 / Mp32 pad n_vir / M to 32
  __bf16 *XvB[2], *XvTB[2];
  int nvp32, Mp32;
  double *d;
  // per-apply workspace, grown on demand (slots WS_*)
  void *ws[WS_N];
  size_t ws_cap[WS_N];     // bytes
  size_t held;
  int prof;                // stage timing on
  // timing events, five per timed apply, read back (and reused) only when
  // the figures are asked for, so a timed apply does not wait for the device
  std::vector<hipEvent_t> ev;
  std::vector<double> ev_flops;   // exchange flops of each pending apply
  double prof_ms[4], prof_flops;
  long prof_launches;
};

namespace {

constexpr int kMaxNocc = 128;
constexpr int kTargetBlocks = 2048;

int grid1(long n) { return (int)std::min<long>((n + 255) / 256, 65535L); }

// W tiles (32 x 64) per 32-row chunk in the packed layout: ceil(M / 64)
// rounded up to even, so that the 128-column workgroups' two tiles exist
int w_tiles_per_chunk(int M) { return ((M + 63) / 64 + 1) / 2 * 2; }

// Timed applies waiting to be read: at most this many, then the oldest
// figures are collected (which waits for them) before more are recorded.
constexpr size_t kMaxPendingApplies = 512;

// Adds the stage times of every pending timed apply to the totals (waiting
// for each one's closing event: the applies may have run on several
// streams) and frees their events for reuse.
int prof_collect(mbpt_sigma *h) {
  const size_t n = h->ev_flops.size();
  for (size_t a = 0; a < n; ++a) {
    MBPT_HIP_TRY(hipEventSynchronize(h->ev[5 * a + 4]));
    for (int q = 0; q < 4; ++q) {
      float ms = 0.f;
      MBPT_HIP_TRY(hipEventElapsedTime(&ms, h->ev[5 * a + q], h->ev[5 * a + q + 1]));
      h->prof_ms[q] += ms;
    }
    h->prof_flops += h->ev_flops[a];
    h->prof_launches += 1;
  }
  h->ev_flops.clear();
  return MBPT_OK;
}

// The five events of the next timed apply (created as the pool grows).
int prof_events(mbpt_sigma *h, hipEvent_t **out) {
  if (h->ev_flops.size() >= kMaxPendingApplies) {
    const int rc = prof_collect(h);
    if (rc) return rc;
  }
  const size_t need = 5 * (h->ev_flops.size() + 1);
  while (h->ev.size() < need) {
    hipEvent_t e;
    MBPT_HIP_TRY(hipEventCreate(&e));
    h->ev.push_back(e);
  }
  *out = &h->ev[need - 5];
  return MBPT_OK;
}

void free_ws(mbpt_sigma *h) {
  for (int q = 0; q < WS_N; ++q) {
    (void)hipFree(h->ws[q]);
    h->ws[q] = nullptr;
    h->ws_cap[q] = 0;
  }
}

int nsplit_for(const mbpt_sigma *h, int k) {
  const int nvb = (h->M + 63) / 64;
  int ns = (kTargetBlocks + nvb * k - 1) / (nvb * k);
  ns = std::max(1, std::min(ns, (h->M + 63) / 64));
  return ns;
}

// grid-index chunks of the tail GEMM: enough workgroups to fill the chip
int nq_for(int rows, int nv, int M) {
  const int tiles = ((rows + 63) / 64) * ((nv + 63) / 64);
  int nq = (512 + tiles - 1) / tiles;
  return std::max(1, std::min(nq, (M + 63) / 64));
}

int ensure_buf(mbpt_sigma *h, int q, size_t bytes) {
  if (bytes <= h->ws_cap[q]) return MBPT_OK;
  (void)hipFree(h->ws[q]);
  h->ws[q] = nullptr;
  h->ws_cap[q] = 0;
  const size_t grow = std::max(bytes, h->ws_cap[q] + h->ws_cap[q] / 4);
  MBPT_HIP_TRY(hipMalloc(&h->ws[q], grow));
  h->ws_cap[q] = grow;
  return MBPT_OK;
}

// The work decomposition of one apply of k vectors: K splits of the
// exchange, the tail's K chunks, the Hartree row chunks.
struct ApplyPlan {
  int nsplit, ldr, nq, kchunk, nqu, hchunk;
};

ApplyPlan plan_for(const mbpt_sigma *h, int k) {
  ApplyPlan a;
  const int M = h->M, no = h->no, nv = h->nv;
  a.nsplit = nsplit_for(h, k);
  const int rows = k * no;
  const bool km = h->esz == 4 && h->prec != MBPT_BF16X3;  // FP32: K-major head and tail
  a.ldr = (rows + 127) / 128 * 128;
  // tail: K = M split in nq chunks of kchunk (multiples of 16), enough
  // workgroups for the skinny output
  if (h->prec == MBPT_BF16X3) {
    // K split by a cost model: rounds of the 512 resident workgroups times
    // the K steps (32 deep) each one runs, plus two steps' worth for its
    // partial tile; a split whose last round is nearly empty loses to a
    // finer or a coarser one
    const int tiles = (a.ldr / 128) * (h->nvq / 128);
    const int ksteps = h->Mp32 / 32;
    double best = 1e300;
    a.nq = 1;
    for (int nq = 1; nq <= std::max(1, h->Mp32 / 64); ++nq) {
      const long rounds = ((long)tiles * nq + 511) / 512;
      const double cost = (double)rounds * ((ksteps + nq - 1) / nq + 2);
      if (cost < best - 1e-9) { best = cost; a.nq = nq; }
    }
    a.kchunk = ((h->Mp32 + a.nq - 1) / a.nq + 31) / 32 * 32;
    a.nq = (h->Mp32 + a.kchunk - 1) / a.kchunk;
  } else if (km) {
    const int tiles = (a.ldr / 128) * ((nv + 127) / 128);
    a.nq = std::max(1, std::min((512 + tiles - 1) / tiles, h->Mp / 64));
    a.kchunk = ((h->Mp + a.nq - 1) / a.nq + 15) / 16 * 16;
    a.nq = (h->Mp + a.kchunk - 1) / a.kchunk;
  } else {
    a.nq = nq_for(rows, nv, M);
    a.kchunk = ((M + a.nq - 1) / a.nq + 15) / 16 * 16;
    a.nq = (M + a.kchunk - 1) / a.kchunk;
  }
  // Hartree u = p V streamed: 256-column blocks x nqu row chunks
  a.nqu = std::max(1, std::min((512 + (M + 255) / 256 - 1) / ((M + 255) / 256), (M + 63) / 64));
  a.hchunk = (M + a.nqu - 1) / a.nqu;
  a.nqu = (M + a.hchunk - 1) / a.hchunk;
  return a;
}

// Bytes of each workspace slot for an apply of k vectors.  The initializer
// must name every slot: its length is checked against WS_N at compile time.
void ws_sizes(const mbpt_sigma *h, int k, const ApplyPlan &a, bool want_b, size_t *out) {
  const size_t e = h->esz, M = h->M, no = h->no, nv = h->nv;
  const bool km = h->esz == 4;   // K-major head/tail operands (FP32 working precision)
  const size_t sizes[] = {e * k * no * M,                                   // WS_ZXV
                          e * k * M,                                        // WS_P
                          e * (size_t)a.nqu * k * M,                        // WS_U
                          e * (size_t)a.nsplit * k * h->nopad * M,          // WS_TPART
                          km ? 0 : e * k * no * M,                          // WS_TC
                          e * (size_t)a.nq * k * no * nv,                   // WS_EP
                          km ? 4 * (size_t)std::max(h->nvp32, h->Mp32) * a.ldr : 0,  // WS_ZT
                          want_b ? e * (size_t)a.nsplit * k * h->nopad * M : 0,       // WS_UPART
                          // split BF16, B block: zXv in fragment order        // WS_ZFRAG
                          want_b && h->prec == MBPT_BF16X3
                              ? 16 * (size_t)k * ((M + 31) / 32) * bf16_shape(h->no).spp
                              : 0};
  static_assert(sizeof(sizes) / sizeof(sizes[0]) == WS_N,
                "every workspace slot needs its size here");
  for (int q = 0; q < WS_N; ++q) out[q] = sizes[q];
}

int ensure_ws(mbpt_sigma *h, int k, const ApplyPlan &a, bool want_b) {
  size_t sizes[WS_N];
  ws_sizes(h, k, a, want_b, sizes);
  for (int q = 0; q < WS_N; ++q) {
    int rc = ensure_buf(h, q, sizes[q]);
    if (rc) return rc;
  }
  return MBPT_OK;
}

template <typename T> int create_impl(mbpt_sigma *h, const double *X, const double *W,
                                      const double *V, hipStream_t st) {
  const long M = h->M;
  k_split_X<T><<<grid1(M * (h->no + h->nv)), 256, 0, st>>>(
      h->M, h->no, h->nv, X, (T *)h->Xo, (T *)h->Xv, (T *)h->XoT);
  int rc = launch_status("k_split_X");
  if (rc) return rc;
  k_convert<T><<<grid1(M * M), 256, 0, st>>>(M * M, W, (T *)h->W);
  if ((rc = launch_status("k_convert(W)"))) return rc;
  if (h->kappa != 0.0) {
    k_convert<T><<<grid1(M * M), 256, 0, st>>>(M * M, V, (T *)h->V);
    if ((rc = launch_status("k_convert(V)"))) return rc;
  }
  return MBPT_OK;
}

template <typename T, int NOPAD>
void launch_exchange(const mbpt_sigma *h, int k, int nsplit, int rps, hipStream_t st) {
  const int nvb = (h->M + 63) / 64;
  const int nblk = pad_grid((long)nvb * k * nsplit);
  k_exchange<T, NOPAD><<<nblk, 256, 0, st>>>(h->M, h->no, k, (const T *)h->Xp,
                                             (const T *)h->W, (const T *)h->ws[WS_ZXV],
                                             (T *)h->ws[WS_TPART], nsplit, rps, nvb, nblk);
}

template <typename T, int NOPAD>
void launch_exchange_b(const mbpt_sigma *h, int k, int nsplit, int rps, T *Upart,
                       hipStream_t st) {
  const int nmb = (h->M + 63) / 64;
  const int nblk = pad_grid((long)nmb * k * nsplit);
  k_exchange_b<T, NOPAD><<<nblk, 256, 0, st>>>(h->M, h->no, k, (const T *)h->Xp,
                                               (const T *)h->W, (const T *)h->ws[WS_ZXV],
                                               Upart, nsplit, rps, nmb, nblk);
}

// two vectors per wave: W streams in its packed tiles; one vector, two column
// tiles: W from global memory
template <int NP, int NIE, int TAIL, int NV>
void launch_exchange_bf16_nv(const mbpt_sigma *h, int k, int nsplit, int rps, hipStream_t st) {
  constexpr int NC = NV == 2 ? 1 : 2, NW = NV == 2 ? 8 : 4;
  const int cols = 16 * NC * NW;
  const int nvb = (h->M + cols - 1) / cols, kp = (k + NV - 1) / NV;
  const int nblk = pad_grid((long)nvb * kp * nsplit);
  k_exchange_bf16<NP, NIE, TAIL, NV, NC, NV == 2, NW><<<nblk, 64 * NW, 0, st>>>(
      h->M, h->no, k, h->frag, (const float *)h->W, h->wtile, (const float *)h->ws[WS_ZXV],
      (float *)h->ws[WS_TPART], nsplit, rps, nvb, nblk, w_tiles_per_chunk(h->M));
}

template <int NP, int NIE, int TAIL>
void launch_exchange_b_bf16_k(const mbpt_sigma *h, int k, int nsplit, int rps, float *Upart,
                              hipStream_t st) {
  constexpr int NC = 2;
  const int nvb = (h->M + 64 * NC - 1) / (64 * NC);
  const int nblk = pad_grid((long)nvb * k * nsplit);
  k_exchange_b_bf16<NP, NIE, TAIL, NC><<<nblk, 256, 0, st>>>(
      h->M, h->no, k, h->frag, (const bf16x8 *)h->ws[WS_ZFRAG], (const float *)h->Xo,
      (const float *)h->W, Upart, nsplit, rps, nvb, nblk);
}

// the B block's split-BF16 exchange: zXv packed, then the kernel of this
// n_o's (row tiles, tail MFMAs)
template <int NP>
int launch_exchange_b_bf16(const mbpt_sigma *h, int k, int nsplit, int rps, float *Upart,
                           hipStream_t st) {
  constexpr int NI = NP / 16;
  const Bf16Shape sh = bf16_shape(h->no);
  if (sh.np != NP) { set_error("internal: bf16 padding"); return MBPT_ERR_VALUE; }
  const long n = (long)k * ((h->M + 31) / 32) * sh.spp;
  k_pack_bf16_zxv<<<(int)std::min<long>((n + 255) / 256, 16384L), 256, 0, st>>>(
      h->M, h->no, k, (const float *)h->ws[WS_ZXV], (bf16x8 *)h->ws[WS_ZFRAG]);
  int rc = launch_status("k_pack_bf16_zxv");
  if (rc) return rc;
  if (sh.nie == NI - 1 && sh.tail == 1) launch_exchange_b_bf16_k<NP, NI - 1, 1>(h, k, nsplit, rps, Upart, st);
  else if (sh.nie == NI - 1 && sh.tail == 2) launch_exchange_b_bf16_k<NP, NI - 1, 2>(h, k, nsplit, rps, Upart, st);
  else if (sh.nie == NI && sh.tail == 2) launch_exchange_b_bf16_k<NP, NI, 2>(h, k, nsplit, rps, Upart, st);
  else if (sh.nie == NI && sh.tail == 3) launch_exchange_b_bf16_k<NP, NI, 3>(h, k, nsplit, rps, Upart, st);
  else { set_error("internal: bf16 shape"); return MBPT_ERR_VALUE; }
  return launch_status("k_exchange_b_bf16");
}

// two vectors per wave whenever there are two
template <int NP, int NIE, int TAIL>
void launch_exchange_bf16_k(const mbpt_sigma *h, int k, int nsplit, int rps, hipStream_t st) {
  if (k >= 2) launch_exchange_bf16_nv<NP, NIE, TAIL, 2>(h, k, nsplit, rps, st);
  else launch_exchange_bf16_nv<NP, NIE, TAIL, 1>(h, k, nsplit, rps, st);
}

// the (row tiles, tail MFMAs) pairs an n_o in (NP - 32, NP] can have
template <int NP>
int launch_exchange_bf16(const mbpt_sigma *h, int k, int nsplit, int rps, hipStream_t st) {
  constexpr int NI = NP / 16;
  const Bf16Shape sh = bf16_shape(h->no);
  if (sh.np != NP) { set_error("internal: bf16 padding"); return MBPT_ERR_VALUE; }
  if (sh.nie == NI - 1 && sh.tail == 1) launch_exchange_bf16_k<NP, NI - 1, 1>(h, k, nsplit, rps, st);
  else if (sh.nie == NI - 1 && sh.tail == 2) launch_exchange_bf16_k<NP, NI - 1, 2>(h, k, nsplit, rps, st);
  else if (sh.nie == NI && sh.tail == 2) launch_exchange_bf16_k<NP, NI, 2>(h, k, nsplit, rps, st);
  else if (sh.nie == NI && sh.tail == 3) launch_exchange_bf16_k<NP, NI, 3>(h, k, nsplit, rps, st);
  else { set_error("internal: bf16 shape"); return MBPT_ERR_VALUE; }
  return MBPT_OK;
}

template <typename T, int KT>
void launch_hartree_u(int M, int k, int chunk, int nqu, const T *p, const T *V, T *u,
                      hipStream_t st) {
  k_hartree_u<T, KT><<<dim3((M + 255) / 256, nqu), 256, 0, st>>>(M, k, chunk, p, V, u);
}

template <typename T>
int apply_impl(mbpt_sigma *h, int k, const double *z, double *Az, double *Bz, hipStream_t st) {
  const int M = h->M, no = h->no, nv = h->nv;
  const ApplyPlan plan = plan_for(h, k);
  const int nsplit = plan.nsplit, rows = k * no, ldr = plan.ldr;
  const int nq = plan.nq, kchunk = plan.kchunk, nqu = plan.nqu, hchunk = plan.hchunk;
  const bool km = sizeof(T) == 4;   // FP32 working precision: K-major head and tail
  const bool b3 = h->prec == MBPT_BF16X3;   // split-BF16 head and tail GEMMs
  int rc = ensure_ws(h, k, plan, Bz != nullptr);
  if (rc) return rc;
  T *zXv = (T *)h->ws[WS_ZXV], *p = (T *)h->ws[WS_P], *u = (T *)h->ws[WS_U];
  T *Tpart = (T *)h->ws[WS_TPART], *Tc = (T *)h->ws[WS_TC], *Ep = (T *)h->ws[WS_EP];
  // rows per split, a multiple of the 64-row LDS stage
  const int rps = (((M + nsplit - 1) / nsplit) + 63) / 64 * 64;
  const int nsplit_eff = (M + rps - 1) / rps;
  hipEvent_t *ev = nullptr;   // this apply's five timing events
  if (h->prof) {
    if ((rc = prof_events(h, &ev))) return rc;
    MBPT_HIP_TRY(hipEventRecord(ev[0], st));
  }

  if (b3) {  // head, split BF16: z as hi / lo rows, times X_v^T
    __bf16 *Zh = (__bf16 *)h->ws[WS_ZT], *Zl = Zh + (size_t)ldr * h->nvp32;
    k_split_rows_bf16<double, false><<<grid1((long)ldr * h->nvp32), 256, 0, st>>>(
        rows, nv, z, nv, ldr, h->nvp32, Zh, Zl);
    if ((rc = launch_status("k_split_rows_bf16(z)"))) return rc;
    const int ntm = ldr / 128, ntn = h->Mq / 128;
    const int nblk = pad_grid((long)ntm * ntn);
    k_gemm_nt3<EpiStore><<<nblk, 256, 0, st>>>(h->nvp32, h->nvp32, Zh, Zl, h->nvp32, h->XvB[0],
                                                h->XvB[1], h->nvp32, ntm, ntn, 1,
                                                EpiStore{(float *)zXv, rows, M, M}, nblk);
    if ((rc = launch_status("k_gemm_nt3(head)"))) return rc;
  } else if (km) {  // head, K-major: Z^T padded, then (Z^T)^T X_v^T
    float *Zt = (float *)h->ws[WS_ZT];
    k_transpose_pad<double><<<dim3(ldr / 64, (h->nvp + 63) / 64), 256, 0, st>>>(
        rows, nv, z, nv, Zt, ldr, ldr, h->nvp);
    if ((rc = launch_status("k_transpose_pad(z)"))) return rc;
    const int ntm = ldr / 128, ntn = (M + 127) / 128;
    const int nblk = pad_grid((long)ntm * ntn);
    k_gemm_km<EpiStore><<<nblk, 512, 0, st>>>(h->nvp, h->nvp, Zt, ldr, h->XvT, h->Mq, ntm,
                                               ntn, 1, EpiStore{(float *)zXv, rows, M, M},
                                               nblk);
    if ((rc = launch_status("k_gemm_km(head)"))) return rc;
  } else {  // head
    const int ntm = (rows + 63) / 64, ntn = (M + 63) / 64;
    const int nblk = pad_grid((long)ntm * ntn);
    k_head<T><<<nblk, 256, 0, st>>>(rows, M, nv, z, (const T *)h->Xv, zXv, ntm, ntn,
                                    nblk);
    if ((rc = launch_status("k_head"))) return rc;
  }
  if (ev) MBPT_HIP_TRY(hipEventRecord(ev[1], st));
  if (h->kappa != 0.0) {  // hartree
    k_hartree_p<T><<<grid1((long)k * M), 256, 0, st>>>(M, no, k, (const T *)h->XoT,
                                                        zXv, p);
    if ((rc = launch_status("k_hartree_p"))) return rc;
    if (k <= 16) launch_hartree_u<T, 1>(M, k, hchunk, nqu, p, (const T *)h->V, u, st);
    else if (k <= 32) launch_hartree_u<T, 2>(M, k, hchunk, nqu, p, (const T *)h->V, u, st);
    else if (k <= 48 || sizeof(T) == 8)   // FP64: 48 vectors a pass keeps two waves per SIMD
      launch_hartree_u<T, 3>(M, k, hchunk, nqu, p, (const T *)h->V, u, st);
    else launch_hartree_u<T, 4>(M, k, hchunk, nqu, p, (const T *)h->V, u, st);
    if ((rc = launch_status("k_hartree_u"))) return rc;
    if (nqu > 1) {   // one sum of the chunk partials, read once per (vector, point) below
      k_sum_partials<T><<<grid1((long)k * M), 256, 0, st>>>((long)k * M, nqu, u);
      if ((rc = launch_status("k_sum_partials"))) return rc;
    }
  }
  if (ev) MBPT_HIP_TRY(hipEventRecord(ev[2], st));
  if (h->prec == MBPT_BF16X3) {  // exchange, split BF16
    switch (h->nopad) {
      case 32: rc = launch_exchange_bf16<32>(h, k, nsplit_eff, rps, st); break;
      case 64: rc = launch_exchange_bf16<64>(h, k, nsplit_eff, rps, st); break;
      case 96: rc = launch_exchange_bf16<96>(h, k, nsplit_eff, rps, st); break;
      case 128: rc = launch_exchange_bf16<128>(h, k, nsplit_eff, rps, st); break;
      default: set_error("internal: nopad"); return MBPT_ERR_VALUE;
    }
    if (rc) return rc;
  } else switch (h->nopad) {  // exchange, FP32 / FP64
    case 16: launch_exchange<T, 16>(h, k, nsplit_eff, rps, st); break;
    case 32: launch_exchange<T, 32>(h, k, nsplit_eff, rps, st); break;
    case 48: launch_exchange<T, 48>(h, k, nsplit_eff, rps, st); break;
    case 64: launch_exchange<T, 64>(h, k, nsplit_eff, rps, st); break;
    case 80: launch_exchange<T, 80>(h, k, nsplit_eff, rps, st); break;
    case 96: launch_exchange<T, 96>(h, k, nsplit_eff, rps, st); break;
    case 112: launch_exchange<T, 112>(h, k, nsplit_eff, rps, st); break;
    case 128: launch_exchange<T, 128>(h, k, nsplit_eff, rps, st); break;
    default: set_error("internal: nopad"); return MBPT_ERR_VALUE;
  }
  if ((rc = launch_status("k_exchange"))) return rc;
  T *Upart = (T *)h->ws[WS_UPART];
  if (Bz && h->prec == MBPT_BF16X3) {  // the B block's exchange partials, split BF16
    switch (h->nopad) {
      case 32: rc = launch_exchange_b_bf16<32>(h, k, nsplit_eff, rps, (float *)Upart, st); break;
      case 64: rc = launch_exchange_b_bf16<64>(h, k, nsplit_eff, rps, (float *)Upart, st); break;
      case 96: rc = launch_exchange_b_bf16<96>(h, k, nsplit_eff, rps, (float *)Upart, st); break;
      case 128: rc = launch_exchange_b_bf16<128>(h, k, nsplit_eff, rps, (float *)Upart, st); break;
      default: set_error("internal: nopad"); return MBPT_ERR_VALUE;
    }
    if (rc) return rc;
  } else if (Bz) {  // the B block's exchange partials, FP32 / FP64
    switch (h->nopad) {
      case 16: launch_exchange_b<T, 16>(h, k, nsplit_eff, rps, Upart, st); break;
      case 32: launch_exchange_b<T, 32>(h, k, nsplit_eff, rps, Upart, st); break;
      case 48: launch_exchange_b<T, 48>(h, k, nsplit_eff, rps, Upart, st); break;
      case 64: launch_exchange_b<T, 64>(h, k, nsplit_eff, rps, Upart, st); break;
      case 80: launch_exchange_b<T, 80>(h, k, nsplit_eff, rps, Upart, st); break;
      case 96: launch_exchange_b<T, 96>(h, k, nsplit_eff, rps, Upart, st); break;
      case 112: launch_exchange_b<T, 112>(h, k, nsplit_eff, rps, Upart, st); break;
      case 128: launch_exchange_b<T, 128>(h, k, nsplit_eff, rps, Upart, st); break;
      default: set_error("internal: nopad"); return MBPT_ERR_VALUE;
    }
    if ((rc = launch_status("k_exchange_b"))) return rc;
  }
  if (ev) MBPT_HIP_TRY(hipEventRecord(ev[3], st));
  // tail: out = dd o z - (sum_s partials_s - kappa X_o^T o u) X_v (dd NULL: no
  // diagonal term, the B block)
  auto tail = [&](const T *src, const double *dd, double *out) -> int {
    if (b3) {  // tail, split BF16: T' as hi / lo rows, times X_v, split over K
      __bf16 *Th = (__bf16 *)h->ws[WS_ZT], *Tl = Th + (size_t)ldr * h->Mp32;
      k_tail_prep_b3<<<grid1((long)ldr * h->Mp32), 256, 0, st>>>(
          k, no, h->nopad, M, h->Mp32, ldr, nsplit_eff, (const float *)src,
          (const float *)h->XoT, h->kappa != 0.0 ? (const float *)u : nullptr,
          (float)h->kappa, Th, Tl);
      if ((rc = launch_status("k_tail_prep_b3"))) return rc;
      const int ntm = ldr / 128, ntn = h->nvq / 128;
      const int nblk = pad_grid((long)ntm * ntn * nq);
      k_gemm_nt3<EpiStore><<<nblk, 256, 0, st>>>(h->Mp32, kchunk, Th, Tl, h->Mp32, h->XvTB[0],
                                                  h->XvTB[1], h->Mp32, ntm, ntn, nq,
                                                  EpiStore{(float *)Ep, rows, nv, nv}, nblk);
      if ((rc = launch_status("k_gemm_nt3(tail)"))) return rc;
    } else if (km) {  // tail, K-major: T'^T padded, then (T'^T)^T X_v split over K
      float *Tct = (float *)h->ws[WS_ZT];
      k_tail_prep_t<<<dim3(ldr / 64, (h->Mp + 63) / 64), 256, 0, st>>>(
          k, no, h->nopad, M, h->Mp, ldr, nsplit_eff, (const float *)src,
          (const float *)h->XoT, h->kappa != 0.0 ? (const float *)u : nullptr, 1,
          (float)h->kappa, Tct);
      if ((rc = launch_status("k_tail_prep_t"))) return rc;
      const int ntm = ldr / 128, ntn = (nv + 127) / 128;
      const int nblk = pad_grid((long)ntm * ntn * nq);
      k_gemm_km<EpiStore><<<nblk, 512, 0, st>>>(h->Mp, kchunk, Tct, ldr, h->XvP, h->nvq, ntm,
                                                 ntn, nq, EpiStore{(float *)Ep, rows, nv, nv},
                                                 nblk);
      if ((rc = launch_status("k_gemm_km(tail)"))) return rc;
    } else {  // tail
      k_tail_prep<T><<<grid1((long)rows * M), 256, 0, st>>>(
          k, no, h->nopad, M, nsplit_eff, src, (const T *)h->XoT,
          h->kappa != 0.0 ? u : nullptr, 1, (T)h->kappa, Tc);
      if ((rc = launch_status("k_tail_prep"))) return rc;
      const int ntm = (rows + 63) / 64, ntn = (nv + 63) / 64;
      const int nblk = pad_grid((long)ntm * ntn * nq);
      k_gemm_nn_splitk<T><<<nblk, 256, 0, st>>>(rows, nv, M, kchunk, Tc, M,
                                                (const T *)h->Xv, nv, Ep, ntm, ntn, nq, nblk);
      if ((rc = launch_status("k_gemm_nn_splitk(tail)"))) return rc;
    }
    k_tail_finish<T><<<grid1((long)rows * nv), 256, 0, st>>>(rows, no, nv, nq, Ep, dd, z,
                                                              out);
    if ((rc = launch_status("k_tail_finish"))) return rc;
    return MBPT_OK;
  };
  if ((rc = tail(Tpart, h->d, Az))) return rc;
  if (ev) {
    MBPT_HIP_TRY(hipEventRecord(ev[4], st));
    h->ev_flops.push_back((double)k * (4.0 * no * (double)M * M + (double)M * M));
  }
  if (Bz && (rc = tail(Upart, nullptr, Bz))) return rc;
  return MBPT_OK;
}

template <typename T> int sdiag_impl(mbpt_sigma *h, double *out, hipStream_t st) {
  const int M = h->M, no = h->no, nv = h->nv;
  void *Q = nullptr;
  MBPT_HIP_TRY(hipMalloc(&Q, sizeof(T) * (size_t)M * nv));
  int rc;
  {
    const int ntm = (M + 63) / 64, ntn = (nv + 63) / 64;
    const int nblk = pad_grid((long)ntm * ntn);
    k_wsq<T><<<nblk, 256, 0, st>>>(M, nv, (const T *)h->W, (const T *)h->Xv, (T *)Q,
                                   ntm, ntn, nblk);
    rc = launch_status("k_wsq");
  }
  if (!rc) {
    const int ntm = (no + 63) / 64, ntn = (nv + 63) / 64;
    const int nblk = pad_grid((long)ntm * ntn);
    k_sdiag<T><<<nblk, 256, 0, st>>>(M, no, nv, (const T *)h->XoT, (const T *)Q, h->d,
                                     out, ntm, ntn, nblk);
    rc = launch_status("k_sdiag");
  }
  hipError_t e = hipStreamSynchronize(st);
  (void)hipFree(Q);
  if (!rc && e != hipSuccess) {
    set_error(std::string("screened diagonal: ") + hipGetErrorString(e));
    rc = MBPT_ERR_HIP;
  }
  return rc;
}

}  // namespace

extern "C" {

int mbpt_sigma_create(mbpt_sigma **out, int M, int nocc, int nvir, int precision,
                      double kappa, const double *X_dev, const double *W_dev,
                      const double *V_dev, const double *eps_host, void *stream) {
  MBPT_REQUIRE(out != nullptr, "mbpt_sigma_create: out is NULL");
  *out = nullptr;
  MBPT_REQUIRE(M > 0 && nocc > 0 && nvir > 0, "mbpt_sigma_create: M, nocc and nvir must be positive");
  MBPT_REQUIRE(nocc <= kMaxNocc, "mbpt_sigma_create: nocc > 128 is not supported by the exchange kernel");
  MBPT_REQUIRE(X_dev && W_dev && eps_host, "mbpt_sigma_create: NULL X, W or eps");
  MBPT_REQUIRE(kappa == 0.0 || V_dev, "mbpt_sigma_create: kappa != 0 needs V");
  if (precision != MBPT_FP64 && precision != MBPT_FP32 && precision != MBPT_BF16X3) {
    set_error("mbpt_sigma_create: precision " + std::to_string(precision) +
              " is not built (MBPT_FP64, MBPT_FP32 or MBPT_BF16X3)");
    return MBPT_ERR_UNSUPPORTED;
  }
  hipStream_t st = (hipStream_t)stream;
  mbpt_sigma *h = new mbpt_sigma();
  h->M = M; h->no = nocc; h->nv = nvir;
  // exchange K width: 16-multiples for the 16x16x4 MFMAs, 32 for 16x16x32 BF16
  h->nopad = precision == MBPT_BF16X3 ? (nocc + 31) / 32 * 32 : (nocc + 15) / 16 * 16;
  h->prec = precision; h->kappa = kappa;
  h->esz = precision == MBPT_FP64 ? 8 : 4;
  const size_t e = h->esz;
  auto fail = [&](int rc) { mbpt_sigma_destroy(h); return rc; };
#define TRY_ALLOC(ptr, bytes)                                                   \
  do {                                                                          \
    hipError_t _e = hipMalloc(&(ptr), (bytes));                                 \
    if (_e != hipSuccess) {                                                     \
      set_error(std::string("hipMalloc: ") + hipGetErrorString(_e));            \
      return fail(MBPT_ERR_HIP);                                                \
    }                                                                           \
  } while (0)
  TRY_ALLOC(h->Xo, e * M * nocc);
  TRY_ALLOC(h->Xv, e * M * nvir);
  TRY_ALLOC(h->XoT, e * M * nocc);
  TRY_ALLOC(h->W, e * M * M);
  if (kappa != 0.0) TRY_ALLOC(h->V, e * M * M);
  TRY_ALLOC(h->d, sizeof(double) * nocc * nvir);
#undef TRY_ALLOC
  h->held = e * ((size_t)M * (2 * nocc + nvir) + (size_t)M * M * (kappa != 0.0 ? 2 : 1)) +
            sizeof(double) * nocc * nvir;
  std::vector<double> dh((size_t)nocc * nvir);
  for (int i = 0; i < nocc; ++i)
    for (int a = 0; a < nvir; ++a) dh[(size_t)i * nvir + a] = eps_host[nocc + a] - eps_host[i];
  hipError_t he = hipMemcpyAsync(h->d, dh.data(), sizeof(double) * dh.size(),
                                 hipMemcpyHostToDevice, st);
  if (he == hipSuccess) he = hipStreamSynchronize(st);  // dh leaves scope
  if (he != hipSuccess) {
    set_error(std::string("upload of d: ") + hipGetErrorString(he));
    return fail(MBPT_ERR_HIP);
  }
  int rc = precision == MBPT_FP64 ? create_impl<double>(h, X_dev, W_dev, V_dev, st)
                                  : create_impl<float>(h, X_dev, W_dev, V_dev, st);
  if (rc) return fail(rc);
  if (precision != MBPT_FP64) {  // K-major, zero-padded X_v copies for head and tail
    h->nvp = (nvir + 15) / 16 * 16;
    h->Mq = (M + 127) / 128 * 128;
    h->Mp = (M + 15) / 16 * 16;
    h->nvq = (nvir + 127) / 128 * 128;
    h->nvp32 = (nvir + 31) / 32 * 32;
    h->Mp32 = (M + 31) / 32 * 32;
  }
  if (precision == MBPT_BF16X3) {  // split-BF16 X_v (head) and X_v^T (tail), hi / lo rows
    const size_t nb = (size_t)h->Mq * h->nvp32, nt = (size_t)h->nvq * h->Mp32;
    for (int hl = 0; hl < 2 && he == hipSuccess; ++hl) {
      he = hipMalloc((void **)&h->XvB[hl], sizeof(__bf16) * nb);
      if (he == hipSuccess) he = hipMalloc((void **)&h->XvTB[hl], sizeof(__bf16) * nt);
    }
    if (he != hipSuccess) {
      set_error(std::string("hipMalloc: ") + hipGetErrorString(he));
      return fail(MBPT_ERR_HIP);
    }
    h->held += sizeof(__bf16) * 2 * (nb + nt);
    k_split_rows_bf16<float, false><<<grid1((long)nb), 256, 0, st>>>(
        M, nvir, (const float *)h->Xv, nvir, h->Mq, h->nvp32, h->XvB[0], h->XvB[1]);
    k_split_rows_bf16<float, true><<<grid1((long)nt), 256, 0, st>>>(
        nvir, M, (const float *)h->Xv, nvir, h->nvq, h->Mp32, h->XvTB[0], h->XvTB[1]);
    if ((rc = launch_status("split-BF16 X_v copies"))) return fail(rc);
  } else if (precision != MBPT_FP64) {
    he = hipMalloc((void **)&h->XvT, sizeof(float) * (size_t)h->nvp * h->Mq);
    if (he == hipSuccess) he = hipMalloc((void **)&h->XvP, sizeof(float) * (size_t)h->Mp * h->nvq);
    if (he != hipSuccess) {
      set_error(std::string("hipMalloc: ") + hipGetErrorString(he));
      return fail(MBPT_ERR_HIP);
    }
    h->held += sizeof(float) * ((size_t)h->nvp * h->Mq + (size_t)h->Mp * h->nvq);
    k_transpose_pad<float><<<dim3(h->Mq / 64, (h->nvp + 63) / 64), 256, 0, st>>>(
        M, nvir, (const float *)h->Xv, nvir, h->XvT, h->Mq, h->Mq, h->nvp);
    k_pad_copy<<<grid1((long)h->Mp * h->nvq), 256, 0, st>>>(M, nvir, (const float *)h->Xv,
                                                            h->Mp, h->nvq, h->XvP);
    if ((rc = launch_status("X_v padded copies"))) return fail(rc);
  }
  {  // the exchange's zero-padded X_o (the B block uses it in every precision)
    const int Mp32 = (M + 31) / 32 * 32;
    const size_t bytes = e * (size_t)Mp32 * h->nopad;
    he = hipMalloc(&h->Xp, bytes);
    if (he != hipSuccess) {
      set_error(std::string("hipMalloc: ") + hipGetErrorString(he));
      return fail(MBPT_ERR_HIP);
    }
    h->held += bytes;
    if (precision == MBPT_FP64)
      k_pack_xo<double><<<grid1((long)Mp32 * h->nopad), 256, 0, st>>>(
          M, nocc, Mp32, h->nopad, (const double *)h->Xo, (double *)h->Xp);
    else
      k_pack_xo<float><<<grid1((long)Mp32 * h->nopad), 256, 0, st>>>(
          M, nocc, Mp32, h->nopad, (const float *)h->Xo, (float *)h->Xp);
    if ((rc = launch_status("k_pack_xo"))) return fail(rc);
  }
  if (precision == MBPT_BF16X3) {
    const long slots = (long)((M + 31) / 32) * bf16_shape(nocc).spc;
    he = hipMalloc((void **)&h->frag, sizeof(bf16x8) * slots);
    if (he != hipSuccess) {
      set_error(std::string("hipMalloc: ") + hipGetErrorString(he));
      return fail(MBPT_ERR_HIP);
    }
    h->held += sizeof(bf16x8) * slots;
    k_pack_bf16_xo<<<grid1(slots), 256, 0, st>>>(M, nocc, (const float *)h->Xo, h->frag);
    if ((rc = launch_status("k_pack_bf16_xo"))) return fail(rc);
    const long wslots = (long)((M + 31) / 32) * w_tiles_per_chunk(M) * 512;
    he = hipMalloc((void **)&h->wtile, sizeof(f32x4) * wslots);
    if (he != hipSuccess) {
      set_error(std::string("hipMalloc: ") + hipGetErrorString(he));
      return fail(MBPT_ERR_HIP);
    }
    h->held += sizeof(f32x4) * wslots;
    k_pack_w_tiles<<<grid1(wslots), 256, 0, st>>>(M, w_tiles_per_chunk(M), (const float *)h->W,
                                                   h->wtile);
    if ((rc = launch_status("k_pack_w_tiles"))) return fail(rc);
  }
  *out = h;
  return MBPT_OK;
}

int mbpt_sigma_apply(mbpt_sigma *h, int k, const double *z_dev, double *Az_dev,
                     void *stream) {
  MBPT_REQUIRE(h != nullptr, "mbpt_sigma_apply: NULL handle");
  MBPT_REQUIRE(k >= 0, "mbpt_sigma_apply: k < 0");
  if (k == 0) return MBPT_OK;
  MBPT_REQUIRE(z_dev && Az_dev, "mbpt_sigma_apply: NULL z or Az");
  MBPT_REQUIRE(z_dev != Az_dev, "mbpt_sigma_apply: z and Az must be distinct buffers");
  hipStream_t st = (hipStream_t)stream;
  return h->prec == MBPT_FP64 ? apply_impl<double>(h, k, z_dev, Az_dev, nullptr, st)
                              : apply_impl<float>(h, k, z_dev, Az_dev, nullptr, st);
}

int mbpt_sigma_apply_ab(mbpt_sigma *h, int k, const double *z_dev, double *Az_dev,
                        double *Bz_dev, void *stream) {
  MBPT_REQUIRE(h != nullptr, "mbpt_sigma_apply_ab: NULL handle");
  MBPT_REQUIRE(k >= 0, "mbpt_sigma_apply_ab: k < 0");
  if (k == 0) return MBPT_OK;
  MBPT_REQUIRE(z_dev && Az_dev && Bz_dev, "mbpt_sigma_apply_ab: NULL z, Az or Bz");
  MBPT_REQUIRE(z_dev != Az_dev && z_dev != Bz_dev && Az_dev != Bz_dev,
               "mbpt_sigma_apply_ab: z, Az and Bz must be distinct buffers");
  hipStream_t st = (hipStream_t)stream;
  return h->prec == MBPT_FP64 ? apply_impl<double>(h, k, z_dev, Az_dev, Bz_dev, st)
                              : apply_impl<float>(h, k, z_dev, Az_dev, Bz_dev, st);
}

int mbpt_sigma_screened_diag(mbpt_sigma *h, double *out_dev, void *stream) {
  MBPT_REQUIRE(h != nullptr && out_dev != nullptr, "mbpt_sigma_screened_diag: NULL argument");
  hipStream_t st = (hipStream_t)stream;
  return h->prec == MBPT_FP64 ? sdiag_impl<double>(h, out_dev, st)
                              : sdiag_impl<float>(h, out_dev, st);
}

size_t mbpt_sigma_workspace_bytes(const mbpt_sigma *h, int k, int with_b) {
  if (!h || k <= 0) return 0;
  size_t sizes[WS_N];
  ws_sizes(h, k, plan_for(h, k), with_b != 0, sizes);
  size_t total = 0;
  for (int q = 0; q < WS_N; ++q) total += sizes[q];
  return total;
}

size_t mbpt_sigma_held_bytes(const mbpt_sigma *h) {
  if (!h) return 0;
  size_t ws = 0;
  for (int q = 0; q < WS_N; ++q) ws += h->ws_cap[q];
  return h->held + ws;
}

int mbpt_sigma_profile(mbpt_sigma *h, int enable) {
  MBPT_REQUIRE(h != nullptr, "mbpt_sigma_profile: NULL handle");
  h->prof = enable ? 1 : 0;
  h->ev_flops.clear();          // pending applies are dropped with the figures
  for (int q = 0; q < 4; ++q) h->prof_ms[q] = 0.0;
  h->prof_flops = 0.0;
  h->prof_launches = 0;
  return MBPT_OK;
}

int mbpt_sigma_profile_read(mbpt_sigma *h, double *ms_out, long *launches,
                            double *exchange_flops) {
  MBPT_REQUIRE(h != nullptr, "mbpt_sigma_profile_read: NULL handle");
  int rc = prof_collect(h);
  if (rc) return rc;
  if (ms_out)
    for (int q = 0; q < 4; ++q) ms_out[q] = h->prof_ms[q];
  if (launches) *launches = h->prof_launches;
  if (exchange_flops) *exchange_flops = h->prof_flops;
  return MBPT_OK;
}

int mbpt_sigma_destroy(mbpt_sigma *h) {
  if (!h) return MBPT_OK;
  for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
  free_ws(h);
  (void)hipFree(h->Xo); (void)hipFree(h->Xp); (void)hipFree(h->Xv); (void)hipFree(h->XoT); (void)hipFree(h->W); (void)hipFree(h->V);
  (void)hipFree(h->frag); (void)hipFree(h->wtile); (void)hipFree(h->XvT); (void)hipFree(h->XvP);
  for (int hl = 0; hl < 2; ++hl) { (void)hipFree(h->XvB[hl]); (void)hipFree(h->XvTB[hl]); }
  (void)hipFree(h->d);
  delete h;
  return MBPT_OK;
}

}  // extern "C"
