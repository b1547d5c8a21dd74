// Batched sigma-vectors of many independent ISDF instances of uneven size
// (n_occ, n_vir, M differ per instance) on gfx950.
//
// Every stage of the single-instance action (sigma.hip) runs as ONE launch
// over all instances: head z X_v^T, Hartree p and u = V p, the fused exchange
// (one launch per 16-wide occupied bucket, the width being a template
// argument), tail prep, the tail's K-chunk products and the FP64 finish
// A z = d o z - T' X_v.  Each stage has a work table -- the prefix sum over
// instances of that instance's work items (64x64 tiles, K chunks, 256-element
// blocks) -- and each workgroup walks items lid, lid + G, ... (persistent: the
// grid G is a plan parameter, at most the item count), finding the instance
// of an item by binary search.  The launch sequence and its arguments depend
// only on the handle, so a caller captures it once in a hipGraph and replays
// it every Davidson cycle; what changes per cycle (the trial vectors and the
// per-instance count of live vectors) is read from device memory.
//
// Vectors use a padded batch layout z[b][v][pair] with pair < npair_max
// (instance b's n_occ n_vir pairs first, occupied slow); the padding is
// never read and never written.
#include "common.h"

#include <algorithm>
#include <vector>

namespace mbpt {
namespace {

struct Inst {
  int id;                      // position in the caller's instance order
  int M, no, nv, nopad;        // nopad: n_o padded to the plan's exchange bucket width
  int ldxo;                    // row stride of the pool's X_o (n_o padded to 32)
  long xo, xv, xot, w, v;      // element offsets into the operand pool
  long d;                      // offset into the pair energies (FP64)
  long zxv, p, u, tp, tc, ep;  // element offsets into the workspace
  int nsplit, rps, nq, kch, nqu, kchu;
};

typedef __attribute__((address_space(3))) void lds_ptr;

struct Tables {
  const Inst *inst;
  const int *pre;   // n + 1 prefix sums of the work items
  int n, total;
};

__device__ __forceinline__ int find_inst(const int *pre, int n, int item) {
  int lo = 0, hi = n;  // pre[lo] <= item < pre[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (pre[mid] <= item) lo = mid; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ int live(const int *kact, int b, int k) {
  return kact ? max(0, min(kact[b], k)) : k;
}

template <typename T> struct LdZb {  // A(m = v n_o + i, k = a) = z[v][i n_v + a]
  const double *z; long ldz; int no, nv, rows;
  __device__ T operator()(int m, int k) const {
    if (m >= rows || k >= nv) return T(0);
    const int v = m / no, i = m - v * no;
    return (T)z[(long)v * ldz + (long)i * nv + k];
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

template <typename T>
__device__ __forceinline__ void store_tile(T *out, int ld, int rows, int cols, int m0, int n0,
                                           typename MF<T>::acc (&acc)[2][2]) {
  for (int fm = 0; fm < 2; ++fm)
    for (int fn = 0; fn < 2; ++fn) {
      const int c = tile_col(n0, fn);
      for (int r = 0; r < 4; ++r) {
        const int m = tile_row<T>(m0, fm, r);
        if (m < rows && c < cols) out[(long)m * ld + c] = acc[fm][fn][r];
      }
    }
}

// ---- head: zXv[b] = z[b] X_v^T, items = 64x64 tiles of (k n_o) x M -----------------
template <typename T>
__global__ __launch_bounds__(256) void g_head(Tables t, int k, const int *__restrict__ kact,
                                              const double *__restrict__ z, long ldz,
                                              long zbatch, const T *__restrict__ pool,
                                              T *__restrict__ ws) {
  __shared__ TileSmem<T> sm;
  const int G = gridDim.x;
  for (int item = xcd_logical(blockIdx.x, G); item < t.total; item += G) {
    const int b = find_inst(t.pre, t.n, item);
    const Inst I = t.inst[b];
    const int rows = live(kact, I.id, k) * I.no, ntn = (I.M + 63) / 64;
    const int local = item - t.pre[b], tm = local / ntn, tn = local % ntn;
    if (tm * kTile >= rows) continue;
    typename MF<T>::acc acc[2][2];
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) zero(acc[i][j]);
    gemm64<T, true, true>(LdZb<T>{z + I.id * zbatch, ldz, I.no, I.nv, rows},
                          LdBnk<T>{pool + I.xv, I.nv, I.nv, I.M}, tm * kTile, tn * kTile,
                          I.nv, acc, sm);
    store_tile<T>(ws + I.zxv, I.M, rows, I.M, tm * kTile, tn * kTile, acc);
  }
}

// ---- Hartree density p[b]_mu = sum_j X_o[mu, j] zXv[b][j, mu], 256 per item ------------
template <typename T>
__global__ __launch_bounds__(256) void g_hartree_p(Tables t, int k, const int *__restrict__ kact,
                                                   const T *__restrict__ pool,
                                                   T *__restrict__ ws) {
  const int G = gridDim.x;
  for (int item = xcd_logical(blockIdx.x, G); item < t.total; item += G) {
    const int b = find_inst(t.pre, t.n, item);
    const Inst I = t.inst[b];
    const long e = (long)(item - t.pre[b]) * 256 + threadIdx.x;
    if (e >= (long)live(kact, I.id, k) * I.M) continue;
    const int v = (int)(e / I.M), mu = (int)(e % I.M);
    const T *XoT = pool + I.xot, *zXv = ws + I.zxv;
    T s = 0;
    for (int j = 0; j < I.no; ++j)
      s += XoT[(long)j * I.M + mu] * zXv[((long)v * I.no + j) * I.M + mu];
    ws[I.p + e] = s;
  }
}

// ---- Hartree u_q[b] = sum_{mu in chunk q} p[b][mu] V[mu][:], V streamed once -------------
// Items: (256-column block, row chunk of kchu rows).  Each thread owns one
// column; the live vectors' p values of 64 rows at a time sit in LDS; 16 rows
// of V are in flight per thread before the FMAs that use them.
template <typename T, int KV>
__global__ __launch_bounds__(256) void g_hartree_u(Tables t, int k, const int *__restrict__ kact,
                                                   const T *__restrict__ pool,
                                                   T *__restrict__ ws) {
  __shared__ T ps[KV][64];
  const int G = gridDim.x;
  for (int item = xcd_logical(blockIdx.x, G); item < t.total; item += G) {
    const int b = find_inst(t.pre, t.n, item);
    const Inst I = t.inst[b];
    const int M = I.M, kl = live(kact, I.id, k), ncb = (M + 255) / 256;
    const int local = item - t.pre[b], cb = local % ncb, q = local / ncb;
    const int nu = cb * 256 + threadIdx.x;
    const int mu_beg = q * I.kchu, mu_end = min(M, mu_beg + I.kchu);
    const T *p = ws + I.p, *V = pool + I.v;
    T *u = ws + I.u;
    for (int v0 = 0; v0 < kl; v0 += KV) {
      const int nvec = min(KV, kl - v0);
      T acc[KV];
#pragma unroll
      for (int v = 0; v < KV; ++v) acc[v] = 0;
      for (int mg = mu_beg; mg < mu_end; mg += 64) {
        __syncthreads();
        for (int e = threadIdx.x; e < KV * 64; e += 256) {
          const int v = e >> 6, r = e & 63, mu = mg + r;
          ps[v][r] = (v < nvec && mu < mu_end) ? p[(long)(v0 + v) * M + mu] : T(0);
        }
        __syncthreads();
        if (nu < M)
          for (int r0 = 0; r0 < 64 && mg + r0 < mu_end; r0 += 16) {
            T x[16];
#pragma unroll
            for (int qq = 0; qq < 16; ++qq) {
              const int mu = mg + r0 + qq;
              x[qq] = mu < mu_end ? V[(long)mu * M + nu] : T(0);
            }
#pragma unroll
            for (int qq = 0; qq < 16; ++qq)
#pragma unroll
              for (int v = 0; v < KV; ++v) acc[v] += ps[v][r0 + qq] * x[qq];
          }
      }
      if (nu < M)
        for (int v = 0; v < nvec; ++v) u[((long)q * kl + v0 + v) * M + nu] = acc[v];
    }
  }
}

// ---- split-K NN products: Hartree u_q = p V (WHICH 0), tail E_q = T' X_v (WHICH 1) ----
// Items: 64x64 output tiles times K chunks.  The chunk partials are laid out
// [q][rows][cols] with rows = live vectors (x n_o for the tail), which is how
// the consumers read them.
template <typename T, int WHICH>
__global__ __launch_bounds__(256) void g_nn(Tables t, int k, const int *__restrict__ kact,
                                            const T *__restrict__ pool, T *__restrict__ ws) {
  __shared__ TileSmem<T> sm;
  const int G = gridDim.x;
  for (int item = xcd_logical(blockIdx.x, G); item < t.total; item += G) {
    const int b = find_inst(t.pre, t.n, item);
    const Inst I = t.inst[b];
    const int kl = live(kact, I.id, k);
    int rows, cols, rows_plan, kchunk, ldb;
    const T *A, *B;
    T *C;
    if (WHICH == 0) {
      rows = kl; rows_plan = k; cols = I.M; kchunk = I.kchu;
      A = ws + I.p; B = pool + I.v; ldb = I.M; C = ws + I.u;
    } else {
      rows = kl * I.no; rows_plan = k * I.no; cols = I.nv; kchunk = I.kch;
      A = ws + I.tc; B = pool + I.xv; ldb = I.nv; C = ws + I.ep;
    }
    const int ntm = (rows_plan + 63) / 64, ntn = (cols + 63) / 64;
    const int local = item - t.pre[b];
    const int tn = local % ntn, tm = (local / ntn) % ntm, q = local / (ntn * ntm);
    if (tm * kTile >= rows) continue;
    const int k0 = q * kchunk, kc = min(kchunk, I.M - k0);
    typename MF<T>::acc acc[2][2];
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) zero(acc[i][j]);
    gemm64<T, true, false>(LdArm<T>{A + k0, I.M, rows, kc},
                           LdBkn<T>{B + (long)k0 * ldb, ldb, kc, cols}, tm * kTile,
                           tn * kTile, kc, acc, sm);
    store_tile<T>(C + (long)q * rows * cols, cols, rows, cols, tm * kTile, tn * kTile, acc);
  }
}

// ---- exchange T_s[b] = X_o^T (W o (X_o zXv[b])) over grid-row split s -------------------
// The fused two-MFMA scheme of the single-instance exchange: per 16 grid rows,
// P = X_o zXv (K = n_o) stays in registers, is scaled by W and is the B
// operand of the second product.  Items: (64 grid columns, vector, split),
// the vector fastest, so the k items sharing a W column block run together.
template <typename T, int NOPAD>
__global__ __launch_bounds__(256) void g_exchange(Tables t, int k, const int *__restrict__ kact,
                                                  const T *__restrict__ pool,
                                                  T *__restrict__ ws) {
  // X_o of an instance sits in the pool zero-padded to (M rounded to 32) x
  // ldxo (>= NOPAD); its first NOPAD columns stream through double-buffered
  // 32-row LDS stages by direct
  // 16-byte global-to-LDS loads (one instruction per row, a stage ahead); the
  // next 16 rows' W are prefetched.  Rows past M are zero rows (P = 0 there,
  // W read at a clamped row).
  constexpr int NS = NOPAD / 4, NI = NOPAD / 16, ROWS = 32;
  constexpr int LD = NOPAD + 16 / (int)sizeof(T), NL = NOPAD * (int)sizeof(T) / 16;
  __shared__ __attribute__((aligned(16))) T xs[2][ROWS * LD];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lr = lane & 15, lq = lane >> 4;
  const int G = gridDim.x;
  for (int item = xcd_logical(blockIdx.x, G); item < t.total; item += G) {
    const int b = find_inst(t.pre, t.n, item);
    const Inst I = t.inst[b];
    const int M = I.M, no = I.no, nvb = (M + 63) / 64;
    const int local = item - t.pre[b];
    const int v = local % k, tt = local / k, vb = tt % nvb, s = tt / nvb;
    if (v >= live(kact, I.id, k)) continue;
    const T *Xp = pool + I.xo, *W = pool + I.w, *zXv = ws + I.zxv;
    const int col = vb * 64 + wave * 16 + lr;
    const bool colok = col < M;
    const int colc = min(col, M - 1);
    T bz[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      const int j = 4 * q + lq;
      bz[q] = (j < no && colok) ? zXv[((long)v * no + j) * M + col] : T(0);
    }
    typename MF<T>::acc acc[NI];
#pragma unroll
    for (int it = 0; it < NI; ++it) zero(acc[it]);
    // rps is a multiple of 64: a stage never straddles two splits
    const int st_beg = (s * I.rps) / ROWS;
    const int st_end = min((M + ROWS - 1) / ROWS, ((s + 1) * I.rps) / ROWS);
    auto fetch = [&](int st, int n) {
#pragma unroll
      for (int q = 0; q < ROWS / 4; ++q) {
        const int r = 4 * q + wave;
        if (lane < NL)
          __builtin_amdgcn_global_load_lds(
              (const void *)((const char *)(Xp + (long)(ROWS * st + r) * I.ldxo) + 16 * lane),
              (lds_ptr *)&xs[n][r * LD], 16, 0, 0);
      }
    };
    auto load_w = [&](int mu0, T (&w)[4]) {
#pragma unroll
      for (int r = 0; r < 4; ++r)
        w[r] = W[(long)min(mu0 + MF<T>::orow(lq, r), M - 1) * M + colc];
    };
    if (st_beg < st_end) {
      __syncthreads();   // the previous item's last stage is no longer read
      fetch(st_beg, 0);
      T w[4];
      load_w(ROWS * st_beg, w);
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      int cur = 0;
      for (int st = st_beg; st < st_end; ++st) {
        if (st + 1 < st_end) fetch(st + 1, cur ^ 1);
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
              acc[it] = MF<T>::mma(x[(r0 + MF<T>::orow(lq, r)) * LD + 16 * it + lr], S[r],
                                   acc[it]);
#pragma unroll
          for (int r = 0; r < 4; ++r) w[r] = wn[r];
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        cur ^= 1;
      }
    }
    if (colok) {
      T *Tp = ws + I.tp;
#pragma unroll
      for (int it = 0; it < NI; ++it)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          Tp[(((long)s * k + v) * NOPAD + 16 * it + MF<T>::orow(lq, r)) * M + col] = acc[it][r];
    }
  }
}

// ---- tail prep: T' = sum_s T_s - kappa X_o^T o u, 256 elements of (k n_o) x M per item -----
template <typename T>
__global__ __launch_bounds__(256) void g_tail_prep(Tables t, int k, const int *__restrict__ kact,
                                                   T kappa, const T *__restrict__ pool,
                                                   T *__restrict__ ws) {
  const int G = gridDim.x;
  for (int item = xcd_logical(blockIdx.x, G); item < t.total; item += G) {
    const int b = find_inst(t.pre, t.n, item);
    const Inst I = t.inst[b];
    const int M = I.M, no = I.no, kl = live(kact, I.id, k);
    const long e = (long)(item - t.pre[b]) * 256 + threadIdx.x;
    if (e >= (long)kl * no * M) continue;
    const int m = (int)(e / M), mu = (int)(e % M);
    const int v = m / no, i = m - v * no;
    const T *Tp = ws + I.tp;
    const long sstride = (long)k * I.nopad * M;
    const long off = ((long)v * I.nopad + i) * M + mu;
    T s = Tp[off];
    for (int q = 1; q < I.nsplit; ++q) s += Tp[off + q * sstride];
    if (kappa != T(0)) {
      const T *u = ws + I.u;
      T ub = u[(long)v * M + mu];
      for (int q = 1; q < I.nqu; ++q) ub += u[((long)q * kl + v) * M + mu];
      s -= kappa * pool[I.xot + (long)i * M + mu] * ub;
    }
    ws[I.tc + e] = s;
  }
}

// ---- finish: Az = d o z - sum_q E_q in FP64, 256 elements of (k n_o) x n_v per item ----------
template <typename T>
__global__ __launch_bounds__(256) void g_tail_finish(Tables t, int k, const int *__restrict__ kact,
                                                     const T *__restrict__ ws,
                                                     const double *__restrict__ dpool,
                                                     const double *__restrict__ z,
                                                     double *__restrict__ Az, long ldz,
                                                     long zbatch) {
  const int G = gridDim.x;
  for (int item = xcd_logical(blockIdx.x, G); item < t.total; item += G) {
    const int b = find_inst(t.pre, t.n, item);
    const Inst I = t.inst[b];
    const int no = I.no, nv = I.nv;
    const long n = (long)live(kact, I.id, k) * no * nv;
    const long e = (long)(item - t.pre[b]) * 256 + threadIdx.x;
    if (e >= n) continue;
    const int m = (int)(e / nv), a = (int)(e % nv), v = m / no, i = m - v * no;
    const T *Ep = ws + I.ep;
    T s = Ep[e];
    for (int q = 1; q < I.nq; ++q) s += Ep[(long)q * n + e];
    const long zo = I.id * zbatch + (long)v * ldz + (long)i * nv + a;
    Az[zo] = dpool[I.d + (long)i * nv + a] * z[zo] - (double)s;
  }
}

// ---- screened diagonal: d - (X_o o X_o)^T W (X_v o X_v), two batched tile GEMMs ----------------
// Y_b = W_b (X_v o X_v) (M x n_v, K = M) into a scratch buffer, then
// out_b = d_b - (X_o o X_o)^T Y_b (n_o x n_v, K = M) into row b of the output.
template <typename T> struct LdBsqT {  // B(k, n) = P[k][n]^2
  const T *p; int ld, K, cols;
  __device__ T operator()(int k, int n) const {
    if (k >= K || n >= cols) return T(0);
    const T x = p[(long)k * ld + n];
    return x * x;
  }
};
template <typename T> struct LdAsqT {  // A(m, k) = P[m][k]^2
  const T *p; int ld, rows, K;
  __device__ T operator()(int m, int k) const {
    if (m >= rows || k >= K) return T(0);
    const T x = p[(long)m * ld + k];
    return x * x;
  }
};

template <typename T, int STAGE>
__global__ __launch_bounds__(256) void g_sdiag(Tables t, const T *__restrict__ pool,
                                               const long *__restrict__ yoff, T *__restrict__ Y,
                                               const double *__restrict__ dpool,
                                               double *__restrict__ out, long ldo) {
  __shared__ TileSmem<T> sm;
  const int G = gridDim.x;
  for (int item = xcd_logical(blockIdx.x, G); item < t.total; item += G) {
    const int b = find_inst(t.pre, t.n, item);
    const Inst I = t.inst[b];
    const int rows = STAGE == 0 ? I.M : I.no, cols = I.nv, ntn = (cols + 63) / 64;
    const int local = item - t.pre[b], tm = local / ntn, tn = local % ntn;
    typename MF<T>::acc acc[2][2];
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) zero(acc[i][j]);
    T *Yb = Y + yoff[I.id];
    if (STAGE == 0) {
      gemm64<T, true, false>(LdArm<T>{pool + I.w, I.M, I.M, I.M},
                             LdBsqT<T>{pool + I.xv, I.nv, I.M, I.nv}, tm * kTile, tn * kTile,
                             I.M, acc, sm);
      store_tile<T>(Yb, cols, rows, cols, tm * kTile, tn * kTile, acc);
    } else {
      gemm64<T, true, false>(LdAsqT<T>{pool + I.xot, I.M, I.no, I.M},
                             LdBkn<T>{Yb, I.nv, I.M, I.nv}, tm * kTile, tn * kTile, I.M, acc,
                             sm);
      for (int fm = 0; fm < 2; ++fm)
        for (int fn = 0; fn < 2; ++fn) {
          const int a = tile_col(tn * kTile, fn);
          for (int r = 0; r < 4; ++r) {
            const int i = tile_row<T>(tm * kTile, fm, r);
            if (i < rows && a < cols) {
              const long p = (long)i * I.nv + a;
              out[I.id * ldo + p] = dpool[I.d + p] - (double)acc[fm][fn][r];
            }
          }
        }
    }
  }
}

// ---- operand upload: X split into X_o, X_v, X_o^T; W and V converted ----------------------
template <typename T>
__global__ void g_split_X(int M, int no, int nv, int nopad, const double *__restrict__ X,
                          T *__restrict__ Xo, T *__restrict__ Xv, T *__restrict__ XoT) {
  const long n = (long)M * (no + nv);
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x) {
    const int mu = (int)(e / (no + nv)), p = (int)(e % (no + nv));
    const T x = (T)X[e];
    if (p < no) {
      Xo[(long)mu * nopad + p] = x;   // the padding stays zero (pool cleared at create)
      XoT[(long)p * M + mu] = x;
    } else {
      Xv[(long)mu * nv + (p - no)] = x;
    }
  }
}

template <typename T>
__global__ void g_convert(long n, const double *__restrict__ src, T *__restrict__ dst) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x)
    dst[e] = (T)src[e];
}

int grid1(long n) { return (int)std::min<long>((n + 255) / 256, 16384L); }

enum { S_HEAD, S_HP, S_HU, S_TPREP, S_TAIL, S_FIN, S_X16, S_X32, S_X48, S_X64, S_N };
constexpr int kBuckets = 4;
constexpr int kStages = 6;   // stages before the exchange buckets

}  // namespace
}  // namespace mbpt

using namespace mbpt;

struct mbpt_batch {
  int n, kmax, npmax, prec;
  double kappa;
  size_t esz;
  std::vector<Inst> hinst;
  long pool_elems;
  void *pool;           // X_o, X_v, X_o^T, W, V of every instance, working precision
  double *dpool;        // pair energies
  // plan
  int planned;
  int rows_target, tail_kchunk, hartree_kchunk, grid_cap, bucket_width;
  Inst *dinst;          // instances in input order
  Inst *dinst_bucket;   // instances regrouped by occupied bucket (exchange)
  int *pre;             // S_N x (n + 1) prefix sums
  int total[S_N], nb[kBuckets], boff[kBuckets];
  void *ws;
  size_t ws_bytes;
};

namespace {

void free_plan(mbpt_batch *h) {
  (void)hipFree(h->dinst); (void)hipFree(h->dinst_bucket); (void)hipFree(h->pre);
  (void)hipFree(h->ws);
  h->dinst = nullptr; h->dinst_bucket = nullptr; h->pre = nullptr; h->ws = nullptr;
  h->ws_bytes = 0;
  h->planned = 0;
}

int grid_for(const mbpt_batch *h, int total) {
  const int g = h->grid_cap > 0 ? std::min(total, h->grid_cap) : total;
  return pad_grid(std::max(g, 1));
}

template <typename T> int upload(mbpt_batch *h, const double *const *X, const double *const *W,
                                 const double *const *V, hipStream_t st) {
  T *pool = (T *)h->pool;
  for (int b = 0; b < h->n; ++b) {
    const Inst &I = h->hinst[b];
    const long M = I.M;
    g_split_X<T><<<grid1(M * (I.no + I.nv)), 256, 0, st>>>(I.M, I.no, I.nv, I.ldxo, X[b],
                                                           pool + I.xo,
                                                           pool + I.xv, pool + I.xot);
    g_convert<T><<<grid1(M * M), 256, 0, st>>>(M * M, W[b], pool + I.w);
    if (h->kappa != 0.0) g_convert<T><<<grid1(M * M), 256, 0, st>>>(M * M, V[b], pool + I.v);
    int rc = launch_status("batched operand upload");
    if (rc) return rc;
  }
  return MBPT_OK;
}

template <typename T>
int apply_impl(mbpt_batch *h, const double *z, double *Az, const int *kact, hipStream_t st) {
  const int n = h->n, k = h->kmax;
  const long ldz = h->npmax, zbatch = (long)k * h->npmax;
  const T *pool = (const T *)h->pool;
  T *ws = (T *)h->ws;
  auto tab = [&](int s) { return Tables{h->dinst, h->pre + s * (n + 1), n, h->total[s]}; };
  int rc;
  g_head<T><<<grid_for(h, h->total[S_HEAD]), 256, 0, st>>>(tab(S_HEAD), k, kact, z, ldz, zbatch,
                                                           pool, ws);
  if ((rc = launch_status("g_head"))) return rc;
  if (h->kappa != 0.0) {
    g_hartree_p<T><<<grid_for(h, h->total[S_HP]), 256, 0, st>>>(tab(S_HP), k, kact, pool, ws);
    if ((rc = launch_status("g_hartree_p"))) return rc;
    if (k <= 16)
      g_hartree_u<T, 16><<<grid_for(h, h->total[S_HU]), 256, 0, st>>>(tab(S_HU), k, kact, pool, ws);
    else
      g_hartree_u<T, 64><<<grid_for(h, h->total[S_HU]), 256, 0, st>>>(tab(S_HU), k, kact, pool, ws);
    if ((rc = launch_status("g_hartree_u"))) return rc;
  }
  for (int q = 0; q < kBuckets; ++q) {
    const int tot = h->total[S_X16 + q];
    if (!tot) continue;
    const Tables t{h->dinst_bucket + h->boff[q], h->pre + (S_X16 + q) * (n + 1), h->nb[q], tot};
    const int g = grid_for(h, tot);
    switch (q) {
      case 0: g_exchange<T, 16><<<g, 256, 0, st>>>(t, k, kact, pool, ws); break;
      case 1: g_exchange<T, 32><<<g, 256, 0, st>>>(t, k, kact, pool, ws); break;
      case 2: g_exchange<T, 48><<<g, 256, 0, st>>>(t, k, kact, pool, ws); break;
      default: g_exchange<T, 64><<<g, 256, 0, st>>>(t, k, kact, pool, ws); break;
    }
    if ((rc = launch_status("g_exchange"))) return rc;
  }
  g_tail_prep<T><<<grid_for(h, h->total[S_TPREP]), 256, 0, st>>>(tab(S_TPREP), k, kact,
                                                                  (T)h->kappa, pool, ws);
  if ((rc = launch_status("g_tail_prep"))) return rc;
  g_nn<T, 1><<<grid_for(h, h->total[S_TAIL]), 256, 0, st>>>(tab(S_TAIL), k, kact, pool, ws);
  if ((rc = launch_status("g_nn(tail)"))) return rc;
  g_tail_finish<T><<<grid_for(h, h->total[S_FIN]), 256, 0, st>>>(tab(S_FIN), k, kact, ws,
                                                                  h->dpool, z, Az, ldz, zbatch);
  return launch_status("g_tail_finish");
}

}  // namespace

extern "C" {

int mbpt_batch_destroy(mbpt_batch *h) {
  if (!h) return MBPT_OK;
  free_plan(h);
  (void)hipFree(h->pool);
  (void)hipFree(h->dpool);
  delete h;
  return MBPT_OK;
}

int mbpt_batch_plan(mbpt_batch *h, int rows_per_split, int tail_kchunk, int hartree_kchunk,
                    int grid_cap, int bucket_width, void *stream) {
  MBPT_REQUIRE(h, "mbpt_batch_plan: NULL handle");
  MBPT_REQUIRE(rows_per_split >= 64 && tail_kchunk >= 16 && hartree_kchunk >= 16 && grid_cap >= 0,
               "mbpt_batch_plan: rows_per_split >= 64, K chunks >= 16, grid_cap >= 0");
  MBPT_REQUIRE(bucket_width == 16 || bucket_width == 32,
               "mbpt_batch_plan: bucket_width is 16 or 32");
  hipStream_t st = (hipStream_t)stream;
  MBPT_HIP_TRY(hipStreamSynchronize(st));
  free_plan(h);
  const int n = h->n, k = h->kmax;
  h->rows_target = rows_per_split; h->tail_kchunk = tail_kchunk;
  h->hartree_kchunk = hartree_kchunk; h->grid_cap = grid_cap; h->bucket_width = bucket_width;
  std::vector<int> pre(S_N * (n + 1), 0);
  long wso = 0;
  for (int b = 0; b < n; ++b) {
    Inst &I = h->hinst[b];
    // the exchange's K padding: n_o up to the bucket width (the pool's X_o
    // rows are padded to 32, so either width reads them in place)
    I.nopad = (I.no + bucket_width - 1) / bucket_width * bucket_width;
    const long Mb = I.M;
    const int Mr = (I.M + 63) / 64 * 64;
    I.rps = std::min(Mr, (rows_per_split + 63) / 64 * 64);
    I.nsplit = (I.M + I.rps - 1) / I.rps;
    I.kch = std::min((I.M + 15) / 16 * 16, (tail_kchunk + 15) / 16 * 16);
    I.nq = (I.M + I.kch - 1) / I.kch;
    I.kchu = std::min((I.M + 15) / 16 * 16, (hartree_kchunk + 15) / 16 * 16);
    I.nqu = (I.M + I.kchu - 1) / I.kchu;
    I.zxv = wso; wso += (long)k * I.no * Mb;
    I.p = wso; wso += (long)k * Mb;
    I.u = wso; wso += (long)I.nqu * k * Mb;
    I.tp = wso; wso += (long)I.nsplit * k * I.nopad * Mb;
    I.tc = wso; wso += (long)k * I.no * Mb;
    I.ep = wso; wso += (long)I.nq * k * I.no * I.nv;
    wso = (wso + 63) / 64 * 64;
    const int rows = k * I.no, nmt = (I.M + 63) / 64;
    const long work[kStages] = {
        (long)((rows + 63) / 64) * nmt,
        ((long)k * I.M + 255) / 256,
        h->kappa != 0.0 ? (long)((I.M + 255) / 256) * I.nqu : 0L,
        ((long)rows * I.M + 255) / 256,
        (long)((rows + 63) / 64) * ((I.nv + 63) / 64) * I.nq,
        ((long)rows * I.nv + 255) / 256};
    for (int s = 0; s < kStages; ++s) {
      const long nx = pre[s * (n + 1) + b] + work[s];
      MBPT_REQUIRE(nx < (1L << 30), "mbpt_batch_plan: work table overflows int");
      pre[s * (n + 1) + b + 1] = (int)nx;
    }
  }
  for (int s = 0; s < kStages; ++s) h->total[s] = pre[s * (n + 1) + n];
  // exchange buckets: instances regrouped by occupied width (16, 32, 48, 64)
  std::vector<Inst> grouped;
  for (int q = 0; q < kBuckets; ++q) {
    h->boff[q] = (int)grouped.size();
    int *pb = &pre[(S_X16 + q) * (n + 1)];
    int j = 0;
    for (int b = 0; b < n; ++b) {
      const Inst &I = h->hinst[b];
      if (I.nopad != 16 * (q + 1)) continue;
      grouped.push_back(I);
      const long nx = pb[j] + (long)((I.M + 63) / 64) * k * I.nsplit;
      MBPT_REQUIRE(nx < (1L << 30), "mbpt_batch_plan: work table overflows int");
      pb[j + 1] = (int)nx;
      ++j;
    }
    h->nb[q] = j;
    h->total[S_X16 + q] = pb[j];
  }
  h->ws_bytes = h->esz * (size_t)std::max(wso, 1L);
  MBPT_HIP_TRY(hipMalloc(&h->dinst, sizeof(Inst) * n));
  MBPT_HIP_TRY(hipMalloc(&h->dinst_bucket, sizeof(Inst) * n));
  MBPT_HIP_TRY(hipMalloc(&h->pre, sizeof(int) * pre.size()));
  MBPT_HIP_TRY(hipMalloc(&h->ws, h->ws_bytes));
  MBPT_HIP_TRY(hipMemcpyAsync(h->dinst, h->hinst.data(), sizeof(Inst) * n,
                              hipMemcpyHostToDevice, st));
  MBPT_HIP_TRY(hipMemcpyAsync(h->dinst_bucket, grouped.data(), sizeof(Inst) * n,
                              hipMemcpyHostToDevice, st));
  MBPT_HIP_TRY(hipMemcpyAsync(h->pre, pre.data(), sizeof(int) * pre.size(),
                              hipMemcpyHostToDevice, st));
  MBPT_HIP_TRY(hipStreamSynchronize(st));
  h->planned = 1;
  return MBPT_OK;
}

int mbpt_batch_create(mbpt_batch **out, int n, const int *M, const int *nocc, const int *nvir,
                      int precision, double kappa, const double *const *X_dev,
                      const double *const *W_dev, const double *const *V_dev,
                      const double *eps_host, int kmax, void *stream) {
  MBPT_REQUIRE(out && M && nocc && nvir && X_dev && W_dev && eps_host,
               "mbpt_batch_create: NULL argument");
  *out = nullptr;
  MBPT_REQUIRE(n > 0 && kmax > 0, "mbpt_batch_create: n and kmax must be positive");
  if (precision == MBPT_BF16X3) {
    set_error("mbpt_batch_create: the batched action runs MBPT_FP32 or MBPT_FP64");
    return MBPT_ERR_UNSUPPORTED;
  }
  MBPT_REQUIRE(precision == MBPT_FP32 || precision == MBPT_FP64,
               "mbpt_batch_create: unknown precision");
  MBPT_REQUIRE(kappa == 0.0 || V_dev, "mbpt_batch_create: V_dev is NULL with kappa != 0");
  for (int b = 0; b < n; ++b) {
    MBPT_REQUIRE(M[b] > 0 && nocc[b] > 0 && nvir[b] > 0, "mbpt_batch_create: sizes must be positive");
    MBPT_REQUIRE(nocc[b] <= 64, "mbpt_batch_create: nocc > 64 per instance is not supported");
    MBPT_REQUIRE(X_dev[b] && W_dev[b] && (kappa == 0.0 || V_dev[b]),
                 "mbpt_batch_create: NULL operand pointer");
  }
  hipStream_t st = (hipStream_t)stream;
  mbpt_batch *h = new mbpt_batch();
  h->n = n; h->kmax = kmax; h->kappa = kappa; h->prec = precision;
  h->esz = precision == MBPT_FP64 ? sizeof(double) : sizeof(float);
  h->hinst.resize(n);
  long fo = 0, dof = 0, eo = 0;
  int npmax = 0;
  for (int b = 0; b < n; ++b) {
    Inst &I = h->hinst[b];
    I = Inst{};
    I.id = b;
    I.M = M[b]; I.no = nocc[b]; I.nv = nvir[b];
    I.nopad = (I.no + 15) / 16 * 16;
    I.ldxo = (I.no + 31) / 32 * 32;
    npmax = std::max(npmax, I.no * I.nv);
    const long Mb = I.M;
    I.xo = fo; fo += (Mb + 31) / 32 * 32 * I.ldxo;   // zero-padded for the exchange
    I.xv = fo; fo += Mb * I.nv;
    I.xot = fo; fo += Mb * I.no;
    I.w = fo; fo += Mb * Mb;
    I.v = fo; fo += (kappa != 0.0) ? Mb * Mb : 0;
    fo = (fo + 63) / 64 * 64;
    I.d = dof; dof += (long)I.no * I.nv;
  }
  h->npmax = npmax;
  h->pool_elems = fo;
  std::vector<double> dh(dof);
  for (int b = 0; b < n; ++b) {
    const Inst &I = h->hinst[b];
    const double *e = eps_host + eo;
    for (int i = 0; i < I.no; ++i)
      for (int a = 0; a < I.nv; ++a) dh[I.d + (long)i * I.nv + a] = e[I.no + a] - e[i];
    eo += I.no + I.nv;
  }
  hipError_t e = hipMalloc(&h->pool, h->esz * fo);
  // zero padding of the packed X_o (the upload writes only the real entries)
  if (e == hipSuccess) e = hipMemsetAsync(h->pool, 0, h->esz * fo, st);
  if (e == hipSuccess) e = hipMalloc(&h->dpool, sizeof(double) * dof);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h->dpool, dh.data(), sizeof(double) * dof, hipMemcpyHostToDevice, st);
  if (e != hipSuccess) {
    set_error(std::string("mbpt_batch_create: ") + hipGetErrorString(e));
    mbpt_batch_destroy(h);
    return MBPT_ERR_HIP;
  }
  int rc = precision == MBPT_FP64 ? upload<double>(h, X_dev, W_dev, V_dev, st)
                                  : upload<float>(h, X_dev, W_dev, V_dev, st);
  if (!rc) rc = mbpt_batch_plan(h, 1024, 512, 512, 0, 16, stream);
  if (rc) {
    mbpt_batch_destroy(h);
    return rc;
  }
  *out = h;
  return MBPT_OK;
}

int mbpt_batch_screened_diag(mbpt_batch *h, double *out_dev, void *stream) {
  MBPT_REQUIRE(h && out_dev, "mbpt_batch_screened_diag: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  const int n = h->n;
  std::vector<int> pre(2 * (n + 1), 0);
  std::vector<long> yoff(n);
  long yo = 0;
  for (int b = 0; b < n; ++b) {
    const Inst &I = h->hinst[b];
    const int ntn = (I.nv + 63) / 64;
    pre[b + 1] = pre[b] + ((I.M + 63) / 64) * ntn;
    pre[(n + 1) + b + 1] = pre[(n + 1) + b] + ((I.no + 63) / 64) * ntn;
    yoff[b] = yo;
    yo += (long)I.M * I.nv;
  }
  int *dpre = nullptr;
  long *dyoff = nullptr;
  void *Y = nullptr;
  hipError_t e = hipMalloc(&dpre, sizeof(int) * pre.size());
  if (!e) e = hipMalloc(&dyoff, sizeof(long) * n);
  if (!e) e = hipMalloc(&Y, h->esz * std::max(yo, 1L));
  if (!e) e = hipMemcpyAsync(dpre, pre.data(), sizeof(int) * pre.size(), hipMemcpyHostToDevice, st);
  if (!e) e = hipMemcpyAsync(dyoff, yoff.data(), sizeof(long) * n, hipMemcpyHostToDevice, st);
  int rc = MBPT_OK;
  if (!e) {
    const Tables t0{h->dinst, dpre, n, pre[n]}, t1{h->dinst, dpre + n + 1, n, pre[2 * n + 1]};
    if (h->prec == MBPT_FP64) {
      g_sdiag<double, 0><<<pad_grid(t0.total), 256, 0, st>>>(t0, (const double *)h->pool, dyoff,
                                                             (double *)Y, h->dpool, out_dev, h->npmax);
      g_sdiag<double, 1><<<pad_grid(t1.total), 256, 0, st>>>(t1, (const double *)h->pool, dyoff,
                                                             (double *)Y, h->dpool, out_dev, h->npmax);
    } else {
      g_sdiag<float, 0><<<pad_grid(t0.total), 256, 0, st>>>(t0, (const float *)h->pool, dyoff,
                                                            (float *)Y, h->dpool, out_dev, h->npmax);
      g_sdiag<float, 1><<<pad_grid(t1.total), 256, 0, st>>>(t1, (const float *)h->pool, dyoff,
                                                            (float *)Y, h->dpool, out_dev, h->npmax);
    }
    rc = launch_status("g_sdiag");
    if (!rc) e = hipStreamSynchronize(st);
  }
  (void)hipFree(dpre); (void)hipFree(dyoff); (void)hipFree(Y);
  if (e != hipSuccess) {
    set_error(std::string("mbpt_batch_screened_diag: ") + hipGetErrorString(e));
    return MBPT_ERR_HIP;
  }
  return rc;
}

int mbpt_batch_npair_max(const mbpt_batch *h) { return h ? h->npmax : 0; }

size_t mbpt_batch_held_bytes(const mbpt_batch *h) {
  return h ? h->esz * (size_t)h->pool_elems + h->ws_bytes : 0;
}

int mbpt_batch_apply(mbpt_batch *h, const double *z_dev, double *Az_dev, const int *kact_dev,
                     void *stream) {
  MBPT_REQUIRE(h && h->planned, "mbpt_batch_apply: handle without a plan");
  MBPT_REQUIRE(z_dev && Az_dev && z_dev != Az_dev, "mbpt_batch_apply: z and Az must be distinct");
  hipStream_t st = (hipStream_t)stream;
  return h->prec == MBPT_FP64 ? apply_impl<double>(h, z_dev, Az_dev, kact_dev, st)
                              : apply_impl<float>(h, z_dev, Az_dev, kact_dev, st);
}

}  // extern "C"
