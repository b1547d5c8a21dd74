// FP64 kernels of the single-instance Davidson iteration on gfx950: the
// products over the long pair index that its Rayleigh-Ritz step, residuals,
// corrections and trial-space orthogonalisation are made of.  The trial
// space V and its image AV are column blocks of row-major buffers (n rows,
// the pair index); every kernel takes an operand as up to two column ranges
// of one such buffer, so [V | AV] or [V | T] pass without a copy.
//
// All products use v_mfma_f64_16x16x4f64 with each operand fragment loaded
// straight from HBM.  The k order inside an MFMA is free as long as both
// operands agree on it, so a lane's four k values are chosen to be adjacent
// in memory: every row access is a 128-byte segment per 16 lanes.
#include "common.h"

#include <algorithm>

namespace mbpt {

// Logical column i of a two-range operand: p[r * ld + (i < n1 ? off1 + i : off2 + i - n1)]
struct Cols2 {
  const double *p;
  int ld, n1, off1, off2;
  __device__ __forceinline__ int col(int i) const { return i < n1 ? off1 + i : off2 + (i - n1); }
};

typedef double f64x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f64x4 mma64(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// ---- C (a x b) = A^T B over rows [r0, r1): a workgroup of four waves per
// (64 x 32 tile, chunk of rows), each wave a quarter of the chunk, their
// tiles summed in LDS before the chunk's partial is written.
//
// Lane l holds columns i = 16 f + l % 16 of A (f < 4) and j = 16 g + l % 16
// of B (g < 2).  A step covers 4 rows: lane l reads row base + l / 16, which
// is the k index of its operand in all eight MFMAs (f, g) of the step.
__global__ __launch_bounds__(256) void k_tns_partial(int n, int a, int b, int chunk, Cols2 A,
                                                    Cols2 B, double *__restrict__ P, int nta,
                                                    int ntb, int nblk) {
  __shared__ double red[3][32 * 64];
  const int lid = xcd_logical(blockIdx.x, nblk);
  const int ntiles = nta * ntb;
  const int c = lid / ntiles, t = lid % ntiles;
  const int ta = t / ntb, tb = t % ntb;
  if (c * chunk >= n) return;   // the whole workgroup
  const int wave = threadIdx.x >> 6, sub = chunk / 4;
  const int r0 = c * chunk + wave * sub;
  const int r1 = min(n, r0 + sub);   // r1 <= r0: this wave has no rows
  const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
  int ca[4], cb[2];
#pragma unroll
  for (int f = 0; f < 4; ++f) ca[f] = A.col(min(a - 1, 64 * ta + 16 * f + lr));
#pragma unroll
  for (int g = 0; g < 2; ++g) cb[g] = B.col(min(b - 1, 32 * tb + 16 * g + lr));
  f64x4 acc[4][2];
#pragma unroll
  for (int f = 0; f < 4; ++f)
#pragma unroll
    for (int g = 0; g < 2; ++g) acc[f][g] = f64x4{0.0, 0.0, 0.0, 0.0};
  // a pipeline of PD steps: the loads of step s + PD are issued as soon as
  // step s's MFMAs have read their operands, so PD steps stay in flight.
  // Rows past r1 are read at r1 - 1 (always in bounds, no branch around the
  // loads) and their B operand is zeroed where it is used.
  constexpr int PD = 8;
  double va[PD][4], vb[PD][2];
  bool ok[PD];
  auto load = [&](int row, double (&xa)[4], double (&xb)[2], bool &in) {
    in = row < r1;
    const long rr = in ? row : r1 - 1;
    const long ra = rr * A.ld, rb = rr * B.ld;
#pragma unroll
    for (int f = 0; f < 4; ++f) xa[f] = A.p[ra + ca[f]];
#pragma unroll
    for (int g = 0; g < 2; ++g) xb[g] = B.p[rb + cb[g]];
  };
#pragma unroll
  for (int d = 0; d < PD; ++d) load(r0 + 4 * d + lq, va[d], vb[d], ok[d]);
  for (int base = r0; base < r1; base += 4 * PD) {
#pragma unroll
    for (int d = 0; d < PD; ++d) {
      double xb[2];
#pragma unroll
      for (int g = 0; g < 2; ++g) xb[g] = ok[d] ? vb[d][g] : 0.0;
#pragma unroll
      for (int f = 0; f < 4; ++f)
#pragma unroll
        for (int g = 0; g < 2; ++g) acc[f][g] = mma64(va[d][f], xb[g], acc[f][g]);
      load(base + 4 * (PD + d) + lq, va[d], vb[d], ok[d]);
    }
  }
  // the four waves' tiles summed in a fixed order by wave 0
  if (wave > 0) {
#pragma unroll
    for (int f = 0; f < 4; ++f)
#pragma unroll
      for (int g = 0; g < 2; ++g)
#pragma unroll
        for (int q = 0; q < 4; ++q) red[wave - 1][((f * 2 + g) * 4 + q) * 64 + lane] = acc[f][g][q];
  }
  __syncthreads();
  if (wave > 0) return;
  // D[l / 16 + 4 q][l % 16] of each 16 x 16 tile (the f64 MFMA output layout)
  double *out = P + (long)c * a * b;
#pragma unroll
  for (int f = 0; f < 4; ++f)
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = 64 * ta + 16 * f + lq + 4 * q, j = 32 * tb + 16 * g + lr;
        const int e = ((f * 2 + g) * 4 + q) * 64 + lane;
        const double v = ((acc[f][g][q] + red[0][e]) + red[1][e]) + red[2][e];
        if (i < a && j < b) out[(long)i * b + j] = v;
      }
}

// Fixed-order sum of the chunk partials: C[i][j] = sum_c P[c][i][j].
__global__ __launch_bounds__(256) void k_tns_sum(int a, int b, int nc,
                                                 const double *__restrict__ P,
                                                 double *__restrict__ C, int ldc) {
  __shared__ double red[16][17];
  const long count = (long)a * b;
  const int el = threadIdx.x & 15, g = threadIdx.x >> 4;
  for (long base = blockIdx.x * 16L; base < count; base += (long)gridDim.x * 16) {
    const long e = base + el;
    double s = 0.0;
    if (e < count)
      for (int c = g; c < nc; c += 16) s += P[(long)c * count + e];
    red[g][el] = s;
    __syncthreads();
    if (g == 0 && e < count) {
      double t = 0.0;
      for (int q = 0; q < 16; ++q) t += red[q][el];
      C[(e / b) * ldc + (e % b)] = t;
    }
    __syncthreads();
  }
}

// ---- C (n x c) = beta C + alpha A (n x k) Y (k x c); with R != null also
// R = A2 Y - X diag(theta) where X = A Y is the first product (the Ritz
// vectors and their residuals from one pass over V and AV).
//
// A workgroup of four waves takes 64 rows, a wave 16 rows and all of c
// (<= 64, four 16-wide tiles); Y (k x c) is staged in LDS by the workgroup,
// kNnsRows rows at a time, padded to c16 + 4 doubles so that the four k rows
// one MFMA reads fall in different bank halves.  Lane l reads row i0 + l % 16
// at the four adjacent columns k0 + 4 (l / 16) + s and uses column s in
// MFMA s; the loads of two 16-wide k steps are issued together.
constexpr int kNnsRows = 96;
template <bool RITZ>
__global__ __launch_bounds__(256) void k_nns(int n, int k, int c, Cols2 A, Cols2 A2,
                                             const double *__restrict__ Y, int ldy,
                                             const double *__restrict__ theta,
                                             double *__restrict__ C, int ldc, int c_off,
                                             double *__restrict__ R, int ldr, double alpha,
                                             double beta) {
  extern __shared__ double ys[];
  const int ntile = (c + 15) / 16, ldys = 16 * ntile + 4;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i0 = blockIdx.x * 64 + 16 * wave;
  const bool active = i0 < n;   // waves past n still take part in the barriers
  const int lr = lane & 15, lq = lane >> 4;
  const int row = min(n - 1, i0 + lr);
  const long ra = (long)row * A.ld, ra2 = (long)row * A2.ld;
  f64x4 acc[4], acc2[4];
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    acc[g] = f64x4{0.0, 0.0, 0.0, 0.0};
    acc2[g] = f64x4{0.0, 0.0, 0.0, 0.0};
  }
  for (int kb = 0; kb < k; kb += kNnsRows) {
    const int kn = min(kNnsRows, k - kb), kn32 = (kn + 31) / 32 * 32;
    __syncthreads();   // the previous rows of Y are no longer read
    for (int e = threadIdx.x; e < kn32 * 16 * ntile; e += 256) {
      const int kk = e / (16 * ntile), j = e % (16 * ntile);
      ys[kk * ldys + j] = (kk < kn && j < c) ? Y[(long)(kb + kk) * ldy + j] : 0.0;
    }
    __syncthreads();
    if (!active) continue;
    for (int k0 = 0; k0 < kn; k0 += 32) {
      double va[2][4], va2[2][4];
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int kc = min(kb + k0 + 16 * h + 4 * lq + s, k - 1);   // past k: Y's row is 0
          va[h][s] = A.p[ra + A.col(kc)];
          if (RITZ) va2[h][s] = A2.p[ra2 + A2.col(kc)];
        }
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          if (g >= ntile) break;
          const int j = 16 * g + lr;
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const double y = ys[(k0 + 16 * h + 4 * lq + s) * ldys + j];
            acc[g] = mma64(va[h][s], y, acc[g]);
            if (RITZ) acc2[g] = mma64(va2[h][s], y, acc2[g]);
          }
        }
    }
  }
  if (!active) return;
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    if (g >= ntile) break;
    const int j = 16 * g + lr;
    if (j >= c) continue;
    const double th = RITZ ? theta[j] : 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = i0 + lq + 4 * q;
      if (i >= n) continue;
      const long e = (long)i * ldc + c_off + j;
      const double x = alpha * acc[g][q];
      C[e] = beta == 0.0 ? x : beta * C[e] + x;
      if (RITZ) R[(long)i * ldr + j] = acc2[g][q] - acc[g][q] * th;
    }
  }
}

// LDS bytes of k_nns's staged rows of Y (at most 96 x 68 doubles)
inline size_t nns_lds(int k, int c) {
  const int ntile = (c + 15) / 16;
  return (size_t)std::min(kNnsRows, (k + 31) / 32 * 32) * (16 * ntile + 4) * sizeof(double);
}

// ---- Unit scaling of a projected correction block before the operator:
// from the two Gram-Schmidt passes' projections (S1 = V^T T, G0 = T^T T,
// S2 = V^T T1, G1 = T1^T T1, laid out as the rows of `pr`, (3m + 2p + 2) x p)
// column j keeps norm a_j = sqrt(max(G1_jj - |S2_:j|^2, 0)) after the
// projections against b_j = sqrt(max(G0_jj, 0)) before; it is kept when
// a_j > drop max(b_j, 1e-300).  Rows 3m + 2p and 3m + 2p + 1 of `pr` receive
// a_j and the 0 / 1 keep flag, and the block's columns are scaled by 1 / a_j
// (kept) or 0 (dropped).
__global__ __launch_bounds__(256) void k_block_norms(int m, int p, double drop,
                                                     double *__restrict__ pr) {
  __shared__ double part[4][64];
  const int j = threadIdx.x & 63, q = threadIdx.x >> 6;   // column, quarter of S2's rows
  const double *S2 = pr + (long)(2 * m + p) * p, *G1 = S2 + (long)m * p;
  double s2 = 0.0;
  if (j < p)
    for (int i = q; i < m; i += 4) s2 += S2[(long)i * p + j] * S2[(long)i * p + j];
  part[q][j] = s2;
  __syncthreads();
  if (q != 0 || j >= p) return;
  s2 = (part[0][j] + part[1][j]) + (part[2][j] + part[3][j]);
  const double a = sqrt(fmax(G1[(long)j * p + j] - s2, 0.0));
  const double b = sqrt(fmax(pr[(long)(m + j) * p + j], 0.0));
  pr[(long)(3 * m + 2 * p) * p + j] = a;
  pr[(long)(3 * m + 2 * p + 1) * p + j] = a > drop * fmax(b, 1e-300) ? 1.0 : 0.0;
}

__global__ __launch_bounds__(256) void k_block_scale(int n, int m, int p,
                                                     const double *__restrict__ pr,
                                                     double *__restrict__ T, int ldt) {
  const long e = blockIdx.x * 256L + threadIdx.x;
  if (e >= (long)n * p) return;
  const int i = (int)(e / p), j = (int)(e % p);
  const double a = pr[(long)(3 * m + 2 * p) * p + j];
  const bool keep = pr[(long)(3 * m + 2 * p + 1) * p + j] > 0.5;
  T[(long)i * ldt + j] *= keep ? 1.0 / fmax(a, 1e-300) : 0.0;
}

// ---- Jacobi-Davidson correction (Solvers/davidson.py:43-77) for k columns:
// t = K^-1 r - alpha K^-1 u, K = diag - theta floored at `floor` in magnitude,
// alpha = u.K^-1 r / u.K^-1 u.  Lane l of a workgroup takes column
// j = 64 blockIdx.y + l, so a wave reads 64 adjacent entries of a row; the
// four waves take every fourth of the workgroup's kJdRows rows.  Pass 1:
// per-workgroup partial sums of u.K^-1 u and u.K^-1 r (fixed order: rows in
// order within a wave, then the waves, then the workgroups); pass 2: alpha
// per column; pass 3: t.
constexpr int kJdRows = 128;

__device__ __forceinline__ double jd_den(double d, double th, double floor) {
  const double x = d - th;
  return fabs(x) < floor ? floor : x;
}

__global__ __launch_bounds__(256) void k_jd_partial(int n, int k, const double *__restrict__ U,
                                                    int ldu, const double *__restrict__ R,
                                                    int ldr, const double *__restrict__ dg,
                                                    const double *__restrict__ theta,
                                                    double floor, double *__restrict__ part) {
  __shared__ double red[2][4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = 64 * blockIdx.y + lane;
  const bool col = j < k;
  const double th = col ? theta[j] : 0.0;
  const int i0 = blockIdx.x * kJdRows + wave, i1 = min(n, (blockIdx.x + 1) * kJdRows);
  double uku = 0.0, ukr = 0.0;
  if (col) {
#pragma unroll 8
    for (int i = i0; i < i1; i += 4) {
      const double den = jd_den(dg[i], th, floor);
      const double u = U[(long)i * ldu + j], r = R[(long)i * ldr + j];
      uku += u * u / den;
      ukr += u * r / den;
    }
  }
  red[0][wave][lane] = uku;
  red[1][wave][lane] = ukr;
  __syncthreads();
  if (wave == 0 && col) {
    const double a = ((red[0][0][lane] + red[0][1][lane]) + red[0][2][lane]) + red[0][3][lane];
    const double b = ((red[1][0][lane] + red[1][1][lane]) + red[1][2][lane]) + red[1][3][lane];
    part[((long)j * gridDim.x + blockIdx.x) * 2] = a;
    part[((long)j * gridDim.x + blockIdx.x) * 2 + 1] = b;
  }
}

// alpha_j from the nblk partials of column j, into alpha[j]: sixteen waves,
// each summing every sixteenth partial, then the waves in order
__global__ __launch_bounds__(1024) void k_jd_alpha(int k, int nblk,
                                                   const double *__restrict__ part,
                                                   double *__restrict__ alpha) {
  __shared__ double red[2][16][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = 64 * blockIdx.x + lane;
  double a = 0.0, b = 0.0;
  if (j < k) {
#pragma unroll 4
    for (int q = wave; q < nblk; q += 16) {
      a += part[((long)j * nblk + q) * 2];
      b += part[((long)j * nblk + q) * 2 + 1];
    }
  }
  red[0][wave][lane] = a;
  red[1][wave][lane] = b;
  __syncthreads();
  if (wave == 0 && j < k) {
    double su = 0.0, sr = 0.0;
    for (int w = 0; w < 16; ++w) {
      su += red[0][w][lane];
      sr += red[1][w][lane];
    }
    alpha[j] = fabs(su) > 1e-300 ? sr / su : 0.0;
  }
}

__global__ __launch_bounds__(256) void k_jd_apply(int n, int k, const double *__restrict__ U,
                                                  int ldu, const double *__restrict__ R, int ldr,
                                                  const double *__restrict__ dg,
                                                  const double *__restrict__ theta,
                                                  double floor,
                                                  const double *__restrict__ alpha,
                                                  double *__restrict__ T, int ldt) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = 64 * blockIdx.y + lane;
  if (j >= k) return;
  const double th = theta[j], al = alpha[j];
  const int i0 = blockIdx.x * kJdRows + wave, i1 = min(n, (blockIdx.x + 1) * kJdRows);
#pragma unroll 8
  for (int i = i0; i < i1; i += 4) {
    const double den = jd_den(dg[i], th, floor);
    T[(long)i * ldt + j] = (R[(long)i * ldr + j] - al * U[(long)i * ldu + j]) / den;
  }
}

}  // namespace mbpt

using namespace mbpt;

namespace {

// Rows a workgroup of k_tns_partial covers: a multiple of 128 (four waves
// of whole 32-row groups), about 2048 waves over all tiles and chunks.
int tns_chunk(int n, int a, int b, int *nc_out) {
  const long tiles = (long)((a + 63) / 64) * ((b + 31) / 32);
  long nc = std::max<long>(1, 512 / tiles);
  nc = std::min<long>(nc, (n + 127) / 128);
  int chunk = (int)((n + nc - 1) / nc);
  chunk = (chunk + 127) / 128 * 128;
  *nc_out = (n + chunk - 1) / chunk;
  return chunk;
}

bool cols_ok(int ncol, int ld, int n1, int off1, int off2) {
  if (ncol < 0 || n1 < 0 || n1 > ncol || off1 < 0 || off2 < 0) return false;
  if (n1 > 0 && off1 + n1 > ld) return false;
  if (ncol > n1 && off2 + (ncol - n1) > ld) return false;
  return true;
}

}  // namespace

extern "C" {

long mbpt_gemm_tn_seg_f64_work(int n, int a, int b) {
  if (n <= 0 || a <= 0 || b <= 0) return 0;
  int nc = 0;
  tns_chunk(n, a, b, &nc);
  return (long)nc * a * b;
}

int mbpt_gemm_tn_seg_f64(int n, int a, int b, const double *A, int lda, int a1, int aoff1,
                         int aoff2, const double *B, int ldb, int b1, int boff1, int boff2,
                         double *C, int ldc, double *work, long work_elems, void *stream) {
  MBPT_REQUIRE(n > 0 && a >= 0 && b >= 0, "mbpt_gemm_tn_seg_f64: n must be positive, a and b non-negative");
  if (a == 0 || b == 0) return MBPT_OK;
  MBPT_REQUIRE(A && B && C && work, "mbpt_gemm_tn_seg_f64: NULL argument");
  MBPT_REQUIRE(cols_ok(a, lda, a1, aoff1, aoff2) && cols_ok(b, ldb, b1, boff1, boff2),
               "mbpt_gemm_tn_seg_f64: a column range runs past its row");
  MBPT_REQUIRE(ldc >= b, "mbpt_gemm_tn_seg_f64: ldc < b");
  int nc = 0;
  const int chunk = tns_chunk(n, a, b, &nc);
  MBPT_REQUIRE(work_elems >= (long)nc * a * b, "mbpt_gemm_tn_seg_f64: work buffer too small");
  hipStream_t st = (hipStream_t)stream;
  const int nta = (a + 63) / 64, ntb = (b + 31) / 32;
  const int nblk = pad_grid((long)nta * ntb * nc);
  k_tns_partial<<<nblk, 256, 0, st>>>(n, a, b, chunk, Cols2{A, lda, a1, aoff1, aoff2},
                                     Cols2{B, ldb, b1, boff1, boff2}, work, nta, ntb, nblk);
  int rc = launch_status("k_tns_partial");
  if (rc) return rc;
  const long count = (long)a * b;
  k_tns_sum<<<(int)std::min<long>((count + 15) / 16, 8192L), 256, 0, st>>>(a, b, nc, work, C,
                                                                          ldc);
  return launch_status("k_tns_sum");
}

int mbpt_gemm_nn_seg_f64(int n, int k, int c, const double *A, int lda, int k1, int aoff1,
                         int aoff2, const double *Y, int ldy, double *C, int ldc, int c_off,
                         double alpha, double beta, void *stream) {
  MBPT_REQUIRE(n > 0 && k >= 0 && c >= 0 && c <= 64, "mbpt_gemm_nn_seg_f64: need n > 0, k >= 0, 0 <= c <= 64");
  if (c == 0) return MBPT_OK;
  MBPT_REQUIRE(A && Y && C, "mbpt_gemm_nn_seg_f64: NULL argument");
  MBPT_REQUIRE(cols_ok(k, lda, k1, aoff1, aoff2), "mbpt_gemm_nn_seg_f64: a column range runs past its row");
  MBPT_REQUIRE(ldy >= c && c_off >= 0 && c_off + c <= ldc, "mbpt_gemm_nn_seg_f64: bad ldy, ldc or c_off");
  hipStream_t st = (hipStream_t)stream;
  const Cols2 a{A, lda, k1, aoff1, aoff2};
  k_nns<false><<<(n + 63) / 64, 256, nns_lds(k, c), st>>>(n, k, c, a, a, Y, ldy, nullptr, C, ldc,
                                                          c_off, nullptr, 0, alpha, beta);
  return launch_status("k_nns");
}

int mbpt_ritz_residual_f64(int n, int m, int k, const double *V, int ldv, int voff,
                           const double *AV, int ldav, int avoff, const double *Y, int ldy,
                           const double *theta, double *X, int ldx, double *R, int ldr,
                           void *stream) {
  MBPT_REQUIRE(n > 0 && m >= 0 && k >= 0 && k <= 64, "mbpt_ritz_residual_f64: need n > 0, m >= 0, 0 <= k <= 64");
  if (k == 0) return MBPT_OK;
  MBPT_REQUIRE(V && AV && Y && theta && X && R, "mbpt_ritz_residual_f64: NULL argument");
  MBPT_REQUIRE(voff >= 0 && voff + m <= ldv && avoff >= 0 && avoff + m <= ldav,
               "mbpt_ritz_residual_f64: a column range runs past its row");
  MBPT_REQUIRE(ldy >= k && ldx >= k && ldr >= k, "mbpt_ritz_residual_f64: leading dimension too small");
  hipStream_t st = (hipStream_t)stream;
  k_nns<true><<<(n + 63) / 64, 256, nns_lds(m, k), st>>>(n, m, k, Cols2{V, ldv, m, voff, 0},
                                            Cols2{AV, ldav, m, avoff, 0}, Y, ldy, theta, X,
                                            ldx, 0, R, ldr, 1.0, 0.0);
  return launch_status("k_nns<ritz>");
}

int mbpt_block_scale_f64(int n, int m, int p, double drop, double *pr, double *T, int ldt,
                         void *stream) {
  MBPT_REQUIRE(n > 0 && m >= 0 && p >= 0 && p <= 64, "mbpt_block_scale_f64: need n > 0, m >= 0, 0 <= p <= 64");
  if (p == 0) return MBPT_OK;
  MBPT_REQUIRE(pr && T && ldt >= p, "mbpt_block_scale_f64: NULL argument or ldt < p");
  MBPT_REQUIRE(drop >= 0.0, "mbpt_block_scale_f64: drop must be >= 0");
  hipStream_t st = (hipStream_t)stream;
  k_block_norms<<<1, 256, 0, st>>>(m, p, drop, pr);
  int rc = launch_status("k_block_norms");
  if (rc) return rc;
  const long total = (long)n * p;
  k_block_scale<<<(int)((total + 255) / 256), 256, 0, st>>>(n, m, p, pr, T, ldt);
  return launch_status("k_block_scale");
}

long mbpt_jd_correction_f64_work(int n, int k) {
  return 2L * k * ((n + kJdRows - 1) / kJdRows) + k;
}

int mbpt_jd_correction_f64(int n, int k, const double *U, int ldu, const double *R, int ldr,
                           const double *diag, const double *theta, double floor, double *T,
                           int ldt, double *work, long work_elems, void *stream) {
  MBPT_REQUIRE(n > 0 && k >= 0 && k <= 65535, "mbpt_jd_correction_f64: need n > 0, 0 <= k <= 65535");
  if (k == 0) return MBPT_OK;
  MBPT_REQUIRE(U && R && diag && theta && T && work, "mbpt_jd_correction_f64: NULL argument");
  MBPT_REQUIRE(ldu >= k && ldr >= k && ldt >= k, "mbpt_jd_correction_f64: leading dimension too small");
  MBPT_REQUIRE(floor > 0.0, "mbpt_jd_correction_f64: floor must be positive");
  const int nblk = (n + kJdRows - 1) / kJdRows, ncg = (k + 63) / 64;
  MBPT_REQUIRE(work_elems >= 2L * k * nblk + k, "mbpt_jd_correction_f64: work buffer too small");
  double *alpha = work + 2L * k * nblk;
  hipStream_t st = (hipStream_t)stream;
  k_jd_partial<<<dim3(nblk, ncg), 256, 0, st>>>(n, k, U, ldu, R, ldr, diag, theta, floor, work);
  int rc = launch_status("k_jd_partial");
  if (rc) return rc;
  k_jd_alpha<<<ncg, 1024, 0, st>>>(k, nblk, work, alpha);
  if ((rc = launch_status("k_jd_alpha"))) return rc;
  k_jd_apply<<<dim3(nblk, ncg), 256, 0, st>>>(n, k, U, ldu, R, ldr, diag, theta, floor, alpha,
                                              T, ldt);
  return launch_status("k_jd_apply");
}

}  // extern "C"
