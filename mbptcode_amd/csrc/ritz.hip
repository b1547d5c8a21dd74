// Device Rayleigh-Ritz step of the single-instance Davidson (FP64): the
// lowest k eigenpairs of the symmetric part of its m x m Rayleigh matrix,
// 1 <= m <= 192, 1 <= k <= min(m, 64), and the two quadratic forms the
// iteration's residual norms come from, without a trip to the host.
//
//   k_ritz_tridiag   Householder reduction to tridiagonal form in one
//                    workgroup, the matrix in registers (a 16 x 16 or 32 x 16
//                    grid of threads, each holding a tile of entries).  Step
//                    j: reflector from column j, p = tau A v,
//                    w = p - (tau / 2)(p.v) v, A -= v w^T + w v^T on the
//                    trailing block (Golub & Van Loan, "Matrix Computations",
//                    8.3.1).
//   k_ritz_tri_eig   one 64-lane workgroup per wanted eigenvalue: Sturm-count
//                    multisection on the Gershgorin interval (64 points per
//                    round) to working precision, then two steps of inverse
//                    iteration with partial pivoting for its eigenvector
//                    (Golub & Van Loan 8.4.1-8.4.2 and 7.6.1).
//   k_ritz_back      eigenvectors of eigenvalues closer than 1e-6 ||T|| are
//                    orthogonalised against their cluster (two classical
//                    Gram-Schmidt passes), then taken back through the
//                    reflectors: y = Q z.
//   k_ritz_quad      y^T G y and y^T H y per eigenvector (H as given, not
//                    symmetrised: a lower-precision operator's Rayleigh matrix
//                    is not exactly symmetric).
#include "common.h"

#include <cfloat>
#include <type_traits>

namespace mbpt {
namespace {

constexpr int kRB = 6;               // 32-wide blocks per dimension
constexpr int kRMax = 32 * kRB;      // largest m
constexpr int kRK = 64;              // most eigenpairs
constexpr int kZLd = kRMax + 1;      // LDS row stride of the eigenvector block

// workspace (doubles): d, e, tau [kRMax each], tnorm [1], the reflectors
// [kRMax x kRMax], the tridiagonal eigenvectors [kRK x kRMax]
constexpr long kOffD = 0, kOffE = kRMax, kOffTau = 2 * kRMax, kOffNorm = 3 * kRMax;
constexpr long kOffHv = 3 * kRMax + 8, kOffZ = kOffHv + (long)kRMax * kRMax;
constexpr long kRitzWork = kOffZ + (long)kRK * kRMax;

// a double moved across lanes by a DPP control (row_mask selects the rows of
// 16 lanes that receive it; the others get 0)
template <int CTRL, int ROWS = 0xF> __device__ __forceinline__ double dpp(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, ROWS, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, ROWS, 0xF, false);
  return __hiloint2double(hi, lo);
}
// every lane: the sum over its row of 16 lanes (row rotations by 8, 4, 2, 1)
__device__ __forceinline__ double sum16(double x) {
  x += dpp<0x128>(x);
  x += dpp<0x124>(x);
  x += dpp<0x122>(x);
  return x + dpp<0x121>(x);
}
// lanes 16-31 and 48-63: the sum over their half-wave (row 0's total added to
// row 1 and row 2's to row 3 by the row-15 broadcast)
__device__ __forceinline__ double sum32_tail(double x) {
  x = sum16(x);
  return x + dpp<0x142, 0xA>(x);
}
// every lane: the sum over the wave (lane 31's half total broadcast to rows 2
// and 3, then lane 63 read out)
__device__ __forceinline__ double wave_sum(double x) {
  x = sum32_tail(x);
  x += dpp<0x143, 0xC>(x);
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), 63),
                          __builtin_amdgcn_readlane(__double2loint(x), 63));
}
// a workgroup barrier that orders LDS only: global stores stay in flight
__device__ __forceinline__ void lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// entry (a, b) of a thread's register tile for run-time block indices
template <int NR, int NC>
__device__ __forceinline__ double tile_get(const double (&A)[NR][NC], int a, int b) {
  double x = 0.0;
#pragma unroll
  for (int p = 0; p < NR; ++p)
#pragma unroll
    for (int q = 0; q < NC; ++q)
      if (p == a && q == b) x = A[p][q];
  return x;
}

// 1 / q to working precision: the hardware reciprocal and two Newton steps
__device__ __forceinline__ double recip(double q) {
  double y = __builtin_amdgcn_rcp(q);
  y = fma(y, fma(-q, y, 1.0), y);
  return fma(y, fma(-q, y, 1.0), y);
}

// f(std::integral_constant<int, v>) for a run-time 0 <= v < N, so that loops
// bounded by v unroll without branches
template <int N, class F> __device__ __forceinline__ void with_const(int v, F &&f) {
  if constexpr (N > 0) {
    if (v == N - 1) f(std::integral_constant<int, N - 1>{});
    else with_const<N - 1>(v, f);
  }
}
// with_const<N>(v, f) when TRIM, else f(0): the whole range, no branches
template <bool TRIM, int N, class F> __device__ __forceinline__ void from_block(int v, F &&f) {
  if constexpr (TRIM) with_const<N>(v, f);
  else f(std::integral_constant<int, 0>{});
}

// a double rotated within each row of 16 lanes (every lane receives)
template <int CTRL> __device__ __forceinline__ double ror(double x) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
// s[a], A0 <= a < N, summed over each row of TC lanes of the thread grid,
// stage by stage so that the rows' chains interleave; every lane of the row
// holds the sum for TC = 16, lanes 16-31 and 48-63 for TC = 32
template <int TC, int A0, int N> __device__ __forceinline__ void row_sums(double (&s)[N]) {
  static_assert(TC == 16 || TC == 32, "row_sums: 16 or 32 lanes");
  double t[N];
#pragma unroll
  for (int a = A0; a < N; ++a) t[a] = ror<0x128>(s[a]);
#pragma unroll
  for (int a = A0; a < N; ++a) s[a] += t[a];
#pragma unroll
  for (int a = A0; a < N; ++a) t[a] = ror<0x124>(s[a]);
#pragma unroll
  for (int a = A0; a < N; ++a) s[a] += t[a];
#pragma unroll
  for (int a = A0; a < N; ++a) t[a] = ror<0x122>(s[a]);
#pragma unroll
  for (int a = A0; a < N; ++a) s[a] += t[a];
#pragma unroll
  for (int a = A0; a < N; ++a) t[a] = ror<0x121>(s[a]);
#pragma unroll
  for (int a = A0; a < N; ++a) s[a] += t[a];
  if constexpr (TC == 32) {
#pragma unroll
    for (int a = A0; a < N; ++a) s[a] += dpp<0x142, 0xA>(s[a]);
  }
}

// The reduction runs on a TR x TC grid of threads, thread (r, c) holding the
// NR x NC entries (TR a + r, TC b + c) in registers.  Two LDS-only barriers a
// step:
//   1. every wave forms the reflector's scalars from column j in LDS by
//      itself and the v entries it needs on demand (v_i = x_i / (alpha -
//      beta) below j + 1); p = tau A v is summed across the TC column lanes
//      with DPP and its rows written to LDS;
//   2. every wave forms w = p - (tau / 2)(p.v) v for itself, updates its
//      part of the trailing block A -= v w^T + w v^T, and the owners of
//      column j + 1 put it in the other LDS buffer for the next step.
// Blocks wholly above the active part or past m are skipped (the tests are
// uniform over the workgroup).  A grid of 16 x 16 threads (one wave per
// SIMD) serves m <= 144; 32 x 16 (two waves per SIMD) the larger sizes: each
// wave repeats the reflector's scalar work, so fewer waves a SIMD make a
// step shorter while the register tile still fits.
template <int TR, int TC, int NR, int NC, bool TRIM>
__global__ __launch_bounds__(TR *TC) void k_ritz_tridiag(int m, const double *__restrict__ H,
                                                         int ldh, double *__restrict__ work) {
  static_assert(TR * TC >= kRMax, "the reflector store needs a thread per row");
  __shared__ double xs[2][kRMax], ps[kRMax];
  double *dd = work + kOffD, *ee = work + kOffE, *taus = work + kOffTau, *hv = work + kOffHv;
  const int tid = threadIdx.x, r = tid / TC, c = tid % TC, lane = tid & 63;
  const int pown = TC == 16 ? 0 : TC - 1;   // a lane of the row that holds row_sum
  double A[NR][NC];
#pragma unroll
  for (int a = 0; a < NR; ++a)
#pragma unroll
    for (int b = 0; b < NC; ++b) {
      const int i = TR * a + r, l = TC * b + c;
      A[a][b] = (i < m && l < m) ? 0.5 * (H[(long)i * ldh + l] + H[(long)l * ldh + i]) : 0.0;
    }
  if (m > 2 && c == 0) {   // column 0 below the diagonal, and d[0]
#pragma unroll
    for (int a = 0; a < NR; ++a) {
      const int i = TR * a + r;
      if (i > 0 && i < m) xs[0][i] = A[a][0];
      if (i == 0) dd[0] = A[a][0];
    }
  }
  lds_sync();
  for (int j = 0; j + 2 < m; ++j) {
    const double *x = xs[j & 1];
    // every LDS value this thread needs, loaded at once (all indices are
    // below kRMax; entries outside (j, m) are stale and selected away)
    double xsq[3], xc[NC];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int i = j + 2 + lane + 64 * q;
      xsq[q] = x[i < m ? i : 0];
    }
#pragma unroll
    for (int b = 0; b < NC; ++b) xc[b] = x[TC * b + c];
    const double xh = x[tid < kRMax ? tid : 0];
    const double alpha = x[j + 1];
    // the reflector I - tau v v^T with v[j + 1] = 1, H x = beta e_{j+1}
    double s2 = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q)
      if (j + 2 + lane + 64 * q < m) s2 += xsq[q] * xsq[q];
    s2 = wave_sum(s2);
    double t = 0.0, scale = 0.0, beta = alpha;
    if (s2 > 0.0) {
      beta = -copysign(sqrt(alpha * alpha + s2), alpha);
      t = (beta - alpha) * recip(beta);
      scale = recip(alpha - beta);
    }
    // v_i from the column's entry x_i
    auto vof = [&](int i, double xi) {
      const double v = i == j + 1 ? 1.0 : xi * scale;
      return (t == 0.0 || i <= j || i >= m) ? 0.0 : v;
    };
    if (tid < kRMax) {
      hv[(long)j * kRMax + tid] = vof(tid, xh);
      if (tid == 0) {
        ee[j] = beta;
        taus[j] = t;
      }
    }
    double vc[NC];
#pragma unroll
    for (int b = 0; b < NC; ++b) vc[b] = vof(TC * b + c, xc[b]);
    // p = tau A v over the rows below j: the blocks from B0 = (j + 1) / TC
    // (column blocks) and A0 = B0 TC / TR (row blocks) on hold row j + 1
    // and all later ones; v = 0 on the columns before them
    const int B0 = (j + 1) / TC;
    from_block<TRIM, NC>(B0, [&](auto bc) {
      constexpr int Bc = decltype(bc)::value, Ac = Bc * TC / TR;
      double s[NR];
#pragma unroll
      for (int a = Ac; a < NR; ++a) {
        s[a] = 0.0;
#pragma unroll
        for (int b = Bc; b < NC; ++b) s[a] = fma(A[a][b], vc[b], s[a]);
      }
      row_sums<TC, Ac>(s);
      if (c == pown) {
#pragma unroll
        for (int a = Ac; a < NR; ++a) ps[TR * a + r] = t * s[a];
      }
    });
    lds_sync();
    double psq[3], pc[NC], pr[NR], xr[NR];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int i = j + 1 + lane + 64 * q;
      psq[q] = ps[i < m ? i : 0];
      xsq[q] = x[i < m ? i : 0];
    }
#pragma unroll
    for (int b = 0; b < NC; ++b) pc[b] = ps[TC * b + c];
#pragma unroll
    for (int a = 0; a < NR; ++a) {
      pr[a] = ps[TR * a + r];
      xr[a] = x[TR * a + r];
    }
    // w = p - (tau / 2)(p.v) v, every wave for itself
    double pv = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int i = j + 1 + lane + 64 * q;
      if (i < m) pv = fma(psq[q], vof(i, xsq[q]), pv);
    }
    const double K = -0.5 * t * wave_sum(pv);
    double wc[NC];
#pragma unroll
    for (int b = 0; b < NC; ++b) {
      const int ic = TC * b + c;
      wc[b] = (ic > j && ic < m) ? fma(K, vc[b], pc[b]) : 0.0;
    }
    // A -= v w^T + w v^T on the trailing block (v = w = 0 outside it);
    // then the owners of column j + 1 put it below the diagonal in the
    // other LDS buffer and d[j + 1], final from here on, in the output
    const int jn = j + 1;
    double *xn = xs[jn & 1];
    from_block<TRIM, NC>(B0, [&](auto bc) {
      constexpr int Bc = decltype(bc)::value, Ac = Bc * TC / TR;
#pragma unroll
      for (int a = Ac; a < NR; ++a) {
        const int ir = TR * a + r;
        const double va = vof(ir, xr[a]);
        const double wa = (ir > j && ir < m) ? fma(K, va, pr[a]) : 0.0;
#pragma unroll
        for (int b = Bc; b < NC; ++b) A[a][b] = fma(-va, wc[b], fma(-wa, vc[b], A[a][b]));
      }
    });
    if (j + 3 < m && c == jn % TC) {
      with_const<NC>(B0, [&](auto bc) {   // column j + 1 lies in block B0
        constexpr int Bc = decltype(bc)::value, Ac = Bc * TC / TR;
        double dj = 0.0;
#pragma unroll
        for (int a = Ac; a < NR; ++a) {
          const int i = TR * a + r;
          xn[i] = A[a][Bc];   // entries outside (jn, m) are never read as such
          dj = i == jn ? A[a][Bc] : dj;
        }
        if (r == jn % TR) dd[jn] = dj;
      });
    }
    lds_sync();
  }
  // the last 2 x 2 (or 1 x 1) block
  for (int q = 0; q < 3; ++q) {
    const int i = q == 0 ? m - 1 : (q == 1 ? m - 2 : m - 1);
    const int l = q == 0 ? m - 1 : (q == 1 ? m - 2 : m - 2);
    if (i < 0 || l < 0) continue;
    if (i % TR == r && l % TC == c) {
      const double v = tile_get(A, i / TR, l / TC);
      if (q < 2) dd[i] = v;
      else ee[m - 2] = v;
    }
  }
}

// Sturm count: eigenvalues of the tridiagonal (d, e^2) below x
__device__ int sturm_count(int m, const double *d, const double *e2, double x, double pivmin) {
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  int cnt = q < 0.0;
  int i = 1;
  for (; i + 3 < m; i += 4) {   // the next four entries' loads ahead of the chain
    double dv[4], fv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      dv[u] = d[i + u] - x;
      fv[u] = e2[i + u - 1];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      q = dv[u] - fv[u] * recip(q);
      if (fabs(q) < pivmin) q = -pivmin;
      cnt += q < 0.0;
    }
  }
  for (; i < m; ++i) {
    q = (d[i] - x) - e2[i - 1] * recip(q);
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0.0;
  }
  return cnt;
}

__global__ __launch_bounds__(64) void k_ritz_tri_eig(int m, double *__restrict__ work,
                                                     double *__restrict__ wq) {
  __shared__ double d[kRMax], e[kRMax], e2[kRMax];
  __shared__ double u0[kRMax], u1[kRMax], u2[kRMax], ml[kRMax], xv[kRMax];
  __shared__ int piv[kRMax];
  const int g = blockIdx.x, lane = threadIdx.x;
  const double *dd = work + kOffD, *ee = work + kOffE;
  for (int i = lane; i < m; i += 64) {
    d[i] = dd[i];
    e[i] = i + 1 < m ? ee[i] : 0.0;
    e2[i] = e[i] * e[i];
  }
  __syncthreads();
  // Gershgorin interval and the pivot floor
  double lo = DBL_MAX, hi = -DBL_MAX, emax = 0.0;
  for (int i = lane; i < m; i += 64) {
    const double rad = fabs(e[i]) + (i > 0 ? fabs(e[i - 1]) : 0.0);
    lo = fmin(lo, d[i] - rad);
    hi = fmax(hi, d[i] + rad);
    emax = fmax(emax, e2[i]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, o, 64));
    hi = fmax(hi, __shfl_xor(hi, o, 64));
    emax = fmax(emax, __shfl_xor(emax, o, 64));
  }
  const double tnorm = fmax(fabs(lo), fabs(hi));
  const double pivmin = DBL_MIN * fmax(1.0, emax);
  const double pad = 4.0 * DBL_EPSILON * tnorm + 2.0 * pivmin;
  double a = lo - pad, b = hi + pad;
  // multisection: keep count(a) <= g < count(b)
  for (int it = 0; it < 24; ++it) {
    if (b - a <= 2.0 * DBL_EPSILON * fmax(fabs(a), fabs(b)) + 2.0 * pivmin) break;
    const double x = a + (b - a) * ((double)(lane + 1) / 65.0);
    const int cnt = sturm_count(m, d, e2, x, pivmin);
    const unsigned long long bal = __ballot(cnt >= g + 1);
    if (bal == 0ull) {
      a = __shfl(x, 63, 64);
    } else {
      const int f = __ffsll((long long)bal) - 1;
      const double nb = __shfl(x, f, 64);
      const double na = __shfl(x, f > 0 ? f - 1 : 0, 64);
      if (f > 0) a = na;
      b = nb;
    }
  }
  const double lam = 0.5 * (a + b);
  if (lane == 0) {
    wq[g] = lam;
    if (g == 0) work[kOffNorm] = tnorm;
  }
  // inverse iteration on T - lam I: LU with partial pivoting (one lane),
  // the pivots' reciprocals (all lanes), two solves (one lane, the running
  // entries carried in registers, the next one loaded a step ahead) each
  // followed by a scaling to unit max-norm (all lanes)
  const double tiny = DBL_EPSILON * fmax(tnorm, DBL_MIN);
  if (lane == 0) {
    double p0 = d[0] - lam, p1 = m > 1 ? e[0] : 0.0, p2 = 0.0;
#pragma unroll 2
    for (int k = 0; k + 1 < m; ++k) {
      const double r0 = e[k], r1 = d[k + 1] - lam, r2 = k + 2 < m ? e[k + 1] : 0.0;
      if (fabs(p0) >= fabs(r0)) {
        const double l = p0 != 0.0 ? r0 / p0 : 0.0;
        u0[k] = p0; u1[k] = p1; u2[k] = p2; ml[k] = l; piv[k] = 0;
        p0 = r1 - l * p1;
        p1 = r2 - l * p2;
      } else {
        const double l = p0 / r0;
        u0[k] = r0; u1[k] = r1; u2[k] = r2; ml[k] = l; piv[k] = 1;
        p0 = p1 - l * r1;
        p1 = p2 - l * r2;
      }
      p2 = 0.0;
    }
    u0[m - 1] = p0; u1[m - 1] = 0.0; u2[m - 1] = 0.0;
  }
  __syncthreads();
  for (int k = lane; k < m; k += 64) {
    double u = u0[k];
    if (fabs(u) < tiny) u = u < 0.0 ? -tiny : tiny;
    u0[k] = 1.0 / u;   // the reciprocal from here on
    // a fixed start with no special direction
    xv[k] = 0.5 + (double)((k * 7919 + g * 104729 + 13) % 1009) / 1009.0;
  }
  __syncthreads();
  for (int step = 0; step < 2; ++step) {
    if (lane == 0) {
      // L: the row swaps and eliminations in order
      double cur = xv[0];
      for (int k = 0; k + 1 < m; ++k) {
        double nxt = xv[k + 1];
        if (piv[k]) {
          const double s = cur;
          cur = nxt;
          nxt = s;
        }
        xv[k] = cur;
        cur = nxt - ml[k] * cur;
      }
      xv[m - 1] = cur;
      // U: back substitution; a growing solution is rescaled, the entries
      // still to be solved (and the one already loaded) with it
      double x1 = 0.0, x2 = 0.0;   // x[k + 1], x[k + 2]
      double xk = xv[m - 1];
      for (int k = m - 1; k >= 0; --k) {
        double xn = k > 0 ? xv[k - 1] : 0.0;
        double s = (xk - u1[k] * x1 - u2[k] * x2) * u0[k];
        if (fabs(s) > 1e100) {
          for (int i = 0; i < k; ++i) xv[i] *= 1e-100;
          for (int i = k + 1; i < m; ++i) xv[i] *= 1e-100;
          x1 *= 1e-100;
          x2 *= 1e-100;
          s *= 1e-100;
          xn *= 1e-100;
        }
        xv[k] = s;
        x2 = x1;
        x1 = s;
        xk = xn;
      }
    }
    __syncthreads();
    double mx = 0.0;
    for (int i = lane; i < m; i += 64) mx = fmax(mx, fabs(xv[i]));
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    const double inv = mx > 0.0 ? 1.0 / mx : 1.0;
    for (int i = lane; i < m; i += 64) xv[i] *= inv;
    __syncthreads();
  }
  double nn = 0.0;
  for (int i = lane; i < m; i += 64) nn += xv[i] * xv[i];
  const double inv = 1.0 / sqrt(wave_sum(nn));
  double *z = work + kOffZ + (long)g * kRMax;
  for (int i = lane; i < m; i += 64) z[i] = xv[i] * inv;
}

__global__ __launch_bounds__(1024) void k_ritz_back(int m, int k, const double *__restrict__ work,
                                                    const double *__restrict__ wq,
                                                    double *__restrict__ Y, int ldy) {
  __shared__ double zs[kRK * kZLd];
  __shared__ double vb[2][8 * kRMax], tb[kRMax];
  __shared__ double dots[kRK], nrm[1];
  const int tid = threadIdx.x, grp = tid >> 4, gl = tid & 15;
  const double *hv = work + kOffHv, *taus = work + kOffTau, *Z = work + kOffZ;
  const double tnorm = work[kOffNorm];
  for (int e = tid; e < k * kRMax; e += 1024) {
    const int g = e / kRMax, i = e % kRMax;
    zs[g * kZLd + i] = i < m ? Z[(long)g * kRMax + i] : 0.0;
  }
  __syncthreads();
  // clusters: two classical Gram-Schmidt passes of z_g against the earlier
  // members of its cluster, then normalisation
  int start = 0;
  for (int g = 1; g < k; ++g) {
    if (wq[g] - wq[g - 1] > 1e-6 * tnorm) {
      start = g;
      continue;
    }
    for (int pass = 0; pass < 2; ++pass) {
      // dots with members h = start .. g-1: 16 lanes per member
      const int h = start + grp;
      double s = 0.0;
      if (h < g)
        for (int i = gl; i < m; i += 16) s += zs[h * kZLd + i] * zs[g * kZLd + i];
      s = sum16(s);
      if (h < g && gl == 0) dots[grp] = s;
      __syncthreads();
      if (tid < m) {
        double v = zs[g * kZLd + tid];
        for (int hh = start; hh < g; ++hh) v -= dots[hh - start] * zs[hh * kZLd + tid];
        zs[g * kZLd + tid] = v;
      }
      __syncthreads();
    }
    if (tid < 64) {
      double s = 0.0;
      for (int i = tid; i < m; i += 64) s += zs[g * kZLd + i] * zs[g * kZLd + i];
      s = wave_sum(s);
      if (tid == 0) nrm[0] = 1.0 / sqrt(s);
    }
    __syncthreads();
    if (tid < m) zs[g * kZLd + tid] *= nrm[0];
    __syncthreads();
  }
  // y = H_0 H_1 ... H_{m-3} z, the reflectors from the last to the first:
  // 16 lanes per eigenvector with its entries i = gl + 16 q in registers;
  // the reflectors pass through LDS eight at a time, the next eight loaded
  // while these are applied
  const int nref = m - 2;   // reflectors j = 0 .. m-3, applied as jj = m-3-j
  double z[kRB * 2];
  if (grp < k) {
#pragma unroll
    for (int q = 0; q < 2 * kRB; ++q) z[q] = zs[grp * kZLd + gl + 16 * q];
  }
  for (int i = tid; i < nref; i += 1024) tb[i] = taus[i];
  auto fetch = [&](int blk, double (&reg)[2]) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int e = tid + 1024 * u, jj = 8 * blk + e / kRMax, i = e % kRMax;
      reg[u] = (e < 8 * kRMax && jj < nref) ? hv[(long)(nref - 1 - jj) * kRMax + i] : 0.0;
    }
  };
  auto stash = [&](int buf, const double (&reg)[2]) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int e = tid + 1024 * u;
      if (e < 8 * kRMax) vb[buf][e] = reg[u];
    }
  };
  const int nblk = nref > 0 ? (nref + 7) / 8 : 0;
  double reg[2];
  if (nblk > 0) {
    fetch(0, reg);
    stash(0, reg);
  }
  lds_sync();
  for (int blk = 0; blk < nblk; ++blk) {
    const bool more = blk + 1 < nblk;
    if (more) fetch(blk + 1, reg);
    if (grp < k) {
      const double *vv = vb[blk & 1];
      for (int u = 0; u < 8 && 8 * blk + u < nref; ++u) {
        const int jj = 8 * blk + u;
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 2 * kRB; ++q) s += vv[u * kRMax + gl + 16 * q] * z[q];
        s = tb[nref - 1 - jj] * sum16(s);
#pragma unroll
        for (int q = 0; q < 2 * kRB; ++q) z[q] -= s * vv[u * kRMax + gl + 16 * q];
      }
    }
    if (more) stash((blk + 1) & 1, reg);
    lds_sync();
  }
  if (grp < k) {
#pragma unroll
    for (int q = 0; q < 2 * kRB; ++q) zs[grp * kZLd + gl + 16 * q] = z[q];
  }
  __syncthreads();
  for (int e = tid; e < k * m; e += 1024) {
    const int i = e / k, g = e % k;
    Y[(long)i * ldy + g] = zs[g * kZLd + i];
  }
}

// y_g^T G y_g and y_g^T H y_g, one workgroup per eigenvector: lane l keeps
// the column sums sum_i y_i G_il over its columns l = lane + 64 q for every
// fourth row i (the rows read whole by a wave), and the products with y_l
// are summed over the wave once, then over the four waves in order
__global__ __launch_bounds__(256) void k_ritz_quad(int m, int k, const double *__restrict__ H,
                                                   int ldh, const double *__restrict__ G,
                                                   int ldg, const double *__restrict__ Y,
                                                   int ldy, double *__restrict__ wq) {
  __shared__ double y[kRMax], red[2][4];
  const int g = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int i = tid; i < m; i += 256) y[i] = Y[(long)i * ldy + g];
  __syncthreads();
  constexpr int NQ = (kRMax + 63) / 64;
  double cg[NQ], ch[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) cg[q] = ch[q] = 0.0;
#pragma unroll 4
  for (int i = wave; i < m; i += 4) {
    const double yi = y[i];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int l = lane + 64 * q;
      if (l < m) {
        cg[q] = fma(yi, G[(long)i * ldg + l], cg[q]);
        ch[q] = fma(yi, H[(long)i * ldh + l], ch[q]);
      }
    }
  }
  double sg = 0.0, sh = 0.0;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int l = lane + 64 * q;
    if (l < m) {
      sg = fma(cg[q], y[l], sg);
      sh = fma(ch[q], y[l], sh);
    }
  }
  sg = wave_sum(sg);
  sh = wave_sum(sh);
  if (lane == 0) {
    red[0][wave] = sg;
    red[1][wave] = sh;
  }
  __syncthreads();
  if (tid == 0) {
    wq[k + g] = ((red[0][0] + red[0][1]) + red[0][2]) + red[0][3];
    wq[2 * k + g] = ((red[1][0] + red[1][1]) + red[1][2]) + red[1][3];
  }
}

}  // namespace
}  // namespace mbpt

using namespace mbpt;

extern "C" {

int mbpt_ritz_max_m(void) { return kRMax; }

long mbpt_ritz_eigh_f64_work(int m, int k) {
  (void)m;
  (void)k;
  return kRitzWork;
}

int mbpt_ritz_eigh_f64(int m, int k, const double *H, int ldh, const double *G, int ldg,
                       double *Y, int ldy, double *wq, double *work, long work_elems,
                       void *stream) {
  MBPT_REQUIRE(m >= 1 && m <= kRMax, "mbpt_ritz_eigh_f64: need 1 <= m <= 192");
  MBPT_REQUIRE(k >= 1 && k <= m && k <= kRK, "mbpt_ritz_eigh_f64: need 1 <= k <= min(m, 64)");
  MBPT_REQUIRE(H && G && Y && wq && work, "mbpt_ritz_eigh_f64: NULL argument");
  MBPT_REQUIRE(ldh >= m && ldg >= m && ldy >= k, "mbpt_ritz_eigh_f64: leading dimension too small");
  MBPT_REQUIRE(work_elems >= kRitzWork, "mbpt_ritz_eigh_f64: work buffer too small");
  hipStream_t st = (hipStream_t)stream;
  switch ((m + 15) / 16) {
    case 1: k_ritz_tridiag<16, 16, 1, 1, true><<<1, 256, 0, st>>>(m, H, ldh, work); break;
    case 2: k_ritz_tridiag<16, 16, 2, 2, true><<<1, 256, 0, st>>>(m, H, ldh, work); break;
    case 3: k_ritz_tridiag<16, 16, 3, 3, true><<<1, 256, 0, st>>>(m, H, ldh, work); break;
    case 4: k_ritz_tridiag<16, 16, 4, 4, true><<<1, 256, 0, st>>>(m, H, ldh, work); break;
    case 5: k_ritz_tridiag<16, 16, 5, 5, true><<<1, 256, 0, st>>>(m, H, ldh, work); break;
    case 6: k_ritz_tridiag<16, 16, 6, 6, true><<<1, 256, 0, st>>>(m, H, ldh, work); break;
    case 7: k_ritz_tridiag<16, 16, 7, 7, true><<<1, 256, 0, st>>>(m, H, ldh, work); break;
    case 8: k_ritz_tridiag<16, 16, 8, 8, true><<<1, 256, 0, st>>>(m, H, ldh, work); break;
    case 9: k_ritz_tridiag<16, 16, 9, 9, true><<<1, 256, 0, st>>>(m, H, ldh, work); break;
    case 10: k_ritz_tridiag<16, 16, 10, 10, true><<<1, 256, 0, st>>>(m, H, ldh, work); break;
    case 11: k_ritz_tridiag<32, 16, 6, 11, false><<<1, 512, 0, st>>>(m, H, ldh, work); break;
    default: k_ritz_tridiag<32, 16, 6, 12, false><<<1, 512, 0, st>>>(m, H, ldh, work); break;
  }
  int rc = launch_status("k_ritz_tridiag");
  if (rc) return rc;
  k_ritz_tri_eig<<<k, 64, 0, st>>>(m, work, wq);
  if ((rc = launch_status("k_ritz_tri_eig"))) return rc;
  k_ritz_back<<<1, 1024, 0, st>>>(m, k, work, wq, Y, ldy);
  if ((rc = launch_status("k_ritz_back"))) return rc;
  k_ritz_quad<<<k, 256, 0, st>>>(m, k, H, ldh, G, ldg, Y, ldy, wq);
  return launch_status("k_ritz_quad");
}

}  // extern "C"
