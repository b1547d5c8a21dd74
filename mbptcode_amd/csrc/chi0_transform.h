// The cosine transform of kept per-time-point chi0 terms onto imaginary
// frequencies, gfx950:
//
//   out[f] = [acc ? out[f] : 0] + sum_t w[f][t] P(t),   P(t) = -2 Go(t) o Gv(t)
//
// Each P(t) is built once (by the main chi0 kernel with one time point per
// chunk) and every frequency reads it here, so a sweep over n_w frequencies
// costs n_tau Green's-function products instead of n_w n_tau.  The transform
// is a streaming pass: a thread holds its elements' n_tau term values in
// registers and writes every frequency's FP64 sum from them, so each term
// is read from HBM once per launch.  The weights travel in the kernel
// arguments (a launch takes as many frequencies as fit).
#pragma once

#include "common.h"

namespace mbpt {

constexpr int kFreqW = 376;   // weights per launch: 3 KB of kernel arguments

struct FreqW {
  int nf, nt;                 // frequencies in this launch, time points
  double w[kFreqW];           // w[f * nt + t]
};

// FP32 terms in the main kernel's layout: P(t) at terms + t Mp^2, row-major
// with row stride Mp, valid on the 128 x 128 tiles (I, J) with I <= J (the
// diagonal tiles whole).  One block per 32 x 32 source tile; a thread owns
// two adjacent columns of one row.  Off-diagonal tiles are also written
// mirrored, through an LDS transpose so both writes stay coalesced.
template <int NT>
__global__ __launch_bounds__(512) void k_chi0_transform_f32(int M, int Mp, int nt128,
                                                            const float *__restrict__ terms,
                                                            FreqW fw, double *__restrict__ out,
                                                            int accumulate) {
  const int p = blockIdx.x >> 4, s = blockIdx.x & 15;
  int I = 0, base = 0;
  while (p >= base + (nt128 - I)) { base += nt128 - I; ++I; }
  const int J = I + (p - base);
  const int r0 = (4 * I + (s >> 2)) * 32, c0 = (4 * J + (s & 3)) * 32;
  if (r0 >= M || c0 >= M) return;
  const int tid = threadIdx.x, r = tid >> 4, c = (tid & 15) * 2;
  const long plane = (long)Mp * Mp;
  const float *src = terms + (long)(r0 + r) * Mp + c0 + c;
  float2 v[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t)
    v[t] = t < fw.nt ? *(const float2 *)(src + t * plane) : make_float2(0.f, 0.f);
  __shared__ double tr[32][33];
  const bool mirror = I != J;
  const int row = r0 + r, col = c0 + c;
  const long MM = (long)M * M;
  for (int f = 0; f < fw.nf; ++f) {
    const double *w = fw.w + f * fw.nt;
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int t = 0; t < NT; ++t)
      if (t < fw.nt) {
        s0 += w[t] * (double)v[t].x;
        s1 += w[t] * (double)v[t].y;
      }
    double *o = out + f * MM;
    if (row < M) {
      const long e = (long)row * M + col;
      if (col < M) o[e] = accumulate ? o[e] + s0 : s0;
      if (col + 1 < M) o[e + 1] = accumulate ? o[e + 1] + s1 : s1;
    }
    if (mirror) {
      tr[r][c] = s0;
      tr[r][c + 1] = s1;
      __syncthreads();
      // element (c0 + r, r0 + c) of the lower triangle is (r0 + c, c0 + r)
      const int orow = c0 + r, ocol = r0 + c;
      if (orow < M) {
        const long e = (long)orow * M + ocol;
        if (ocol < M) o[e] = accumulate ? o[e] + tr[c][r] : tr[c][r];
        if (ocol + 1 < M) o[e + 1] = accumulate ? o[e + 1] + tr[c + 1][r] : tr[c + 1][r];
      }
      __syncthreads();
    }
  }
}

// FP64 terms: P(t) at terms + t M^2, whole symmetric matrices; elementwise.
template <int NT>
__global__ __launch_bounds__(256) void k_chi0_transform_f64(long n, const double *__restrict__ terms,
                                                            FreqW fw, double *__restrict__ out,
                                                            int accumulate) {
  for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
    double v[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) v[t] = t < fw.nt ? terms[t * n + e] : 0.0;
    for (int f = 0; f < fw.nf; ++f) {
      const double *w = fw.w + f * fw.nt;
      double s = 0.0;
#pragma unroll
      for (int t = 0; t < NT; ++t)
        if (t < fw.nt) s += w[t] * v[t];
      double *o = out + f * n;
      o[e] = accumulate ? o[e] + s : s;
    }
  }
}

}  // namespace mbpt
