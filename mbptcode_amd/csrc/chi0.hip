// Space-time polarizability on the interpolation grid, gfx950.
//
//   out = [acc ? out : 0] + sum_t w_t (-2) Go(t) o Gv(t)
//   Go(t) = (X_o o e^{e_o tau_t}) X_o^T     Gv(t) = (X_v o e^{-e_v tau_t}) X_v^T
//
// One workgroup owns a 64x64 tile of the symmetric output (upper-triangular
// tiles only) and loops over every time point: per point the two Green's
// function tiles are accumulated by MFMA from LDS-staged operands (the
// exponential scaling is applied while staging, so it costs no extra pass),
// multiplied in registers and folded into the tile's running sum.  The output
// is written once; off-diagonal tiles are mirrored into the lower triangle.
#include "common.h"
#include "chi0_f32.h"
#include "chi0_transform.h"
#include "gw_sigma_f32.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace mbpt {

struct TauSet {
  int n;
  double w[MBPT_CHI0_MAX_TAU];
};

// s_o[t][i] = exp(e_o[i] tau_t), s_v[t][a] = exp(-e_v[a] tau_t)
struct TauPoints {
  int n;
  double tau[MBPT_CHI0_MAX_TAU];
};

template <typename T>
__global__ void k_exp_tables(int no, int nv, const double *__restrict__ eo,
                             const double *__restrict__ ev, TauPoints tp,
                             T *__restrict__ so, T *__restrict__ sv) {
  const int norb = no + nv;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= tp.n * norb) return;
  const int t = idx / norb, p = idx % norb;
  if (p < no)
    so[(long)t * no + p] = (T)exp(eo[p] * tp.tau[t]);
  else
    sv[(long)t * nv + (p - no)] = (T)exp(-ev[p - no] * tp.tau[t]);
}

template <typename T> struct LdScaled {  // A(m = mu, k = i) = X[mu][i] s[i]
  const T *X, *s;
  int ld, rows, K;
  __device__ T operator()(int m, int k) const {
    return (m < rows && k < K) ? X[(long)m * ld + k] * s[k] : T(0);
  }
};
template <typename T> struct LdPlainT {  // B(k = i, n = nu) = X[nu][i]
  const T *X;
  int ld, K, cols;
  __device__ T operator()(int k, int n) const {
    return (k < K && n < cols) ? X[(long)n * ld + k] : T(0);
  }
};

__device__ __forceinline__ void tri_decode(int lid, int nt, int &ti, int &tj) {
  // lid enumerates (ti <= tj) row by row
  int r = 0, base = 0;
  while (lid >= base + (nt - r)) { base += nt - r; ++r; }
  ti = r;
  tj = r + (lid - base);
}

template <typename T>
__global__ __launch_bounds__(256) void k_chi0(int M, int no, int nv,
                                              const T *__restrict__ Xo,
                                              const T *__restrict__ Xv,
                                              const T *__restrict__ so,
                                              const T *__restrict__ sv, TauSet ts,
                                              double *__restrict__ out, int accumulate,
                                              int nt, int nblk) {
  const int lid = xcd_logical(blockIdx.x, nblk);
  if (lid >= nt * (nt + 1) / 2) return;
  int ti, tj;
  tri_decode(lid, nt, ti, tj);
  const int m0 = ti * kTile, n0 = tj * kTile;
  __shared__ TileSmem<T> sm;
  typename MF<T>::acc chi[2][2], go[2][2], gv[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) zero(chi[i][j]);
  for (int t = 0; t < ts.n; ++t) {
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) { zero(go[i][j]); zero(gv[i][j]); }
    gemm64<T, true, true>(LdScaled<T>{Xo, so + (long)t * no, no, M, no},
                          LdPlainT<T>{Xo, no, no, M}, m0, n0, no, go, sm);
    gemm64<T, true, true>(LdScaled<T>{Xv, sv + (long)t * nv, nv, M, nv},
                          LdPlainT<T>{Xv, nv, nv, M}, m0, n0, nv, gv, sm);
    const T c = (T)(-2.0 * ts.w[t]);
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j)
        for (int r = 0; r < 4; ++r) chi[i][j][r] += c * go[i][j][r] * gv[i][j][r];
  }
  for (int fm = 0; fm < 2; ++fm)
    for (int fn = 0; fn < 2; ++fn) {
      const int col = tile_col(n0, fn);
      for (int r = 0; r < 4; ++r) {
        const int row = tile_row<T>(m0, fm, r);
        if (row < M && col < M) {
          const double v = (double)chi[fm][fn][r];
          const long o = (long)row * M + col;
          out[o] = accumulate ? out[o] + v : v;
          if (ti != tj) {
            const long ot = (long)col * M + row;
            out[ot] = accumulate ? out[ot] + v : v;
          }
        }
      }
    }
}

template <typename T>
__global__ void k_split_X2(int M, int no, int nv, const double *__restrict__ X,
                           T *__restrict__ Xo, T *__restrict__ Xv) {
  const long n = (long)M * (no + nv);
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x) {
    const int mu = (int)(e / (no + nv)), p = (int)(e % (no + nv));
    if (p < no)
      Xo[(long)mu * no + p] = (T)X[e];
    else
      Xv[(long)mu * nv + (p - no)] = (T)X[e];
  }
}

}  // namespace mbpt

using namespace mbpt;

struct mbpt_chi0 {
  int M, no, nv, prec;
  size_t esz;
  void *Xo, *Xv, *so, *sv;
  double *eo, *ev;
  int prof;
  hipEvent_t e0, e1;
  double prof_ms, prof_flops;
  float *Yt, *part, *XT;         // FP32 path: scaled branches, per-chunk sums, X K-major
  size_t cap_yt, cap_part;       // bytes
  long prof_launches;
  // GW self-energy: packed state columns, their column indices, tile partials
  float *gw_xs, *gw_part;
  int *gw_cols;
  size_t cap_xs, cap_gwpart, cap_cols;
  double gw_ms, gw_flops;
  long gw_launches;
  // kept per-time-point terms P(t) = -2 Go(t) o Gv(t) for the frequency
  // transform: FP32 in the main kernel's tile layout (Mp^2 per point) or
  // FP64 whole matrices (M^2 per point); terms_nt points held
  void *terms;
  size_t cap_terms;
  int terms_nt;
  double tr_ms;
  long tr_launches;
};

namespace {

template <typename T>
int build_impl(mbpt_chi0 *h, int ntau, const double *tau, const double *w, double *out,
               int accumulate, hipStream_t st) {
  TauPoints tp;
  TauSet ts;
  tp.n = ts.n = ntau;
  for (int t = 0; t < ntau; ++t) {
    tp.tau[t] = tau[t];
    ts.w[t] = w[t];
  }
  const int norb = h->no + h->nv;
  k_exp_tables<T><<<(ntau * norb + 255) / 256, 256, 0, st>>>(
      h->no, h->nv, h->eo, h->ev, tp, (T *)h->so, (T *)h->sv);
  int rc = launch_status("k_exp_tables");
  if (rc) return rc;
  const int nt = (h->M + kTile - 1) / kTile;
  const int nblk = pad_grid((long)nt * (nt + 1) / 2);
  if (h->prof) MBPT_HIP_TRY(hipEventRecord(h->e0, st));
  k_chi0<T><<<nblk, 256, 0, st>>>(h->M, h->no, h->nv, (const T *)h->Xo,
                                  (const T *)h->Xv, (const T *)h->so,
                                  (const T *)h->sv, ts, out, accumulate, nt, nblk);
  if ((rc = launch_status("k_chi0"))) return rc;
  if (h->prof) {
    MBPT_HIP_TRY(hipEventRecord(h->e1, st));
    MBPT_HIP_TRY(hipEventSynchronize(h->e1));
    float ms = 0.f;
    MBPT_HIP_TRY(hipEventElapsedTime(&ms, h->e0, h->e1));
    h->prof_ms += ms;
    h->prof_flops += (double)ntau * (double)h->M * (h->M + 1.0) * (norb + 1.0);
    h->prof_launches += 1;
  }
  return MBPT_OK;
}

int grow(float *&buf, size_t &cap, size_t bytes) {
  if (bytes <= cap) return MBPT_OK;
  (void)hipFree(buf);
  buf = nullptr;
  cap = 0;
  MBPT_HIP_TRY(hipMalloc(&buf, bytes));
  cap = bytes;
  return MBPT_OK;
}

// time-point chunks per tile.  A split into c chunks takes
// ceil(ntiles c / 256) rounds of one workgroup per CU, each ceil(ntau / c)
// time points long, plus c partial planes to write and sum (about half a
// time-point round each).  The split with the least of that is taken.  Two
// workgroups fit per CU (the kernel is held to 128 VGPRs), but counting 512
// slots chose more chunks and measured 3-9% slower: more partial sums for a
// tail the second resident workgroup already shortens.
int chi0_chunks(int ntiles, int ntau) {
  const int slots = 256;
  int best = 1;
  double best_cost = 1e300;
  for (int c = 1; c <= std::min(ntau, 8); ++c) {
    const long rounds = ((long)ntiles * c + slots - 1) / slots;
    const double cost = (double)rounds * ((ntau + c - 1) / c) + 0.5 * c;
    if (cost < best_cost - 1e-9) { best_cost = cost; best = c; }
  }
  return best;
}

int build_f32(mbpt_chi0 *h, int ntau, const double *tau, const double *w, double *out,
              int accumulate, hipStream_t st) {
  const int M = h->M, Mp = (M + kCT - 1) / kCT * kCT;
  const int ko = (h->no + 15) / 16 * 16, kv = (h->nv + 15) / 16 * 16;
  const int nt = Mp / kCT, ntiles = nt * (nt + 1) / 2;
  const int nchunk = chi0_chunks(ntiles, ntau);
  const int tpc = (ntau + nchunk - 1) / nchunk;
  int rc = grow(h->Yt, h->cap_yt, sizeof(float) * (size_t)ntau * (ko + kv) * Mp);
  if (rc) return rc;
  rc = grow(h->part, h->cap_part, sizeof(float) * (size_t)nchunk * Mp * Mp);
  if (rc) return rc;
  TauW tw;
  tw.n = ntau;
  for (int t = 0; t < ntau; ++t) {
    tw.tau_half[t] = 0.5 * tau[t];
    tw.w[t] = (float)w[t];
  }
  if (h->prof) MBPT_HIP_TRY(hipEventRecord(h->e0, st));
  k_chi0_prep<<<ntau * (ko + kv), 256, 0, st>>>(M, Mp, h->no, h->nv, ko, kv, h->XT,
                                                  h->eo, h->ev, tw, h->Yt);
  if ((rc = launch_status("k_chi0_prep"))) return rc;
  const int nblk = pad_grid((long)ntiles * nchunk);
  k_chi0_f32<<<nblk, 512, 0, st>>>(Mp, ko, kv, nt, nchunk, tpc, h->Yt, tw, h->part, nblk);
  if ((rc = launch_status("k_chi0_f32"))) return rc;
  const int nto = (M + 63) / 64;
  k_chi0_reduce<<<nto * (nto + 1) / 2, 256, 0, st>>>(M, Mp, nchunk, nto, h->part, out,
                                                       accumulate);
  if ((rc = launch_status("k_chi0_reduce"))) return rc;
  if (h->prof) {
    MBPT_HIP_TRY(hipEventRecord(h->e1, st));
    MBPT_HIP_TRY(hipEventSynchronize(h->e1));
    float ms = 0.f;
    MBPT_HIP_TRY(hipEventElapsedTime(&ms, h->e0, h->e1));
    h->prof_ms += ms;
    h->prof_flops += (double)ntau * (double)M * (M + 1.0) * (h->no + h->nv + 1.0);
    h->prof_launches += 1;
  }
  return MBPT_OK;
}

template <typename B>
int grow_bytes(B *&buf, size_t &cap, size_t bytes);

// P(t) for every time point, kept for mbpt_chi0_transform.  FP32: the main
// kernel with one time point per chunk and unit weights writes each P(t)
// into its own plane.  FP64: one fused launch per time point.
int terms_f32(mbpt_chi0 *h, int ntau, const double *tau, hipStream_t st) {
  const int M = h->M, Mp = (M + kCT - 1) / kCT * kCT;
  const int ko = (h->no + 15) / 16 * 16, kv = (h->nv + 15) / 16 * 16;
  const int nt = Mp / kCT, ntiles = nt * (nt + 1) / 2;
  int rc = grow(h->Yt, h->cap_yt, sizeof(float) * (size_t)ntau * (ko + kv) * Mp);
  if (!rc) rc = grow_bytes(h->terms, h->cap_terms, sizeof(float) * (size_t)ntau * Mp * Mp);
  if (rc) return rc;
  TauW tw;
  tw.n = ntau;
  for (int t = 0; t < ntau; ++t) {
    tw.tau_half[t] = 0.5 * tau[t];
    tw.w[t] = 1.f;
  }
  if (h->prof) MBPT_HIP_TRY(hipEventRecord(h->e0, st));
  k_chi0_prep<<<ntau * (ko + kv), 256, 0, st>>>(M, Mp, h->no, h->nv, ko, kv, h->XT,
                                                  h->eo, h->ev, tw, h->Yt);
  if ((rc = launch_status("k_chi0_prep"))) return rc;
  const int nblk = pad_grid((long)ntiles * ntau);
  k_chi0_f32<<<nblk, 512, 0, st>>>(Mp, ko, kv, nt, ntau, 1, h->Yt, tw, (float *)h->terms,
                                   nblk);
  if ((rc = launch_status("k_chi0_f32"))) return rc;
  if (h->prof) {
    MBPT_HIP_TRY(hipEventRecord(h->e1, st));
    MBPT_HIP_TRY(hipEventSynchronize(h->e1));
    float ms = 0.f;
    MBPT_HIP_TRY(hipEventElapsedTime(&ms, h->e0, h->e1));
    h->prof_ms += ms;
    h->prof_flops += (double)ntau * (double)M * (M + 1.0) * (h->no + h->nv + 1.0);
    h->prof_launches += 1;
  }
  return MBPT_OK;
}

int terms_f64(mbpt_chi0 *h, int ntau, const double *tau, hipStream_t st) {
  const size_t plane = (size_t)h->M * h->M;
  int rc = grow_bytes(h->terms, h->cap_terms, sizeof(double) * ntau * plane);
  if (rc) return rc;
  const double one = 1.0;
  for (int t = 0; t < ntau; ++t) {
    rc = build_impl<double>(h, 1, tau + t, &one, (double *)h->terms + t * plane, 0, st);
    if (rc) return rc;
  }
  return MBPT_OK;
}

template <int NT>
int transform_launch(mbpt_chi0 *h, const FreqW &fw, double *out, int accumulate,
                     hipStream_t st) {
  const int M = h->M;
  if (h->prec == MBPT_FP32) {
    const int Mp = (M + kCT - 1) / kCT * kCT, nt128 = Mp / kCT;
    const int nblk = 16 * nt128 * (nt128 + 1) / 2;
    k_chi0_transform_f32<NT><<<nblk, 512, 0, st>>>(M, Mp, nt128, (const float *)h->terms,
                                                   fw, out, accumulate);
    return launch_status("k_chi0_transform_f32");
  }
  const long n = (long)M * M;
  const int nblk = (int)std::min<long>((n + 255) / 256, 8192L);
  k_chi0_transform_f64<NT><<<nblk, 256, 0, st>>>(n, (const double *)h->terms, fw, out,
                                                 accumulate);
  return launch_status("k_chi0_transform_f64");
}

int transform_impl(mbpt_chi0 *h, int nfreq, const double *w, double *out, int accumulate,
                   hipStream_t st) {
  const int nt = h->terms_nt;
  const int per = std::max(1, kFreqW / nt);
  const long MM = (long)h->M * h->M;
  if (h->prof) MBPT_HIP_TRY(hipEventRecord(h->e0, st));
  for (int f0 = 0; f0 < nfreq; f0 += per) {
    FreqW fw;
    fw.nt = nt;
    fw.nf = std::min(per, nfreq - f0);
    for (int q = 0; q < fw.nf * nt; ++q) fw.w[q] = w[(long)f0 * nt + q];
    double *o = out + f0 * MM;
    const int rc = nt <= 16   ? transform_launch<16>(h, fw, o, accumulate, st)
                   : nt <= 32 ? transform_launch<32>(h, fw, o, accumulate, st)
                              : transform_launch<64>(h, fw, o, accumulate, st);
    if (rc) return rc;
  }
  if (h->prof) {
    MBPT_HIP_TRY(hipEventRecord(h->e1, st));
    MBPT_HIP_TRY(hipEventSynchronize(h->e1));
    float ms = 0.f;
    MBPT_HIP_TRY(hipEventElapsedTime(&ms, h->e0, h->e1));
    h->tr_ms += ms;
    h->tr_launches += 1;
  }
  return MBPT_OK;
}

template <typename B>
int grow_bytes(B *&buf, size_t &cap, size_t bytes) {
  if (bytes <= cap) return MBPT_OK;
  (void)hipFree(buf);
  buf = nullptr;
  cap = 0;
  MBPT_HIP_TRY(hipMalloc(&buf, bytes));
  cap = bytes;
  return MBPT_OK;
}

int sigma_f32(mbpt_chi0 *h, int ntau, const double *tau, const float *Zt, int ns,
              const int *states, double *out, hipStream_t st) {
  const int M = h->M, Mp = (M + kCT - 1) / kCT * kCT;
  const int ko = (h->no + 15) / 16 * 16, kv = (h->nv + 15) / 16 * 16;
  const int nt = Mp / kCT, ntiles = nt * (nt + 1) / 2;
  const int nsp = (ns + 15) / 16 * 16;
  int rc = grow(h->Yt, h->cap_yt, sizeof(float) * (size_t)ntau * (ko + kv) * Mp);
  if (!rc) rc = grow_bytes(h->gw_xs, h->cap_xs, sizeof(float) * (size_t)Mp * nsp);
  if (!rc) rc = grow_bytes(h->gw_part, h->cap_gwpart,
                           sizeof(float) * (size_t)ntau * ntiles * 2 * nsp);
  if (!rc) rc = grow_bytes(h->gw_cols, h->cap_cols, sizeof(int) * (size_t)ns);
  if (rc) return rc;
  std::vector<int> cols(ns);
  for (int j = 0; j < ns; ++j) cols[j] = states[j] < h->no ? states[j] : ko + (states[j] - h->no);
  MBPT_HIP_TRY(hipMemcpyAsync(h->gw_cols, cols.data(), sizeof(int) * ns, hipMemcpyHostToDevice, st));
  MBPT_HIP_TRY(hipStreamSynchronize(st));   // cols is a host temporary
  TauW tw;
  tw.n = ntau;
  for (int t = 0; t < ntau; ++t) {
    tw.tau_half[t] = 0.5 * tau[t];
    tw.w[t] = 0.f;
  }
  if (h->prof) MBPT_HIP_TRY(hipEventRecord(h->e0, st));
  k_chi0_prep<<<ntau * (ko + kv), 256, 0, st>>>(M, Mp, h->no, h->nv, ko, kv, h->XT, h->eo,
                                                  h->ev, tw, h->Yt);
  if ((rc = launch_status("k_chi0_prep"))) return rc;
  const long nx = (long)Mp * nsp;
  k_gw_pack_states<<<(int)std::min<long>((nx + 255) / 256, 4096L), 256, 0, st>>>(
      M, Mp, ns, nsp, h->gw_cols, h->XT, h->gw_xs);
  if ((rc = launch_status("k_gw_pack_states"))) return rc;
  const int nblk = pad_grid((long)ntiles * ntau);
  if (nsp <= 64)
    k_gw_sigma_f32<4><<<nblk, 512, 0, st>>>(M, Mp, ko, kv, nt, ntau, h->Yt, Zt, h->gw_xs, nsp,
                                            h->gw_part, nblk);
  else
    k_gw_sigma_f32<1><<<nblk, 512, 0, st>>>(M, Mp, ko, kv, nt, ntau, h->Yt, Zt, h->gw_xs, nsp,
                                            h->gw_part, nblk);
  if ((rc = launch_status("k_gw_sigma_f32"))) return rc;
  k_gw_sigma_reduce<<<2 * ntau * (nsp / 16), 256, 0, st>>>(ntau, ntiles, ns, nsp, h->gw_part,
                                                            out);
  if ((rc = launch_status("k_gw_sigma_reduce"))) return rc;
  if (h->prof) {
    MBPT_HIP_TRY(hipEventRecord(h->e1, st));
    MBPT_HIP_TRY(hipEventSynchronize(h->e1));
    float ms = 0.f;
    MBPT_HIP_TRY(hipEventElapsedTime(&ms, h->e0, h->e1));
    h->gw_ms += ms;
    h->gw_flops += (double)ntau * (double)M * (M + 1.0) * (h->no + h->nv + 1.0 + 2.0 * ns);
    h->gw_launches += 1;
  }
  return MBPT_OK;
}

template <typename T> int split_impl(mbpt_chi0 *h, const double *X, hipStream_t st) {
  const long n = (long)h->M * (h->no + h->nv);
  k_split_X2<T><<<(int)std::min<long>((n + 255) / 256, 65535L), 256, 0, st>>>(
      h->M, h->no, h->nv, X, (T *)h->Xo, (T *)h->Xv);
  return launch_status("k_split_X2");
}

}  // namespace

extern "C" {

int mbpt_chi0_create(mbpt_chi0 **out, int M, int nocc, int nvir, int precision,
                     const double *X_dev, const double *eps_host, double mu,
                     void *stream) {
  MBPT_REQUIRE(out != nullptr, "mbpt_chi0_create: out is NULL");
  *out = nullptr;
  MBPT_REQUIRE(M > 0 && nocc > 0 && nvir > 0, "mbpt_chi0_create: M, nocc and nvir must be positive");
  MBPT_REQUIRE(X_dev && eps_host, "mbpt_chi0_create: NULL X or eps");
  if (precision != MBPT_FP64 && precision != MBPT_FP32) {
    set_error("mbpt_chi0_create: precision " + std::to_string(precision) +
              " is not built (MBPT_FP64 or MBPT_FP32)");
    return MBPT_ERR_UNSUPPORTED;
  }
  for (int i = 0; i < nocc; ++i)
    MBPT_REQUIRE(eps_host[i] - mu < 0.0, "mbpt_chi0_create: mu must lie above every occupied level");
  for (int a = 0; a < nvir; ++a)
    MBPT_REQUIRE(eps_host[nocc + a] - mu > 0.0, "mbpt_chi0_create: mu must lie below every virtual level");
  hipStream_t st = (hipStream_t)stream;
  mbpt_chi0 *h = new mbpt_chi0();
  h->M = M; h->no = nocc; h->nv = nvir; h->prec = precision;
  h->esz = precision == MBPT_FP64 ? 8 : 4;
  const size_t e = h->esz;
  std::vector<double> eo(nocc), ev(nvir);
  for (int i = 0; i < nocc; ++i) eo[i] = eps_host[i] - mu;
  for (int a = 0; a < nvir; ++a) ev[a] = eps_host[nocc + a] - mu;
  hipError_t he = hipSuccess;
  if (he == hipSuccess) he = hipMalloc(&h->Xo, e * M * nocc);
  if (he == hipSuccess) he = hipMalloc(&h->Xv, e * M * nvir);
  if (he == hipSuccess) he = hipMalloc(&h->so, e * MBPT_CHI0_MAX_TAU * nocc);
  if (he == hipSuccess) he = hipMalloc(&h->sv, e * MBPT_CHI0_MAX_TAU * nvir);
  if (he == hipSuccess) he = hipMalloc(&h->eo, sizeof(double) * nocc);
  if (he == hipSuccess) he = hipMalloc(&h->ev, sizeof(double) * nvir);
  if (he == hipSuccess)
    he = hipMemcpyAsync(h->eo, eo.data(), sizeof(double) * nocc, hipMemcpyHostToDevice, st);
  if (he == hipSuccess)
    he = hipMemcpyAsync(h->ev, ev.data(), sizeof(double) * nvir, hipMemcpyHostToDevice, st);
  if (he == hipSuccess) he = hipStreamSynchronize(st);
  if (he != hipSuccess) {
    set_error(std::string("mbpt_chi0_create: ") + hipGetErrorString(he));
    mbpt_chi0_destroy(h);
    return MBPT_ERR_HIP;
  }
  int rc = precision == MBPT_FP64 ? split_impl<double>(h, X_dev, st)
                                  : split_impl<float>(h, X_dev, st);
  if (!rc && precision == MBPT_FP32) {
    const int ko = (nocc + 15) / 16 * 16, kv = (nvir + 15) / 16 * 16;
    he = hipMalloc(&h->XT, sizeof(float) * (size_t)(ko + kv) * M);
    if (he != hipSuccess) {
      set_error(std::string("mbpt_chi0_create: ") + hipGetErrorString(he));
      rc = MBPT_ERR_HIP;
    } else {
      const long n = (long)(ko + kv) * M;
      k_chi0_xt<<<(int)std::min<long>((n + 255) / 256, 16384L), 256, 0, st>>>(
          M, nocc, nvir, ko, kv, X_dev, h->XT);
      rc = launch_status("k_chi0_xt");
    }
  }
  if (rc) {
    mbpt_chi0_destroy(h);
    return rc;
  }
  *out = h;
  return MBPT_OK;
}

int mbpt_chi0_build(mbpt_chi0 *h, int ntau, const double *tau_host,
                    const double *weight_host, double *out_dev, int accumulate,
                    void *stream) {
  MBPT_REQUIRE(h != nullptr, "mbpt_chi0_build: NULL handle");
  MBPT_REQUIRE(ntau >= 1 && ntau <= MBPT_CHI0_MAX_TAU, "mbpt_chi0_build: ntau must be in [1, 64]");
  MBPT_REQUIRE(tau_host && weight_host && out_dev, "mbpt_chi0_build: NULL argument");
  for (int t = 0; t < ntau; ++t)
    MBPT_REQUIRE(tau_host[t] >= 0.0, "mbpt_chi0_build: time points must be non-negative");
  hipStream_t st = (hipStream_t)stream;
  return h->prec == MBPT_FP64
             ? build_impl<double>(h, ntau, tau_host, weight_host, out_dev, accumulate, st)
             : build_f32(h, ntau, tau_host, weight_host, out_dev, accumulate, st);
}

int mbpt_chi0_terms(mbpt_chi0 *h, int ntau, const double *tau_host, void *stream) {
  MBPT_REQUIRE(h != nullptr, "mbpt_chi0_terms: NULL handle");
  MBPT_REQUIRE(ntau >= 1 && ntau <= MBPT_CHI0_MAX_TAU, "mbpt_chi0_terms: ntau must be in [1, 64]");
  MBPT_REQUIRE(tau_host != nullptr, "mbpt_chi0_terms: NULL tau");
  for (int t = 0; t < ntau; ++t)
    MBPT_REQUIRE(tau_host[t] >= 0.0, "mbpt_chi0_terms: time points must be non-negative");
  hipStream_t st = (hipStream_t)stream;
  h->terms_nt = 0;
  const int rc = h->prec == MBPT_FP64 ? terms_f64(h, ntau, tau_host, st)
                                      : terms_f32(h, ntau, tau_host, st);
  if (!rc) h->terms_nt = ntau;
  return rc;
}

int mbpt_chi0_transform(mbpt_chi0 *h, int nfreq, const double *weight_host, double *out_dev,
                        int accumulate, void *stream) {
  MBPT_REQUIRE(h != nullptr, "mbpt_chi0_transform: NULL handle");
  MBPT_REQUIRE(h->terms_nt > 0, "mbpt_chi0_transform: no terms held (call mbpt_chi0_terms)");
  MBPT_REQUIRE(nfreq >= 0, "mbpt_chi0_transform: nfreq < 0");
  if (nfreq == 0) return MBPT_OK;
  MBPT_REQUIRE(weight_host && out_dev, "mbpt_chi0_transform: NULL argument");
  return transform_impl(h, nfreq, weight_host, out_dev, accumulate, (hipStream_t)stream);
}

int mbpt_chi0_transform_profile_read(const mbpt_chi0 *h, double *ms, long *calls) {
  MBPT_REQUIRE(h != nullptr, "mbpt_chi0_transform_profile_read: NULL handle");
  if (ms) *ms = h->tr_ms;
  if (calls) *calls = h->tr_launches;
  return MBPT_OK;
}

int mbpt_chi0_sigma(mbpt_chi0 *h, int ntau, const double *tau_host, const float *Zt_dev,
                    int nstates, const int *states_host, double *out_dev, void *stream) {
  MBPT_REQUIRE(h != nullptr, "mbpt_chi0_sigma: NULL handle");
  if (h->prec != MBPT_FP32) {
    set_error("mbpt_chi0_sigma: the self-energy sums are built for MBPT_FP32 handles");
    return MBPT_ERR_UNSUPPORTED;
  }
  MBPT_REQUIRE(ntau >= 1 && ntau <= MBPT_CHI0_MAX_TAU, "mbpt_chi0_sigma: ntau must be in [1, 64]");
  MBPT_REQUIRE(nstates >= 1 && nstates <= MBPT_GW_MAX_STATES,
               "mbpt_chi0_sigma: nstates must be in [1, MBPT_GW_MAX_STATES]");
  MBPT_REQUIRE(tau_host && Zt_dev && states_host && out_dev, "mbpt_chi0_sigma: NULL argument");
  for (int t = 0; t < ntau; ++t)
    MBPT_REQUIRE(tau_host[t] >= 0.0, "mbpt_chi0_sigma: time points must be non-negative");
  for (int j = 0; j < nstates; ++j)
    MBPT_REQUIRE(states_host[j] >= 0 && states_host[j] < h->no + h->nv,
                 "mbpt_chi0_sigma: a state index is outside the orbitals");
  return sigma_f32(h, ntau, tau_host, Zt_dev, nstates, states_host, out_dev,
                   (hipStream_t)stream);
}

int mbpt_chi0_sigma_profile_read(const mbpt_chi0 *h, double *ms, long *launches,
                                 double *flops) {
  MBPT_REQUIRE(h != nullptr, "mbpt_chi0_sigma_profile_read: NULL handle");
  if (ms) *ms = h->gw_ms;
  if (launches) *launches = h->gw_launches;
  if (flops) *flops = h->gw_flops;
  return MBPT_OK;
}

int mbpt_chi0_profile(mbpt_chi0 *h, int enable) {
  MBPT_REQUIRE(h != nullptr, "mbpt_chi0_profile: NULL handle");
  if (enable && !h->prof) {
    MBPT_HIP_TRY(hipEventCreate(&h->e0));
    MBPT_HIP_TRY(hipEventCreate(&h->e1));
  }
  if (!enable && h->prof) {
    (void)hipEventDestroy(h->e0);
    (void)hipEventDestroy(h->e1);
  }
  h->prof = enable ? 1 : 0;
  h->prof_ms = h->prof_flops = 0.0;
  h->prof_launches = 0;
  h->gw_ms = h->gw_flops = 0.0;
  h->gw_launches = 0;
  h->tr_ms = 0.0;
  h->tr_launches = 0;
  return MBPT_OK;
}

int mbpt_chi0_profile_read(const mbpt_chi0 *h, double *ms, long *launches,
                           double *flops) {
  MBPT_REQUIRE(h != nullptr, "mbpt_chi0_profile_read: NULL handle");
  if (ms) *ms = h->prof_ms;
  if (launches) *launches = h->prof_launches;
  if (flops) *flops = h->prof_flops;
  return MBPT_OK;
}

int mbpt_chi0_destroy(mbpt_chi0 *h) {
  if (!h) return MBPT_OK;
  if (h->prof) {
    (void)hipEventDestroy(h->e0);
    (void)hipEventDestroy(h->e1);
  }
  (void)hipFree(h->Xo); (void)hipFree(h->Xv); (void)hipFree(h->so); (void)hipFree(h->sv);
  (void)hipFree(h->eo); (void)hipFree(h->ev);
  (void)hipFree(h->Yt); (void)hipFree(h->part); (void)hipFree(h->XT);
  (void)hipFree(h->gw_xs); (void)hipFree(h->gw_part); (void)hipFree(h->gw_cols);
  (void)hipFree(h->terms);
  delete h;
  return MBPT_OK;
}

}  // extern "C"
