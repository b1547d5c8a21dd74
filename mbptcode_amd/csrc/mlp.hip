// The learned Davidson preconditioner: a 3 -> 256 -> 256 -> 1 SiLU network
// with BF16 weights and activations and FP32 accumulation
// (v_mfma_f32_16x16x32_bf16), its training step (forward, backward, Adam) and
// its use as the Davidson correction.
//
//   z1 = x W1 + b1, h1 = silu(z1);  z2 = h1 W2 + b2, h2 = silu(z2);  y = h2.w3 + b3
//
// Layer 1 (K = 3) and the output layer (N = 1) are FMA work done in
// registers; the 256 x 256 layer is the MFMA part.  W2 is kept in FP32
// (master copy for Adam) and packed to BF16 in MFMA fragment order twice:
// as the B operand of h1 W2 (forward) and of dz2 W2^T (backward).  A fragment
// of the 16x16x32 MFMA is 64 lanes x 8 bf16 = 1 KiB, stored contiguously, so
// staging it to LDS and reading it back are both conflict-free 16-byte moves.
//
// MFMA lane layout (measured on gfx950): lane l supplies A[l%16][8(l/16)+s]
// and B[8(l/16)+s][l%16] in slot s; the result leaves D[4(l/16)+r][l%16] in r.
#include "common.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace mbpt {
namespace {

constexpr int H = 256;          // hidden width
constexpr int KB = H / 32;      // 32-deep K blocks
constexpr int NT = H / 16;      // 16-wide output tiles
constexpr int NPARAM = H * H + 6 * H + 1;
// flattened parameter offsets: W1 [3][H], b1, W2 [H][H], b2, w3, b3
constexpr int OFF_W1 = 0, OFF_B1 = 3 * H, OFF_W2 = 4 * H, OFF_B2 = 4 * H + H * H,
              OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + H;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ float silu(float z) { return z / (1.f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = 1.f / (1.f + __expf(-z));
  return s * (1.f + z * (1.f - s));
}
// SiLU with the hardware exponential and reciprocal (inference: the result is
// rounded to BF16 or scaled by a BF16-trained weight next)
__device__ __forceinline__ float silu_fast(float z) {
  return z * __builtin_amdgcn_rcpf(1.f + __expf(-z));
}
__device__ __forceinline__ f32x4 mma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// fragment (kb, t) slot s of lane l:  fwd  B[k = 32kb+8q+s][n = 16t+lr] = W2[k][n]
//                                      bwd  B[k = 32kb+8q+s][m = 16t+lr] = W2[m][k]
__global__ void k_mlp_pack(const float *__restrict__ P, bf16x8 *__restrict__ fwd,
                           bf16x8 *__restrict__ bwd) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;  // (kb, t, lane)
  if (idx >= KB * NT * 64) return;
  const int l = idx & 63, t = (idx >> 6) % NT, kb = (idx >> 6) / NT;
  const int lr = l & 15, lq = l >> 4;
  const float *W2 = P + OFF_W2;
  bf16x8 f, b;
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int k = 32 * kb + 8 * lq + s, c = 16 * t + lr;
    f[s] = (__bf16)W2[k * H + c];
    b[s] = (__bf16)W2[c * H + k];
  }
  fwd[idx] = f;
  bwd[idx] = b;
}

// Inputs of one row: explicit features, or a Davidson residual block.
// Batched residual rows (mode 2): nb instances whose pair blocks are P rows
// apart in R, d and T, with off[b] the prefix sum of their real pair counts;
// network row = j ntot + g over real pairs g only, instance b = the block
// holding g.  theta and rms are (nb, k).
struct Batch {
  const int *off;
  long ntot;
  int nb, P, k;
  __device__ void locate(long row, long &pos, int &col) const {
    col = (int)(row / ntot);
    const int g = (int)(row % ntot);
    int lo = 0, hi = nb;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (off[mid] <= g) lo = mid; else hi = mid;
    }
    pos = (long)lo * P + (g - off[lo]);
    col += lo * k;   // index into theta / rms; the column itself is col % k
  }
};

struct Feat {
  const float *X;                  // mode 0: (N, 3)
  const double *R;                 // mode 1: residuals (n, k), row stride ldr
  const double *d, *theta, *rms;   // (n), (k), (k)
  int kc, ldr;                     // mode 1: network row = i k + j (pair i, column j)
  Batch bt;                        // mode 2 when bt.off is set
  __device__ void get(long row, float (&x)[3]) const {
    if (X) {
      x[0] = X[row * 3]; x[1] = X[row * 3 + 1]; x[2] = X[row * 3 + 2];
    } else if (bt.off) {
      long pos;
      int c;
      bt.locate(row, pos, c);
      x[0] = (float)(R[pos * ldr + c % bt.k] / rms[c]);
      x[1] = (float)(d[pos] - theta[c]);
      x[2] = (float)theta[c];
    } else {
      const int i = (int)(row / kc), j = (int)(row % kc);
      x[0] = (float)(R[(long)i * ldr + j] / rms[j]);
      x[1] = (float)(d[i] - theta[j]);
      x[2] = (float)theta[j];
    }
  }
};

struct Out {
  float *y;                        // mode 0: (N)
  double *T;                       // mode 1: corrections (n, k), row stride ldt
  const double *rms;
  int kc, ldt;                     // mode 1: network row = i k + j
  Batch bt;                        // mode 2 when bt.off is set
  __device__ void put(long row, float v) const {
    if (y) {
      y[row] = v;
    } else if (bt.off) {
      long pos;
      int c;
      bt.locate(row, pos, c);
      T[pos * ldt + c % bt.k] = (double)v * rms[c];
    } else {
      const int i = (int)(row / kc), j = (int)(row % kc);
      T[(long)i * ldt + j] = (double)v * rms[j];
    }
  }
};

// Forward over N rows, 16 rows per wave-step; W2 (fwd pack) staged in LDS
// once per workgroup, workgroups stride over the row tiles.  save: z2 (N, H)
// FP32 and h1^T (H, N) BF16 for the backward pass.
__global__ __launch_bounds__(256) void k_mlp_fwd(long N, const float *__restrict__ P,
                                                 const bf16x8 *__restrict__ wfwd,
                                                 Feat feat, Out out,
                                                 float *__restrict__ z2s,
                                                 __bf16 *__restrict__ h1t) {
  extern __shared__ unsigned char smem_raw[];
  bf16x8 *sw = (bf16x8 *)smem_raw;                        // KB*NT*64 fragments
  float *sp = (float *)(smem_raw + sizeof(bf16x8) * KB * NT * 64);  // W1,b1,b2,w3
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lq = lane >> 4;
  for (int e = tid; e < KB * NT * 64; e += 256) sw[e] = wfwd[e];
  for (int e = tid; e < 4 * H; e += 256) sp[e] = P[e];    // W1 (3H) and b1 (H)
  for (int e = tid; e < H; e += 256) {
    sp[4 * H + e] = P[OFF_B2 + e];
    sp[5 * H + e] = P[OFF_W3 + e];
  }
  __syncthreads();
  const float b3 = P[OFF_B3];
  const long ntiles = (N + 15) / 16;
  for (long tile = (long)blockIdx.x * 4 + wave; tile < ntiles; tile += (long)gridDim.x * 4) {
    const long row0 = tile * 16;
    const long rowA = row0 + lr;                 // this lane's row as an A operand
    float x[3] = {0.f, 0.f, 0.f};
    if (rowA < N) feat.get(rowA, x);
    bf16x8 a[KB];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        const int hh = 32 * kb + 8 * lq + s;
        const float z = x[0] * sp[hh] + x[1] * sp[H + hh] + x[2] * sp[2 * H + hh] +
                        sp[3 * H + hh];
        const __bf16 hb = (__bf16)silu(z);
        a[kb][s] = hb;
        if (h1t && rowA < N) h1t[(long)hh * N + rowA] = hb;
      }
    float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      f32x4 acc;
      zero(acc);
#pragma unroll
      for (int kb = 0; kb < KB; ++kb) acc = mma(a[kb], sw[(kb * NT + t) * 64 + lane], acc);
      const int n = 16 * t + lr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float z = acc[r] + sp[4 * H + n];
        const long row = row0 + 4 * lq + r;
        if (z2s && row < N) z2s[row * H + n] = z;
        part[r] += silu(z) * sp[5 * H + n];
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float v = part[r];
      v += __shfl_xor(v, 1);
      v += __shfl_xor(v, 2);
      v += __shfl_xor(v, 4);
      v += __shfl_xor(v, 8);
      const long row = row0 + 4 * lq + r;
      if (lr == 0 && row < N) out.put(row, v + b3);
    }
  }
}

// Inference forward (no saved activations): 8 waves share one LDS copy of W2;
// a wave takes 64 rows (four 16-row tiles) per step, so each W2 fragment read
// from LDS feeds four MFMAs, and each lane's layer-1 weights (8 consecutive
// hidden units per K block, two 16-byte reads per array) serve all four tiles.
constexpr int kInferRT = 4;   // 16-row tiles per wave step
__global__ __launch_bounds__(512) void k_mlp_infer(long N, const float *__restrict__ P,
                                                   const bf16x8 *__restrict__ wfwd,
                                                   Feat feat, Out out) {
  extern __shared__ unsigned char smem_raw[];
  bf16x8 *sw = (bf16x8 *)smem_raw;
  float *sp = (float *)(smem_raw + sizeof(bf16x8) * KB * NT * 64);  // W1, b1, b2, w3
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lq = lane >> 4;
  for (int e = tid; e < KB * NT * 64; e += 512) sw[e] = wfwd[e];
  for (int e = tid; e < 4 * H; e += 512) sp[e] = P[e];
  for (int e = tid; e < H; e += 512) {
    sp[4 * H + e] = P[OFF_B2 + e];
    sp[5 * H + e] = P[OFF_W3 + e];
  }
  __syncthreads();
  const float b3 = P[OFF_B3];
  constexpr int RT = kInferRT;
  const long nsteps = (N + 16 * RT - 1) / (16 * RT);
  for (long step = (long)blockIdx.x * 8 + wave; step < nsteps; step += (long)gridDim.x * 8) {
    const long row0 = step * 16 * RT;
    float x[RT][3];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
      const long row = row0 + 16 * rt + lr;
      x[rt][0] = x[rt][1] = x[rt][2] = 0.f;
      if (row < N) feat.get(row, x[rt]);
    }
    bf16x8 a[RT][KB];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      const int h0 = 32 * kb + 8 * lq;
      float wa[8], wb[8], wc[8], bb[8];
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        wa[s] = sp[h0 + s];
        wb[s] = sp[H + h0 + s];
        wc[s] = sp[2 * H + h0 + s];
        bb[s] = sp[3 * H + h0 + s];
      }
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int s = 0; s < 8; ++s)
          a[rt][kb][s] = (__bf16)silu_fast(x[rt][0] * wa[s] + x[rt][1] * wb[s] +
                                           x[rt][2] * wc[s] + bb[s]);
      __builtin_amdgcn_sched_barrier(0);   // keep each block's weight loads local
    }
    float part[RT][4];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int r = 0; r < 4; ++r) part[rt][r] = 0.f;
#pragma unroll 2
    for (int t = 0; t < NT; ++t) {
      f32x4 acc[RT];
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) zero(acc[rt]);
#pragma unroll
      for (int kb = 0; kb < KB; ++kb) {
        const bf16x8 bw = sw[(kb * NT + t) * 64 + lane];
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) acc[rt] = mma(a[rt][kb], bw, acc[rt]);
      }
      const int n = 16 * t + lr;
      const float b2 = sp[4 * H + n], w3 = sp[5 * H + n];
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int r = 0; r < 4; ++r) part[rt][r] += silu_fast(acc[rt][r] + b2) * w3;
    }
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float v = part[rt][r];
        v += __shfl_xor(v, 1);
        v += __shfl_xor(v, 2);
        v += __shfl_xor(v, 4);
        v += __shfl_xor(v, 8);
        const long row = row0 + 16 * rt + 4 * lq + r;
        if (lr == 0 && row < N) out.put(row, v + b3);
      }
  }
}

// Backward over row tiles.  g = dL/dy = 2 (y - t) / N.  Writes dz2^T (H, N)
// BF16 for the dW2 product and one partial per workgroup of
// [dW1 (3H) | db1 (H) | db2 (H) | dw3 (H) | db3] into part[blockIdx].
constexpr int NPART = 6 * H + 1;

__global__ __launch_bounds__(256) void k_mlp_bwd_rows(
    long N, const float *__restrict__ P, const bf16x8 *__restrict__ wbwd,
    const float *__restrict__ X, const float *__restrict__ tgt,
    const float *__restrict__ y, const float *__restrict__ z2s,
    __bf16 *__restrict__ dz2t, float *__restrict__ part) {
  extern __shared__ unsigned char smem_raw[];
  bf16x8 *sw = (bf16x8 *)smem_raw;
  float *sp = (float *)(smem_raw + sizeof(bf16x8) * KB * NT * 64);  // W1, b1, w3
  float *acc_s = sp + 5 * H;                                          // NPART
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lq = lane >> 4;
  for (int e = tid; e < KB * NT * 64; e += 256) sw[e] = wbwd[e];
  for (int e = tid; e < 4 * H; e += 256) sp[e] = P[e];
  for (int e = tid; e < H; e += 256) sp[4 * H + e] = P[OFF_W3 + e];
  for (int e = tid; e < NPART; e += 256) acc_s[e] = 0.f;
  __syncthreads();
  const float invN = 2.f / (float)N;
  // per-lane partials: dW1/db1 for h = 16t + lr (all t), over this lane's rows
  float pw1[3][NT], pb1[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) { pw1[0][t] = pw1[1][t] = pw1[2][t] = 0.f; pb1[t] = 0.f; }
  float pdb3 = 0.f;
  const long ntiles = (N + 15) / 16;
  for (long tile = (long)blockIdx.x * 4 + wave; tile < ntiles; tile += (long)gridDim.x * 4) {
    const long row0 = tile * 16;
    const long rowA = row0 + lr;
    const float g = rowA < N ? (y[rowA] - tgt[rowA]) * invN : 0.f;
    if (lq == 0) pdb3 += g;
    // dz2 for (row lr, n = 32kb+8q+s): A operand of dz2 W2^T
    bf16x8 a[KB];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        const int n = 32 * kb + 8 * lq + s;
        float dz = 0.f, hg = 0.f;
        if (rowA < N) {
          const float z = z2s[rowA * H + n];
          dz = g * sp[4 * H + n] * silu_grad(z);
          hg = g * silu(z);
        }
        // db2, dw3: sum over the 16 rows (lanes lr) first, then one atomic
        float s1 = dz, s2 = hg;
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) {
          s1 += __shfl_xor(s1, off);
          s2 += __shfl_xor(s2, off);
        }
        if (lr == 0) {
          atomicAdd(&acc_s[4 * H + n], s1);
          atomicAdd(&acc_s[5 * H + n], s2);
        }
        const __bf16 db = (__bf16)dz;
        a[kb][s] = db;
        if (rowA < N) dz2t[(long)n * N + rowA] = db;
      }
    // dh1 = dz2 W2^T, result rows 4q+r, h = 16t+lr
    float xr[4][3];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const long row = row0 + 4 * lq + r;
      xr[r][0] = xr[r][1] = xr[r][2] = 0.f;
      if (row < N) { xr[r][0] = X[row * 3]; xr[r][1] = X[row * 3 + 1]; xr[r][2] = X[row * 3 + 2]; }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      f32x4 acc;
      zero(acc);
#pragma unroll
      for (int kb = 0; kb < KB; ++kb) acc = mma(a[kb], sw[(kb * NT + t) * 64 + lane], acc);
      const int hh = 16 * t + lr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long row = row0 + 4 * lq + r;
        if (row >= N) continue;
        const float z1 = xr[r][0] * sp[hh] + xr[r][1] * sp[H + hh] + xr[r][2] * sp[2 * H + hh] +
                         sp[3 * H + hh];
        const float dz1 = acc[r] * silu_grad(z1);
        pw1[0][t] += xr[r][0] * dz1;
        pw1[1][t] += xr[r][1] * dz1;
        pw1[2][t] += xr[r][2] * dz1;
        pb1[t] += dz1;
      }
    }
  }
  // fold the per-lane dW1/db1 partials over the four row groups, then the waves
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    float v[4] = {pw1[0][t], pw1[1][t], pw1[2][t], pb1[t]};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      v[c] += __shfl_xor(v[c], 16);
      v[c] += __shfl_xor(v[c], 32);
    }
    if (lq == 0) {
      const int hh = 16 * t + lr;
      atomicAdd(&acc_s[hh], v[0]);
      atomicAdd(&acc_s[H + hh], v[1]);
      atomicAdd(&acc_s[2 * H + hh], v[2]);
      atomicAdd(&acc_s[3 * H + hh], v[3]);
    }
  }
  if (lq == 0) {
    float v = pdb3;
    v += __shfl_xor(v, 1); v += __shfl_xor(v, 2); v += __shfl_xor(v, 4); v += __shfl_xor(v, 8);
    if (lr == 0) atomicAdd(&acc_s[6 * H], v);
  }
  __syncthreads();
  for (int e = tid; e < NPART; e += 256) part[(long)blockIdx.x * NPART + e] = acc_s[e];
}

// dW2[k][n] = sum_rows h1[row][k] dz2[row][n], rows split in chunks: one
// 64x64 tile of one chunk per workgroup (4 waves x 16 k-rows, 4 n-tiles each).
__global__ __launch_bounds__(256) void k_mlp_dw2(long N, long chunk,
                                                 const __bf16 *__restrict__ h1t,
                                                 const __bf16 *__restrict__ dz2t,
                                                 float *__restrict__ out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lq = lane >> 4;
  const int tk = blockIdx.x % 4, tn = (blockIdx.x / 4) % 4;
  const long c = blockIdx.x / 16;
  const long r0 = c * chunk, r1 = std::min<long>(N, r0 + chunk);
  const int k = 64 * tk + 16 * wave + lr;          // A row (output row k)
  f32x4 acc[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) zero(acc[j]);
  for (long rb = r0; rb < r1; rb += 32) {
    bf16x8 a, b[4];
    const long base = rb + 8 * lq;
#pragma unroll
    for (int s = 0; s < 8; ++s) a[s] = base + s < r1 ? h1t[(long)k * N + base + s] : (__bf16)0.f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = 64 * tn + 16 * j + lr;
#pragma unroll
      for (int s = 0; s < 8; ++s)
        b[j][s] = base + s < r1 ? dz2t[(long)n * N + base + s] : (__bf16)0.f;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] = mma(a, b[j], acc[j]);
  }
  float *o = out + c * (long)H * H;
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int kk = 64 * tk + 16 * wave + 4 * lq + r, n = 64 * tn + 16 * j + lr;
      o[kk * H + n] = acc[j][r];
    }
}

// grad = sum of the partials, in the flattened parameter order
__global__ void k_mlp_grad_sum(int nparts, const float *__restrict__ part, int nchunks,
                               const float *__restrict__ dw2p, float *__restrict__ grad) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= NPARAM) return;
  float s = 0.f;
  if (e >= OFF_W2 && e < OFF_B2) {
    const int q = e - OFF_W2;
    for (int c = 0; c < nchunks; ++c) s += dw2p[(long)c * H * H + q];
  } else {
    int src;  // index into a partial [dW1 3H | db1 H | db2 H | dw3 H | db3]
    if (e < OFF_B1) src = e;
    else if (e < OFF_W2) src = 3 * H + (e - OFF_B1);
    else if (e < OFF_W3) src = 4 * H + (e - OFF_B2);
    else if (e < OFF_B3) src = 5 * H + (e - OFF_W3);
    else src = 6 * H;
    for (int p = 0; p < nparts; ++p) s += part[(long)p * NPART + src];
  }
  grad[e] = s;
}

__global__ void k_mlp_adam(float *__restrict__ P, const float *__restrict__ grad,
                           float *__restrict__ m, float *__restrict__ v, float lr,
                           float b1, float b2, float eps, float c1, float c2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= NPARAM) return;
  const float g = grad[e];
  const float mm = b1 * m[e] + (1.f - b1) * g;
  const float vv = b2 * v[e] + (1.f - b2) * g * g;
  m[e] = mm;
  v[e] = vv;
  P[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

__global__ void k_mse(long N, const float *__restrict__ y, const float *__restrict__ t,
                      float *__restrict__ partial) {
  float s = 0.f;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < N;
       e += (long)gridDim.x * blockDim.x) {
    const float d = y[e] - t[e];
    s += d * d;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  __shared__ float ws[4];
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

// rms over the n entries of each of the k columns of R (row stride ldr)
__global__ void k_col_rms(int n, int k, const double *__restrict__ R, int ldr,
                          double *__restrict__ rms) {
  const int j = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double v = R[(long)i * ldr + j];
    s += v * v;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  __shared__ double ws[4];
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) rms[j] = sqrt((ws[0] + ws[1] + ws[2] + ws[3]) / n) + 1e-300;
}

// rms per (instance, column) over the instance's real pairs, mode-2 layout
__global__ void k_col_rms_batched(Batch bt, const double *__restrict__ R, int ldr,
                                  double *__restrict__ rms) {
  const int b = blockIdx.x / bt.k, j = blockIdx.x % bt.k;
  const int n = bt.off[b + 1] - bt.off[b];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double v = R[((long)b * bt.P + i) * ldr + j];
    s += v * v;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  __shared__ double ws[4];
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0)
    rms[blockIdx.x] = sqrt((ws[0] + ws[1] + ws[2] + ws[3]) / (n > 0 ? n : 1)) + 1e-300;
}

constexpr size_t kFwdSmem = sizeof(bf16x8) * KB * NT * 64 + sizeof(float) * 6 * H;
constexpr size_t kBwdSmem = sizeof(bf16x8) * KB * NT * 64 + sizeof(float) * (5 * H + NPART);
constexpr int kRowBlocks = 256;   // persistent workgroups for the row kernels

}  // namespace
}  // namespace mbpt

using namespace mbpt;

struct mbpt_mlp {
  float *P, *grad, *m, *v;       // FP32 parameters, gradient, Adam moments
  bf16x8 *wfwd, *wbwd;           // packed BF16 W2
  float *part, *dw2p, *loss;     // partials
  long cap_rows;                 // rows the training buffers hold
  float *y, *z2;
  __bf16 *h1t, *dz2t;
  double *rms;
  int cap_k;
  long step;
};

namespace {

int pack(mbpt_mlp *h, hipStream_t st) {
  k_mlp_pack<<<(KB * NT * 64 + 255) / 256, 256, 0, st>>>(h->P, h->wfwd, h->wbwd);
  return launch_status("k_mlp_pack");
}

int ensure_rows(mbpt_mlp *h, long N) {
  if (N <= h->cap_rows) return MBPT_OK;
  (void)hipFree(h->y); (void)hipFree(h->z2); (void)hipFree(h->h1t); (void)hipFree(h->dz2t);
  h->y = h->z2 = nullptr; h->h1t = h->dz2t = nullptr; h->cap_rows = 0;
  MBPT_HIP_TRY(hipMalloc(&h->y, sizeof(float) * N));
  MBPT_HIP_TRY(hipMalloc(&h->z2, sizeof(float) * N * H));
  MBPT_HIP_TRY(hipMalloc(&h->h1t, sizeof(__bf16) * N * H));
  MBPT_HIP_TRY(hipMalloc(&h->dz2t, sizeof(__bf16) * N * H));
  h->cap_rows = N;
  return MBPT_OK;
}

int grid_rows(long N) {
  return (int)std::max<long>(1, std::min<long>(kRowBlocks, (N + 63) / 64));
}

// persistent inference workgroups: at most one per CU (W2 fills most of the LDS)
int grid_infer(long N) {
  return (int)std::max<long>(1, std::min<long>(kRowBlocks, (N + 511) / 512));
}

constexpr long kDw2Chunks = 64;

int gradient(mbpt_mlp *h, long N, const float *X, const float *T, hipStream_t st) {
  int rc = ensure_rows(h, N);
  if (rc) return rc;
  Feat f{X, nullptr, nullptr, nullptr, nullptr, 0, 0};
  Out o{h->y, nullptr, nullptr, 0, 0};
  const int gr = grid_rows(N);
  k_mlp_fwd<<<gr, 256, kFwdSmem, st>>>(N, h->P, h->wfwd, f, o, h->z2, h->h1t);
  if ((rc = launch_status("k_mlp_fwd"))) return rc;
  k_mlp_bwd_rows<<<gr, 256, kBwdSmem, st>>>(N, h->P, h->wbwd, X, T, h->y, h->z2, h->dz2t,
                                            h->part);
  if ((rc = launch_status("k_mlp_bwd_rows"))) return rc;
  const long chunk = std::max<long>(32, ((N + kDw2Chunks - 1) / kDw2Chunks + 31) / 32 * 32);
  const long nchunks = (N + chunk - 1) / chunk;
  k_mlp_dw2<<<(int)(16 * nchunks), 256, 0, st>>>(N, chunk, h->h1t, h->dz2t, h->dw2p);
  if ((rc = launch_status("k_mlp_dw2"))) return rc;
  k_mlp_grad_sum<<<(NPARAM + 255) / 256, 256, 0, st>>>(gr, h->part, (int)nchunks, h->dw2p,
                                                       h->grad);
  if ((rc = launch_status("k_mlp_grad_sum"))) return rc;
  k_mse<<<64, 256, 0, st>>>(N, h->y, T, h->loss);
  return launch_status("k_mse");
}

}  // namespace

extern "C" {

int mbpt_mlp_nparams(void) { return NPARAM; }

int mbpt_mlp_create(mbpt_mlp **out, const float *params_host, void *stream) {
  MBPT_REQUIRE(out && params_host, "mbpt_mlp_create: NULL argument");
  *out = nullptr;
  hipStream_t st = (hipStream_t)stream;
  mbpt_mlp *h = new mbpt_mlp();
  hipError_t e = hipSuccess;
  const size_t pb = sizeof(float) * NPARAM;
  if (!e) e = hipMalloc(&h->P, pb);
  if (!e) e = hipMalloc(&h->grad, pb);
  if (!e) e = hipMalloc(&h->m, pb);
  if (!e) e = hipMalloc(&h->v, pb);
  if (!e) e = hipMalloc(&h->wfwd, sizeof(bf16x8) * KB * NT * 64);
  if (!e) e = hipMalloc(&h->wbwd, sizeof(bf16x8) * KB * NT * 64);
  if (!e) e = hipMalloc(&h->part, sizeof(float) * kRowBlocks * NPART);
  if (!e) e = hipMalloc(&h->dw2p, sizeof(float) * kDw2Chunks * H * H);
  if (!e) e = hipMalloc(&h->loss, sizeof(float) * 64);
  if (!e) e = hipMemcpyAsync(h->P, params_host, pb, hipMemcpyHostToDevice, st);
  if (!e) e = hipMemsetAsync(h->m, 0, pb, st);
  if (!e) e = hipMemsetAsync(h->v, 0, pb, st);
  if (!e) e = hipStreamSynchronize(st);
  if (!e) e = hipFuncSetAttribute((const void *)k_mlp_fwd,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)kFwdSmem);
  if (!e) e = hipFuncSetAttribute((const void *)k_mlp_infer,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)kFwdSmem);
  if (!e) e = hipFuncSetAttribute((const void *)k_mlp_bwd_rows,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)kBwdSmem);
  if (e != hipSuccess) {
    set_error(std::string("mbpt_mlp_create: ") + hipGetErrorString(e));
    mbpt_mlp_destroy(h);
    return MBPT_ERR_HIP;
  }
  int rc = pack(h, st);
  if (rc) { mbpt_mlp_destroy(h); return rc; }
  *out = h;
  return MBPT_OK;
}

int mbpt_mlp_get_params(mbpt_mlp *h, float *params_host, void *stream) {
  MBPT_REQUIRE(h && params_host, "mbpt_mlp_get_params: NULL argument");
  hipStream_t st = (hipStream_t)stream;
  MBPT_HIP_TRY(hipMemcpyAsync(params_host, h->P, sizeof(float) * NPARAM,
                              hipMemcpyDeviceToHost, st));
  MBPT_HIP_TRY(hipStreamSynchronize(st));
  return MBPT_OK;
}

int mbpt_mlp_forward(mbpt_mlp *h, long N, const float *X_dev, float *y_dev, void *stream) {
  MBPT_REQUIRE(h && N >= 0, "mbpt_mlp_forward: bad arguments");
  if (N == 0) return MBPT_OK;
  MBPT_REQUIRE(X_dev && y_dev, "mbpt_mlp_forward: NULL argument");
  hipStream_t st = (hipStream_t)stream;
  Feat f{X_dev, nullptr, nullptr, nullptr, nullptr, 0, 0};
  Out o{y_dev, nullptr, nullptr, 0, 0};
  k_mlp_infer<<<grid_infer(N), 512, kFwdSmem, st>>>(N, h->P, h->wfwd, f, o);
  return launch_status("k_mlp_infer");
}

int mbpt_mlp_gradient(mbpt_mlp *h, long N, const float *X_dev, const float *T_dev,
                      float *grad_dev, float *loss_host, void *stream) {
  MBPT_REQUIRE(h && N > 0 && X_dev && T_dev, "mbpt_mlp_gradient: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  int rc = gradient(h, N, X_dev, T_dev, st);
  if (rc) return rc;
  if (grad_dev)
    MBPT_HIP_TRY(hipMemcpyAsync(grad_dev, h->grad, sizeof(float) * NPARAM,
                                hipMemcpyDeviceToDevice, st));
  if (loss_host) {
    float part[64];
    MBPT_HIP_TRY(hipMemcpyAsync(part, h->loss, sizeof(part), hipMemcpyDeviceToHost, st));
    MBPT_HIP_TRY(hipStreamSynchronize(st));
    double s = 0.0;
    for (int q = 0; q < 64; ++q) s += part[q];
    *loss_host = (float)(s / (double)N);
  }
  return MBPT_OK;
}

int mbpt_mlp_train_step(mbpt_mlp *h, long N, const float *X_dev, const float *T_dev,
                        float lr, float *loss_host, void *stream) {
  MBPT_REQUIRE(h && N > 0 && X_dev && T_dev && lr > 0.f, "mbpt_mlp_train_step: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  int rc = mbpt_mlp_gradient(h, N, X_dev, T_dev, nullptr, loss_host, stream);
  if (rc) return rc;
  h->step += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)h->step), c2 = 1.f - powf(b2, (float)h->step);
  k_mlp_adam<<<(NPARAM + 255) / 256, 256, 0, st>>>(h->P, h->grad, h->m, h->v, lr, b1, b2,
                                                   1e-8f, c1, c2);
  if ((rc = launch_status("k_mlp_adam"))) return rc;
  return pack(h, st);
}

int mbpt_mlp_correction(mbpt_mlp *h, int n, int k, const double *R_dev, int ldr,
                        const double *d_dev, const double *theta_dev, double *T_dev,
                        int ldt, void *stream) {
  MBPT_REQUIRE(h && n > 0 && k >= 0, "mbpt_mlp_correction: bad arguments");
  if (k == 0) return MBPT_OK;
  MBPT_REQUIRE(R_dev && d_dev && theta_dev && T_dev && ldr >= k && ldt >= k,
               "mbpt_mlp_correction: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  if (k > h->cap_k) {
    (void)hipFree(h->rms);
    h->rms = nullptr;
    MBPT_HIP_TRY(hipMalloc(&h->rms, sizeof(double) * k));
    h->cap_k = k;
  }
  k_col_rms<<<k, 256, 0, st>>>(n, k, R_dev, ldr, h->rms);
  int rc = launch_status("k_col_rms");
  if (rc) return rc;
  const long N = (long)n * k;
  Feat f{nullptr, R_dev, d_dev, theta_dev, h->rms, k, ldr};
  Out o{nullptr, T_dev, h->rms, k, ldt};
  k_mlp_infer<<<grid_infer(N), 512, kFwdSmem, st>>>(N, h->P, h->wfwd, f, o);
  return launch_status("k_mlp_infer");
}

int mbpt_mlp_correction_batched(mbpt_mlp *h, int nbatch, int P, int k, const int *off_dev,
                                long ntot, const double *R_dev, int ldr, const double *d_dev,
                                const double *theta_dev, double *T_dev, int ldt,
                                void *stream) {
  MBPT_REQUIRE(h && nbatch > 0 && P > 0 && k >= 0 && ntot >= 0,
               "mbpt_mlp_correction_batched: bad arguments");
  if (k == 0 || ntot == 0) return MBPT_OK;
  MBPT_REQUIRE(off_dev && R_dev && d_dev && theta_dev && T_dev && ldr >= k && ldt >= k,
               "mbpt_mlp_correction_batched: bad arguments");
  MBPT_REQUIRE(ntot * k < (1L << 40), "mbpt_mlp_correction_batched: too many rows");
  hipStream_t st = (hipStream_t)stream;
  const int nk = nbatch * k;
  if (nk > h->cap_k) {
    (void)hipFree(h->rms);
    h->rms = nullptr;
    MBPT_HIP_TRY(hipMalloc(&h->rms, sizeof(double) * nk));
    h->cap_k = nk;
  }
  const Batch bt{off_dev, ntot, nbatch, P, k};
  k_col_rms_batched<<<nk, 256, 0, st>>>(bt, R_dev, ldr, h->rms);
  int rc = launch_status("k_col_rms_batched");
  if (rc) return rc;
  const long N = ntot * k;
  Feat f{nullptr, R_dev, d_dev, theta_dev, h->rms, 0, ldr, bt};
  Out o{nullptr, T_dev, h->rms, 0, ldt, bt};
  k_mlp_infer<<<grid_infer(N), 512, kFwdSmem, st>>>(N, h->P, h->wfwd, f, o);
  return launch_status("k_mlp_infer");
}

int mbpt_mlp_destroy(mbpt_mlp *h) {
  if (!h) return MBPT_OK;
  (void)hipFree(h->P); (void)hipFree(h->grad); (void)hipFree(h->m); (void)hipFree(h->v);
  (void)hipFree(h->wfwd); (void)hipFree(h->wbwd); (void)hipFree(h->part);
  (void)hipFree(h->dw2p); (void)hipFree(h->loss); (void)hipFree(h->y); (void)hipFree(h->z2);
  (void)hipFree(h->h1t); (void)hipFree(h->dz2t); (void)hipFree(h->rms);
  delete h;
  return MBPT_OK;
}

}  // extern "C"
