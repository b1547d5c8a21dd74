// Shared pieces of the gfx950 kernels: status handling, MFMA traits, the
// XCD-aware block remap and the 64x64 LDS-staged MFMA tile.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/mbpt_hip.h"

namespace mbpt {

// ---- status ----------------------------------------------------------------
void set_error(const std::string &msg);

#define MBPT_HIP_TRY(expr)                                                      \
  do {                                                                         \
    hipError_t _e = (expr);                                                    \
    if (_e != hipSuccess) {                                                    \
      ::mbpt::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));   \
      return MBPT_ERR_HIP;                                                     \
    }                                                                          \
  } while (0)

#define MBPT_REQUIRE(cond, msg)                                                \
  do {                                                                         \
    if (!(cond)) {                                                             \
      ::mbpt::set_error(msg);                                                  \
      return MBPT_ERR_VALUE;                                                   \
    }                                                                          \
  } while (0)

inline int launch_status(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string("launch of ") + what + ": " + hipGetErrorString(e));
    return MBPT_ERR_HIP;
  }
  return MBPT_OK;
}

// ---- MFMA traits -------------------------------------------------------------
// v_mfma_{f32,f64}_16x16x4: lane l supplies A[l%16][l/16] and B[l/16][l%16].
// The 16x16 result's row for element r of lane l differs by type (measured on
// gfx950): f32 keeps D[4*(l/16)+r][l%16], f64 interleaves D[l/16+4*r][l%16].
// orow(l/16, r) gives that row.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename T> struct MF;
template <> struct MF<float> {
  typedef f32x4 acc;
  static __device__ __forceinline__ acc mma(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int orow(int lq, int r) { return 4 * lq + r; }
};
template <> struct MF<double> {
  typedef f64x4 acc;
  static __device__ __forceinline__ acc mma(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int orow(int lq, int r) { return lq + 4 * r; }
};

template <typename A> __device__ __forceinline__ void zero(A &a) {
  a[0] = 0; a[1] = 0; a[2] = 0; a[3] = 0;
}

// ---- XCD-aware remap -----------------------------------------------------------
// Consecutive workgroup ids are dealt round-robin over the 8 XCDs, each with
// its own L2.  With the grid padded to a multiple of 8, this gives every XCD a
// contiguous range of logical ids, so neighbouring tiles (which share operand
// rows) land in one L2.
constexpr int kXcds = 8;
inline int pad_grid(long n) { return (int)((n + kXcds - 1) / kXcds * kXcds); }
__device__ __forceinline__ int xcd_logical(int bid, int nblk_padded) {
  return (bid % kXcds) * (nblk_padded / kXcds) + bid / kXcds;
}

// ---- 64x64 MFMA tile -------------------------------------------------------------
// acc += A[m0:m0+64, 0:K] B[0:K, n0:n0+64] with 256 threads (4 waves in 2x2,
// 32x32 each = 2x2 MFMA fragments), K staged through LDS 16 at a time.
// Loaders return 0 outside the matrix.  AKF/BKF: the loader's k index is the
// unit-stride one, which sets the thread->element map for coalescing.
constexpr int kTile = 64;
constexpr int kBK = 16;
constexpr int kPad = 4;

template <typename T> struct TileSmem {
  T a[kBK][kTile + kPad];
  T b[kBK][kTile + kPad];
};

template <typename T, bool AKF, bool BKF, class LA, class LB>
__device__ __forceinline__ void gemm64(const LA &la, const LB &lb, int m0, int n0,
                                       int K, typename MF<T>::acc (&acc)[2][2],
                                       TileSmem<T> &sm) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int lr = lane & 15, lq = lane >> 4;
  for (int k0 = 0; k0 < K; k0 += kBK) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = tid + 256 * j;
      int m, k;
      if (AKF) { m = e >> 4; k = e & 15; } else { m = e & 63; k = e >> 6; }
      sm.a[k][m] = la(m0 + m, k0 + k);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = tid + 256 * j;
      int n, k;
      if (BKF) { n = e >> 4; k = e & 15; } else { n = e & 63; k = e >> 6; }
      sm.b[k][n] = lb(k0 + k, n0 + n);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kBK / 4; ++kk) {
      T a0 = sm.a[4 * kk + lq][32 * wm + lr];
      T a1 = sm.a[4 * kk + lq][32 * wm + 16 + lr];
      T b0 = sm.b[4 * kk + lq][32 * wn + lr];
      T b1 = sm.b[4 * kk + lq][32 * wn + 16 + lr];
      acc[0][0] = MF<T>::mma(a0, b0, acc[0][0]);
      acc[0][1] = MF<T>::mma(a0, b1, acc[0][1]);
      acc[1][0] = MF<T>::mma(a1, b0, acc[1][0]);
      acc[1][1] = MF<T>::mma(a1, b1, acc[1][1]);
    }
    __syncthreads();
  }
}

// Output coordinates of fragment (fm, fn), element r, for this lane.
template <typename T> __device__ __forceinline__ int tile_row(int m0, int fm, int r) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  return m0 + 32 * (wave >> 1) + 16 * fm + MF<T>::orow(lane >> 4, r);
}
__device__ __forceinline__ int tile_col(int n0, int fn) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  return n0 + 32 * (wave & 1) + 16 * fn + (lane & 15);
}

}  // namespace mbpt
