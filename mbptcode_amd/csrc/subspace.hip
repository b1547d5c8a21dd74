// FP64 reductions over the pair index for the Davidson subspace on gfx950:
// C (a x b) = A^T B with A (n x a) and B (n x b), n the pair count (10^3 to
// 10^5) and a, b the trial-space widths (tens to a few hundred).  A library
// GEMM tiles the tiny output and runs the long contraction in one or two
// workgroups; here the contraction is cut into row chunks, each chunk's
// partial product is an MFMA tile (v_mfma_f64_16x16x4f64), and a second pass
// sums the partials in chunk order (deterministic).
#include "common.h"

#include <algorithm>
#include <mutex>
#include <string>

namespace mbpt {

struct LdAt {  // A(m = i, k = r) = A[r][i]
  const double *p;
  int ld, rows, K;
  __device__ double operator()(int m, int k) const {
    return (m < rows && k < K) ? p[(long)k * ld + m] : 0.0;
  }
};
struct LdBr {  // B(k = r, n = j) = B[r][j]
  const double *p;
  int ld, K, cols;
  __device__ double operator()(int k, int n) const {
    return (k < K && n < cols) ? p[(long)k * ld + n] : 0.0;
  }
};

__global__ __launch_bounds__(256) void k_tn_partial(int n, int a, int b, int chunk,
                                                    const double *__restrict__ A, int lda,
                                                    const double *__restrict__ B, int ldb,
                                                    double *__restrict__ P, int nta,
                                                    int ntb, int nc, int nblk) {
  const int lid = xcd_logical(blockIdx.x, nblk);
  if (lid >= nta * ntb * nc) return;
  const int tb = lid % ntb, ta = (lid / ntb) % nta, c = lid / (ntb * nta);
  const int r0 = c * chunk, rc = min(chunk, n - r0);
  __shared__ TileSmem<double> sm;
  MF<double>::acc acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) zero(acc[i][j]);
  gemm64<double, false, false>(LdAt{A + (long)r0 * lda, lda, a, rc},
                               LdBr{B + (long)r0 * ldb, ldb, rc, b}, ta * kTile,
                               tb * kTile, rc, acc, sm);
  double *out = P + (long)c * a * b;
  for (int fm = 0; fm < 2; ++fm)
    for (int fn = 0; fn < 2; ++fn) {
      const int j = tile_col(tb * kTile, fn);
      for (int r = 0; r < 4; ++r) {
        const int i = tile_row<double>(ta * kTile, fm, r);
        if (i < a && j < b) out[(long)i * b + j] = acc[fm][fn][r];
      }
    }
}

// 16 output elements per block, each summed by 16 threads over interleaved
// chunks (coalesced 128-byte rows of the partial buffer) and then across the
// 16 threads in LDS: enough loads in flight for hundreds of chunks, and a
// fixed summation order.
__global__ __launch_bounds__(256) void k_tn_sum(int a, int b, int nc,
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

}  // namespace mbpt

using namespace mbpt;

namespace {
// Rows per chunk: about a thousand workgroups in all, at least 256 rows each.
int tn_chunk(int n, int a, int b) {
  const long tiles = (long)((a + kTile - 1) / kTile) * ((b + kTile - 1) / kTile);
  long nc = std::max<long>(1, 1024 / tiles);
  nc = std::min<long>(nc, (n + 255) / 256);
  const int c = (int)((n + nc - 1) / nc);
  return (c + 15) / 16 * 16;
}

// The two launches of C = A^T B with the chunk partials in `work`.
int tn_launch(int n, int a, int b, const double *A, int lda, const double *B, int ldb,
              double *C, int ldc, double *work, long work_elems, hipStream_t st) {
  const int chunk = tn_chunk(n, a, b);
  const int nc = (n + chunk - 1) / chunk;
  MBPT_REQUIRE(work_elems >= (long)nc * a * b, "mbpt_gemm_tn_f64: work buffer too small");
  const int nta = (a + kTile - 1) / kTile, ntb = (b + kTile - 1) / kTile;
  const int nblk = pad_grid((long)nta * ntb * nc);
  k_tn_partial<<<nblk, 256, 0, st>>>(n, a, b, chunk, A, lda, B, ldb, work, nta, ntb, nc,
                                     nblk);
  int rc = launch_status("k_tn_partial");
  if (rc) return rc;
  const long count = (long)a * b;
  k_tn_sum<<<(int)std::min<long>((count + 15) / 16, 8192L), 256, 0, st>>>(a, b, nc, work,
                                                                          C, ldc);
  return launch_status("k_tn_sum");
}

// Chunk partials of the allocation-free entry point.  The chunking above
// makes nc * a * b at most 1024 * 64 * 64 whenever the output has at most
// 1024 tiles of 64 x 64 (a, b <= 2048 both), so this array covers every such
// product; larger ones (one chunk, a * b above the size) are refused.  Each
// GPU holds its own copy from the library's load to the process's end.
constexpr long kTnScratchElems = 1024L * kTile * kTile;  // 4 Mi doubles, 32 MiB
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) double scratch[kTnScratchElems];

namespace {
// Device address of the calling thread's current GPU's copy of `scratch`,
// looked up once per device.  The array's host-side name is not a device
// address: only this pointer may reach a kernel.
int tn_scratch(double **out) {
  int dev = 0;
  MBPT_HIP_TRY(hipGetDevice(&dev));
  MBPT_REQUIRE(dev >= 0 && dev < kMaxDevices, "mbpt_gemm_tn_f64_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    MBPT_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    MBPT_REQUIRE(p != nullptr, "mbpt_gemm_tn_f64_noalloc: no device address for the scratch array");
    g_scratch_dev[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_dev[dev];
  return MBPT_OK;
}
}  // namespace

extern "C" {

long mbpt_gemm_tn_f64_work(int n, int a, int b) {
  if (n <= 0 || a <= 0 || b <= 0) return 0;
  const int chunk = tn_chunk(n, a, b);
  return (long)((n + chunk - 1) / chunk) * a * b;
}

int mbpt_gemm_tn_f64(int n, int a, int b, const double *A, int lda, const double *B,
                     int ldb, double *C, int ldc, double *work, long work_elems,
                     void *stream) {
  MBPT_REQUIRE(n > 0 && a >= 0 && b >= 0, "mbpt_gemm_tn_f64: n must be positive, a and b non-negative");
  if (a == 0 || b == 0) return MBPT_OK;
  MBPT_REQUIRE(A && B && C && work, "mbpt_gemm_tn_f64: NULL argument");
  MBPT_REQUIRE(lda >= a && ldb >= b && ldc >= b, "mbpt_gemm_tn_f64: leading dimension too small");
  return tn_launch(n, a, b, A, lda, B, ldb, C, ldc, work, work_elems, (hipStream_t)stream);
}

long mbpt_gemm_tn_f64_noalloc_max(void) { return kTnScratchElems; }

int mbpt_gemm_tn_f64_noalloc(int n, int a, int b, const double *A, int lda,
                             const double *B, int ldb, double *C, int ldc, void *stream) {
  MBPT_REQUIRE(n > 0 && a >= 0 && b >= 0, "mbpt_gemm_tn_f64_noalloc: n must be positive, a and b non-negative");
  if (a == 0 || b == 0) return MBPT_OK;
  MBPT_REQUIRE(A && B && C, "mbpt_gemm_tn_f64_noalloc: NULL argument");
  MBPT_REQUIRE(lda >= a && ldb >= b && ldc >= b, "mbpt_gemm_tn_f64_noalloc: leading dimension too small");
  const long need = mbpt_gemm_tn_f64_work(n, a, b);
  if (need > kTnScratchElems) {
    set_error("mbpt_gemm_tn_f64_noalloc: the product needs " + std::to_string(need) +
              " doubles of partials, above the fixed scratch of " +
              std::to_string(kTnScratchElems) + " (a * b too large)");
    return MBPT_ERR_VALUE;
  }
  double *work = nullptr;
  const int rc = tn_scratch(&work);
  if (rc) return rc;
  return tn_launch(n, a, b, A, lda, B, ldb, C, ldc, work, kTnScratchElems, (hipStream_t)stream);
}

}  // extern "C"
