// Dense matrix products on the CDNA4 matrix cores: C = alpha op(A) op(B) + beta C, column-major, FP64 through
// v_mfma_f64_16x16x4f64 and FP32 through v_mfma_f32_16x16x4f32 (FP32 accumulation). Every GEMM-shaped step of the
// eigensolver goes through here: the congruence H' = X^T H X and the back-solve C = X C', the syr2k-shaped trailing
// update of the dense-to-band reduction, the merge products of divide and conquer, the compact-WY back-transforms
// and the refinement residuals.
//
// The tile itself is in gemm_tile.hpp.
#include "common.hpp"
#include "gemm_tile.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <type_traits>

namespace xtbeig {

namespace {

using namespace gemm_detail;

// Tile shapes of gemm_kernel: 64 rows by TBN columns, eight wavefronts (two per SIMD): 32 x 16 each for TBN 64,
// 32 x 32 each for TBN 128.
template <int TBN> struct Shape;
template <> struct Shape<64> { static constexpr int NT = 512; };
template <> struct Shape<128> { static constexpr int NT = 512; };

template <typename T, bool TA, bool TB, int TBN>
__global__ __launch_bounds__(Shape<TBN>::NT) void gemm_kernel(int M, int N, int K, T alpha, const T* __restrict__ A,
                                                              int lda, int64_t sA, const T* __restrict__ B, int ldb,
                                                              int64_t sB, T beta, T* __restrict__ C, int ldc,
                                                              int64_t sC, Zb zb) {
  __shared__ __attribute__((aligned(16))) T lds[tile_lds_elems<T, TBN>()];
  const int z = zb.z();
  A = zb.at(A) + z * sA;
  B = zb.at(B) + z * sB;
  C = zb.at(C) + z * sC;
  gemm_tile<T, TA, TB, Shape<TBN>::NT, TBN>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, blockIdx.x * BM,
                                            blockIdx.y * TBN, lds);
}

// The tiles (I, J) with I <= J of an n x n C only, one workgroup each (tile q = J (J + 1) / 2 + I).
template <typename T, bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_upper_kernel(int N, int K, T alpha, const T* __restrict__ A, int lda,
                                                         int64_t sA, const T* __restrict__ B, int ldb, int64_t sB,
                                                         T beta, T* __restrict__ C, int ldc, int64_t sC, Zb zb) {
  __shared__ __attribute__((aligned(16))) T lds[tile_lds_elems<T>()];
  A = zb.at(A) + blockIdx.y * sA;
  B = zb.at(B) + blockIdx.y * sB;
  C = zb.at(C) + blockIdx.y * sC;
  const int q = blockIdx.x;
  int J = (int)((sqrtf(8.0f * (float)q + 1.0f) - 1.0f) * 0.5f);
  while (J * (J + 1) / 2 > q) --J;
  while ((J + 1) * (J + 2) / 2 <= q) ++J;
  const int I = q - J * (J + 1) / 2;
  gemm_tile<T, TA, TB>(N, N, K, alpha, A, lda, B, ldb, beta, C, ldc, I * BM, J * BN, lds);
}

template <typename T, bool TA, bool TB, int TBN>
void launch_w(hipStream_t s, int m, int n, int k, T alpha, const T* A, int lda, int64_t sA, const T* B, int ldb,
              int64_t sB, T beta, T* C, int ldc, int64_t sC, int batch) {
  dim3 grid(cdiv(m, BM), cdiv(n, TBN), zgrid(batch));
  hipLaunchKernelGGL((gemm_kernel<T, TA, TB, TBN>), grid, dim3(Shape<TBN>::NT), 0, s, m, n, k, alpha, A, lda, sA, B,
                     ldb, sB, beta, C, ldc, sC, zb(batch));
  XTB_LAUNCH_CHECK();
}

template <typename T, bool TA, bool TB>
void launch(hipStream_t s, int m, int n, int k, T alpha, const T* A, int lda, int64_t sA, const T* B, int ldb,
            int64_t sB, T beta, T* C, int ldc, int64_t sC, int batch) {
  const int w = gemm_tile_width(m, n, batch * current_batch().count, std::is_same<T, double>::value);
  if (w == 128) launch_w<T, TA, TB, 128>(s, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
  else launch_w<T, TA, TB, 64>(s, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
}

// C = beta C when K = 0 (alpha op(A) op(B) vanishes).
template <typename T>
__global__ void scale_kernel(int M, int N, T beta, T* C, int ldc, int64_t sC, Zb zb) {
  C = zb.at(C) + zb.z() * sC;
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i < M) C[idx(i, j, ldc)] = beta == T(0) ? T(0) : beta * C[idx(i, j, ldc)];
}

}  // namespace

namespace {
std::atomic<int> tile_override{0};
// tuned width per bucket of the product's count of 64 x 64 tiles (0: not tuned), [0] FP64, [1] FP32
std::atomic<int> tile_table[2][GEMM_TILE_BUCKETS];

int device_cus() {
  static std::atomic<int> cus[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int c = cus[dev].load(std::memory_order_relaxed);
  if (c == 0) {
    if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0) c = 256;
    cus[dev].store(c, std::memory_order_relaxed);
  }
  return c;
}

// The round model's cost per tile column of the 128-wide tile relative to the 64-wide one (x 1000), from n x n x n
// products at n = 600, 1204 and 1788 on MI355X: the 64-wide tiles run two workgroups per CU and win in FP64 where
// the round counts tie; in FP32 the 128-wide tile is the cheaper per column.
constexpr int MODEL_COST_128[2] = {1000, 950};
}  // namespace

int gemm_tile_bucket(int64_t tiles64) {
  int q = 0;
  while (q + 1 < GEMM_TILE_BUCKETS && gemm_tile_bucket_start(q + 1) <= tiles64) ++q;
  return q;
}
// bucket q starts at ceil(2^(q / 2)) tiles of 64 x 64 (half-octave buckets), at least one past bucket q - 1's start
int64_t gemm_tile_bucket_start(int q) {
  int64_t start = 1;
  for (int r = 1; r <= q; ++r) start = std::max(start + 1, (int64_t)std::ceil(std::pow(2.0, 0.5 * r) - 1e-9));
  return start;
}

void set_gemm_tile(int width) { tile_override.store(width, std::memory_order_relaxed); }
int get_gemm_tile() { return tile_override.load(std::memory_order_relaxed); }
void set_gemm_tile_table(bool fp64, const int* width) {
  for (int q = 0; q < GEMM_TILE_BUCKETS; ++q) tile_table[fp64 ? 0 : 1][q].store(width[q], std::memory_order_relaxed);
}
void get_gemm_tile_table(bool fp64, int* width) {
  for (int q = 0; q < GEMM_TILE_BUCKETS; ++q) width[q] = tile_table[fp64 ? 0 : 1][q].load(std::memory_order_relaxed);
}

// The tile width of an m x n product batched `batch` times: a forced width, else the tuned width of the product's
// bucket (by its count of 64 x 64 tiles, which sets how the grid fills the CUs), else a round model: a launch of T
// tiles on C compute units takes ceil(T / C) rounds of one tile each, a tile of width w costing w x its cost per
// column. The width only partitions C's columns among workgroups: every element sums the same products in the same
// order whatever the width, so the results do not depend on it.
int gemm_tile_width(int m, int n, int batch, bool fp64) {
  const int ov = tile_override.load(std::memory_order_relaxed);
  if (ov == 64 || ov == 128) return ov;
  const int64_t t64 = (int64_t)batch * cdiv(m, BM) * cdiv(n, 64);
  const int tuned = tile_table[fp64 ? 0 : 1][gemm_tile_bucket(t64)].load(std::memory_order_relaxed);
  if (tuned == 64 || tuned == 128) return tuned;
  const int64_t cu = device_cus();
  const int64_t t128 = (int64_t)batch * cdiv(m, BM) * cdiv(n, 128);
  const int64_t c64 = (t64 + cu - 1) / cu * 64 * 1000;
  const int64_t c128 = (t128 + cu - 1) / cu * 128 * MODEL_COST_128[fp64 ? 0 : 1];
  return c128 < c64 ? 128 : 64;
}

// Autotuning of the tile width: for each bucket of 64 x 64 tile counts that the square products of n in [nmin, nmax]
// reach, an n x n x n product with about the bucket's middle tile count, timed with events on stream s in both
// widths (best of 3 after a warm-up), the faster stored. a, b, c: nmax x nmax scratch each. Returns the buckets tuned
// (first tile count, width) through the output arrays (GEMM_TILE_BUCKETS entries each, may be null).
template <typename T>
int autotune_gemm_tiles(hipStream_t s, int nmin, int nmax, T* a, T* b, T* c, int* bucket_tiles, int* widths) {
  const bool fp64 = std::is_same<T, double>::value;
  const int qlo = gemm_tile_bucket((int64_t)cdiv(nmin, 64) * cdiv(nmin, 64));
  const int qhi = gemm_tile_bucket((int64_t)cdiv(nmax, 64) * cdiv(nmax, 64));
  hipEvent_t e0, e1;
  XTB_HIP(hipEventCreate(&e0));
  XTB_HIP(hipEventCreate(&e1));
  const int keep = get_gemm_tile();
  int count = 0;
  for (int q = qlo; q <= qhi; ++q) {
    const double lo = (double)gemm_tile_bucket_start(q);
    const double hi = (double)(q + 1 < GEMM_TILE_BUCKETS ? gemm_tile_bucket_start(q + 1) - 1 : 4 * lo);
    const int side = std::max(1, (int)std::lround(std::sqrt(std::sqrt(lo * hi))));  // tiles per side
    const int n = std::min(nmax, std::max(nmin, 64 * side - 8));
    int best_w = 64;
    float best_ms = 1e30f;
    for (int w : {64, 128}) {
      set_gemm_tile(w);
      float tmin = 1e30f;
      for (int r = 0; r < 4; ++r) {
        XTB_HIP(hipEventRecord(e0, s));
        gemm<T>(s, 'N', 'N', n, n, n, T(1), a, n, b, n, T(0), c, n);
        XTB_HIP(hipEventRecord(e1, s));
        XTB_HIP(hipEventSynchronize(e1));
        float t = 0;
        XTB_HIP(hipEventElapsedTime(&t, e0, e1));
        if (r > 0) tmin = std::min(tmin, t);
      }
      if (tmin < best_ms) {
        best_ms = tmin;
        best_w = w;
      }
    }
    tile_table[fp64 ? 0 : 1][q].store(best_w, std::memory_order_relaxed);
    if (bucket_tiles) bucket_tiles[count] = (int)gemm_tile_bucket_start(q);
    if (widths) widths[count] = best_w;
    ++count;
  }
  set_gemm_tile(keep);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return count;
}
template int autotune_gemm_tiles<double>(hipStream_t, int, int, double*, double*, double*, int*, int*);
template int autotune_gemm_tiles<float>(hipStream_t, int, int, float*, float*, float*, int*, int*);

template <typename T>
void gemm(hipStream_t s, char ta, char tb, int m, int n, int k, T alpha, const T* A, int lda, const T* B, int ldb,
          T beta, T* C, int ldc, int batch, int64_t sA, int64_t sB, int64_t sC) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  if (k <= 0) {
    if (beta == T(1)) return;
    hipLaunchKernelGGL(scale_kernel<T>, dim3(cdiv(m, 256), n, zgrid(batch)), dim3(256), 0, s, m, n, beta, C, ldc, sC,
                       zb(batch));
    XTB_LAUNCH_CHECK();
    return;
  }
  const bool TA = (ta == 'T' || ta == 't'), TB = (tb == 'T' || tb == 't');
  if (!TA && !TB) launch<T, false, false>(s, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
  else if (!TA && TB) launch<T, false, true>(s, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
  else if (TA && !TB) launch<T, true, false>(s, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
  else launch<T, true, true>(s, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
}

// C_z = op(A)(:, K_z) op(B)(K_z, :) for K split into chunks of kc: partial products written to c + z M N (ld M), one
// grid layer per chunk. Returns the number of chunks. A is M x K (ld lda), B K x N (ld ldb), no transposes.
namespace {
template <typename T>
__global__ __launch_bounds__(256) void gemm_splitk_kernel(int M, int N, int K, int kc, const T* __restrict__ A,
                                                          int lda, const T* __restrict__ B, int ldb,
                                                          T* __restrict__ C, Zb zb) {
  __shared__ __attribute__((aligned(16))) T lds[tile_lds_elems<T>()];
  const int z = zb.z();
  A = zb.at(A);
  B = zb.at(B);
  C = zb.at(C);
  const int k0 = z * kc, kz = min(kc, K - k0);
  gemm_tile<T, false, false>(M, N, kz, T(1), A + (int64_t)k0 * lda, lda, B + k0, ldb, T(0),
                             C + (int64_t)z * M * N, M, blockIdx.x * BM, blockIdx.y * BN, lds);
}
}  // namespace

template <typename T>
int gemm_splitk(hipStream_t s, int m, int n, int k, const T* A, int lda, const T* B, int ldb, T* C, int kc) {
  if (m <= 0 || n <= 0 || k <= 0) return 0;
  const int splits = cdiv(k, kc);
  hipLaunchKernelGGL(gemm_splitk_kernel<T>, dim3(cdiv(m, BM), cdiv(n, BN), zgrid(splits)), dim3(256), 0, s, m, n, k,
                     kc, A, lda, B, ldb, C, zb(splits));
  XTB_LAUNCH_CHECK();
  return splits;
}
// C = op(A) B for a thin B (n <= 64 columns: one column tile): the K range split in chunks of kc, one grid layer per
// chunk writing its partial product to part + z M N (ld M), then the partials summed in chunk order (fixed, so the
// result does not depend on scheduling). A grid of cdiv(M, 64) x 1 tiles alone would leave most CUs idle.
namespace {
template <typename T, bool TA>
__global__ __launch_bounds__(256) void gemm_thin_kernel(int M, int N, int K, int kc, const T* __restrict__ A,
                                                        int lda, const T* __restrict__ B, int ldb,
                                                        T* __restrict__ part, Zb zb) {
  __shared__ __attribute__((aligned(16))) T lds[tile_lds_elems<T>()];
  const int z = blockIdx.y;
  A = zb.at(A);
  B = zb.at(B);
  part = zb.at(part);
  const int k0 = z * kc, kz = min(kc, K - k0);
  const T* a0 = TA ? A + k0 : A + (int64_t)k0 * lda;
  gemm_tile<T, TA, false>(M, N, kz, T(1), a0, lda, B + k0, ldb, T(0), part + (int64_t)z * M * N, M,
                          blockIdx.x * BM, 0, lds);
}
template <typename T>
__global__ void sum_parts_kernel(int M, int N, int splits, const T* __restrict__ part, T* __restrict__ C, int ldc,
                                 Zb zb) {
  part = zb.at(part);
  C = zb.at(C);
  const int64_t mn = (int64_t)M * N;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= mn) return;
  T acc = part[q];
  for (int z = 1; z < splits; ++z) acc += part[z * mn + q];
  const int i = (int)(q % M), j = (int)(q / M);
  C[idx(i, j, ldc)] = acc;
}
}  // namespace

size_t gemm_thin_parts(int m, int n, int k) {
  const int splits = cdiv(k, THIN_KC);
  return splits > 1 ? (size_t)splits * m * n : 0;
}

template <typename T>
void gemm_thin(hipStream_t s, char ta, int m, int n, int k, const T* A, int lda, const T* B, int ldb, T* C, int ldc,
               T* part) {
  if (n > BN) throw std::runtime_error("gemm_thin: more than 64 columns");
  if (m <= 0 || n <= 0) return;
  const int splits = cdiv(k, THIN_KC);
  if (splits <= 1) {  // one chunk: the plain product, no partials
    gemm<T>(s, ta, 'N', m, n, k, T(1), A, lda, B, ldb, T(0), C, ldc);
    return;
  }
  if (ta == 'T' || ta == 't')
    hipLaunchKernelGGL((gemm_thin_kernel<T, true>), dim3(cdiv(m, BM), splits, zgrid()), dim3(256), 0, s, m, n, k,
                       THIN_KC, A, lda, B, ldb, part, zb());
  else
    hipLaunchKernelGGL((gemm_thin_kernel<T, false>), dim3(cdiv(m, BM), splits, zgrid()), dim3(256), 0, s, m, n, k,
                       THIN_KC, A, lda, B, ldb, part, zb());
  XTB_LAUNCH_CHECK();
  const int64_t mn = (int64_t)m * n;
  hipLaunchKernelGGL(sum_parts_kernel<T>, dim3((unsigned)((mn + 255) / 256), 1, zgrid()), dim3(256), 0, s, m, n,
                     splits, part, C, ldc, zb());
  XTB_LAUNCH_CHECK();
}
template void gemm_thin<double>(hipStream_t, char, int, int, int, const double*, int, const double*, int, double*,
                                int, double*);
template void gemm_thin<float>(hipStream_t, char, int, int, int, const float*, int, const float*, int, float*, int,
                               float*);

template <typename T>
void gemm_upper_tn(hipStream_t s, int n, int k, T alpha, const T* A, int lda, const T* B, int ldb, T beta, T* C,
                   int ldc, int batch, int64_t sA, int64_t sB, int64_t sC) {
  static_assert(BM == BN, "square tiles");
  if (n <= 0 || batch <= 0) return;
  if (k <= 0) {
    gemm<T>(s, 'T', 'N', n, n, k, alpha, A, lda, B, ldb, beta, C, ldc, batch, sA, sB, sC);
    return;
  }
  const int nt = cdiv(n, BM);
  hipLaunchKernelGGL((gemm_upper_kernel<T, true, false>), dim3(nt * (nt + 1) / 2, batch, zgrid()), dim3(256), 0, s,
                     n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, zb());
  XTB_LAUNCH_CHECK();
}
template void gemm_upper_tn<double>(hipStream_t, int, int, double, const double*, int, const double*, int, double,
                                    double*, int, int, int64_t, int64_t, int64_t);
template void gemm_upper_tn<float>(hipStream_t, int, int, float, const float*, int, const float*, int, float, float*,
                                   int, int, int64_t, int64_t, int64_t);

template int gemm_splitk<double>(hipStream_t, int, int, int, const double*, int, const double*, int, double*, int);
template int gemm_splitk<float>(hipStream_t, int, int, int, const float*, int, const float*, int, float*, int);

template void gemm<double>(hipStream_t, char, char, int, int, int, double, const double*, int, const double*, int,
                           double, double*, int, int, int64_t, int64_t, int64_t);
template void gemm<float>(hipStream_t, char, char, int, int, int, float, const float*, int, const float*, int, float,
                          float*, int, int, int64_t, int64_t, int64_t);

}  // namespace xtbeig
