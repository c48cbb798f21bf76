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

namespace xtbeig {

namespace {

using namespace gemm_detail;

constexpr int GEMM_THREADS = 512;  // eight wavefronts per 64 x 64 tile, two per SIMD

template <typename T, bool TA, bool TB>
__global__ __launch_bounds__(GEMM_THREADS) void gemm_kernel(int M, int N, int K, T alpha, const T* __restrict__ A,
                                                            int lda, int64_t sA, const T* __restrict__ B, int ldb,
                                                            int64_t sB, T beta, T* __restrict__ C, int ldc,
                                                            int64_t sC) {
  __shared__ T As[BK][BM + PAD];
  __shared__ T Bs[BK][BN + PAD];
  A += blockIdx.z * sA;
  B += blockIdx.z * sB;
  C += blockIdx.z * sC;
  gemm_tile<T, TA, TB, GEMM_THREADS>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, blockIdx.x * BM, blockIdx.y * BN,
                                     As, Bs);
}

// The tiles (I, J) with I <= J of an n x n C only, one workgroup each (tile q = J (J + 1) / 2 + I).
template <typename T, bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_upper_kernel(int N, int K, T alpha, const T* __restrict__ A, int lda,
                                                         int64_t sA, const T* __restrict__ B, int ldb, int64_t sB,
                                                         T beta, T* __restrict__ C, int ldc, int64_t sC) {
  __shared__ T As[BK][BM + PAD];
  __shared__ T Bs[BK][BN + PAD];
  A += blockIdx.y * sA;
  B += blockIdx.y * sB;
  C += blockIdx.y * sC;
  const int q = blockIdx.x;
  int J = (int)((sqrtf(8.0f * (float)q + 1.0f) - 1.0f) * 0.5f);
  while (J * (J + 1) / 2 > q) --J;
  while ((J + 1) * (J + 2) / 2 <= q) ++J;
  const int I = q - J * (J + 1) / 2;
  gemm_tile<T, TA, TB>(N, N, K, alpha, A, lda, B, ldb, beta, C, ldc, I * BM, J * BN, As, Bs);
}

template <typename T, bool TA, bool TB>
void launch(hipStream_t s, int m, int n, int k, T alpha, const T* A, int lda, int64_t sA, const T* B, int ldb,
            int64_t sB, T beta, T* C, int ldc, int64_t sC, int batch) {
  dim3 grid(cdiv(m, BM), cdiv(n, BN), batch);
  hipLaunchKernelGGL((gemm_kernel<T, TA, TB>), grid, dim3(GEMM_THREADS), 0, s, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta,
                     C, ldc, sC);
  XTB_LAUNCH_CHECK();
}

// C = beta C when K = 0 (alpha op(A) op(B) vanishes).
template <typename T>
__global__ void scale_kernel(int M, int N, T beta, T* C, int ldc, int64_t sC) {
  C += blockIdx.z * sC;
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i < M) C[idx(i, j, ldc)] = beta == T(0) ? T(0) : beta * C[idx(i, j, ldc)];
}

}  // namespace

template <typename T>
void gemm(hipStream_t s, char ta, char tb, int m, int n, int k, T alpha, const T* A, int lda, const T* B, int ldb,
          T beta, T* C, int ldc, int batch, int64_t sA, int64_t sB, int64_t sC) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  if (k <= 0) {
    if (beta == T(1)) return;
    hipLaunchKernelGGL(scale_kernel<T>, dim3(cdiv(m, 256), n, batch), dim3(256), 0, s, m, n, beta, C, ldc, sC);
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
                                                          T* __restrict__ C) {
  __shared__ T As[BK][BM + PAD];
  __shared__ T Bs[BK][BN + PAD];
  const int z = blockIdx.z;
  const int k0 = z * kc, kz = min(kc, K - k0);
  gemm_tile<T, false, false>(M, N, kz, T(1), A + (int64_t)k0 * lda, lda, B + k0, ldb, T(0),
                             C + (int64_t)z * M * N, M, blockIdx.x * BM, blockIdx.y * BN, As, Bs);
}
}  // namespace

template <typename T>
int gemm_splitk(hipStream_t s, int m, int n, int k, const T* A, int lda, const T* B, int ldb, T* C, int kc) {
  if (m <= 0 || n <= 0 || k <= 0) return 0;
  const int splits = cdiv(k, kc);
  hipLaunchKernelGGL(gemm_splitk_kernel<T>, dim3(cdiv(m, BM), cdiv(n, BN), splits), dim3(256), 0, s, m, n, k, kc, A,
                     lda, B, ldb, C);
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
                                                        T* __restrict__ part) {
  __shared__ T As[BK][BM + PAD];
  __shared__ T Bs[BK][BN + PAD];
  const int z = blockIdx.y;
  const int k0 = z * kc, kz = min(kc, K - k0);
  const T* a0 = TA ? A + k0 : A + (int64_t)k0 * lda;
  gemm_tile<T, TA, false>(M, N, kz, T(1), a0, lda, B + k0, ldb, T(0), part + (int64_t)z * M * N, M,
                          blockIdx.x * BM, 0, As, Bs);
}
template <typename T>
__global__ void sum_parts_kernel(int M, int N, int splits, const T* __restrict__ part, T* __restrict__ C, int ldc) {
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
    hipLaunchKernelGGL((gemm_thin_kernel<T, true>), dim3(cdiv(m, BM), splits), dim3(256), 0, s, m, n, k, THIN_KC, A,
                       lda, B, ldb, part);
  else
    hipLaunchKernelGGL((gemm_thin_kernel<T, false>), dim3(cdiv(m, BM), splits), dim3(256), 0, s, m, n, k, THIN_KC,
                       A, lda, B, ldb, part);
  XTB_LAUNCH_CHECK();
  const int64_t mn = (int64_t)m * n;
  hipLaunchKernelGGL(sum_parts_kernel<T>, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, s, m, n, splits, part, C,
                     ldc);
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
  hipLaunchKernelGGL((gemm_upper_kernel<T, true, false>), dim3(nt * (nt + 1) / 2, batch), dim3(256), 0, s, n, k,
                     alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC);
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
