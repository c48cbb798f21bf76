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

template <typename T, bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_kernel(int M, int N, int K, T alpha, const T* __restrict__ A, int lda,
                                                   int64_t sA, const T* __restrict__ B, int ldb, int64_t sB, T beta,
                                                   T* __restrict__ C, int ldc, int64_t sC) {
  __shared__ T As[BK][BM + PAD];
  __shared__ T Bs[BK][BN + PAD];
  A += blockIdx.z * sA;
  B += blockIdx.z * sB;
  C += blockIdx.z * sC;
  gemm_tile<T, TA, TB>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, blockIdx.x * BM, blockIdx.y * BN, As, Bs);
}

template <typename T, bool TA, bool TB>
void launch(hipStream_t s, int m, int n, int k, T alpha, const T* A, int lda, int64_t sA, const T* B, int ldb,
            int64_t sB, T beta, T* C, int ldc, int64_t sC, int batch) {
  dim3 grid(cdiv(m, BM), cdiv(n, BN), batch);
  hipLaunchKernelGGL((gemm_kernel<T, TA, TB>), grid, dim3(256), 0, s, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta,
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

template void gemm<double>(hipStream_t, char, char, int, int, int, double, const double*, int, const double*, int,
                           double, double*, int, int, int64_t, int64_t, int64_t);
template void gemm<float>(hipStream_t, char, char, int, int, int, float, const float*, int, const float*, int, float,
                          float*, int, int, int64_t, int64_t, int64_t);

}  // namespace xtbeig
