// Cholesky factor of the overlap S = U^T U and its inverse X = U^{-1}, FP64 (the generalized solve keeps S's factor
// in FP64 in both precisions). Blocked right-looking potrf('U') with 64-column blocks: the diagonal block is factored
// and inverted in LDS by one workgroup, the row panel is U_kk^{-T} A_k* as a product with the inverted block, and
// the trailing update is a GEMM on the matrix cores. X is then assembled block column by block column,
// X(0:j, J) = -X(0:j, 0:j) U(0:j, J) X_JJ, two GEMMs per block column. X is what the congruence H' = X^T H X and the
// back-solve C = X C' use (xTB's fact_solve path; scripts/kernels/congruence.c:1-8 of the reference).
#include "common.hpp"
#include "kernels.hpp"

namespace xtbeig {

namespace {

constexpr int NB = 64;

// One workgroup of 256 threads: factor A(k0:k0+kb, k0:k0+kb) = U^T U in LDS, write U to u and U^{-1} to x (both at
// (k0, k0); the lower parts of those blocks zero).
__global__ __launch_bounds__(256) void potrf_diag_kernel(int n, int k0, double* __restrict__ a, int lda,
                                                         double* __restrict__ u, int ldu, double* __restrict__ x,
                                                         int ldx, int* info) {
  __shared__ double L[NB][NB + 1];  // L[i][j] = A(k0 + i, k0 + j)
  __shared__ double Y[NB][NB + 1];
  __shared__ int bad;
  const int kb = min(NB, n - k0), t = threadIdx.x;
  if (*info != 0) return;
  for (int e = t; e < kb * kb; e += 256) {
    const int i = e % kb, j = e / kb;
    L[i][j] = a[idx(k0 + i, k0 + j, lda)];
  }
  if (t == 0) bad = 0;
  __syncthreads();
  for (int j = 0; j < kb; ++j) {
    // pivot
    if (t == 0) {
      const double p = L[j][j];
      if (!(p > 0.0)) bad = j + 1;
      else L[j][j] = sqrt(p);
    }
    __syncthreads();
    if (bad) break;
    const double r = L[j][j];
    for (int c = j + 1 + t; c < kb; c += 256) L[j][c] /= r;  // row j of U
    __syncthreads();
    // trailing upper triangle: L[p][q] -= U[j][p] U[j][q], j < p <= q
    const int m = kb - j - 1;
    for (int e = t; e < m * m; e += 256) {
      const int p = j + 1 + e % m, q = j + 1 + e / m;
      if (p <= q) L[p][q] -= L[j][p] * L[j][q];
    }
    __syncthreads();
  }
  if (bad) {
    if (t == 0) atomicCAS(info, 0, k0 + bad);
    return;
  }
  // Y = U^{-1} by columns: thread c solves U y = e_c (back substitution).
  if (t < kb) {
    const int c = t;
    for (int i = kb - 1; i >= 0; --i) {
      double v = (i == c) ? 1.0 : 0.0;
      if (i <= c) {
        for (int l = i + 1; l <= c; ++l) v -= L[i][l] * Y[l][c];
        v /= L[i][i];
      } else {
        v = 0.0;
      }
      Y[i][c] = v;
    }
  }
  __syncthreads();
  for (int e = t; e < kb * kb; e += 256) {
    const int i = e % kb, j = e / kb;
    u[idx(k0 + i, k0 + j, ldu)] = i <= j ? L[i][j] : 0.0;
    x[idx(k0 + i, k0 + j, ldx)] = i <= j ? Y[i][j] : 0.0;
  }
}

__global__ void check_info_zero(const int*) {}

}  // namespace

void potrf_inverse(hipStream_t s, int n, double* a, int lda, double* u, int ldu, double* x, int ldx, double* tmp,
                   int* info) {
  if (n <= 0) return;
  set_zero(s, n, n, u, ldu);
  set_zero(s, n, n, x, ldx);
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int kb = std::min(NB, n - k0), r = n - k0 - kb;
    hipLaunchKernelGGL(potrf_diag_kernel, dim3(1), dim3(256), 0, s, n, k0, a, lda, u, ldu, x, ldx, info);
    XTB_LAUNCH_CHECK();
    if (r > 0) {
      // U(k0 block, rest) = U_kk^{-T} A(k0 block, rest)
      gemm<double>(s, 'T', 'N', kb, r, kb, 1.0, x + idx(k0, k0, ldx), ldx, a + idx(k0, k0 + kb, lda), lda, 0.0,
                   u + idx(k0, k0 + kb, ldu), ldu);
      // A(rest, rest) -= U(k0 block, rest)^T U(k0 block, rest)
      gemm<double>(s, 'T', 'N', r, r, kb, -1.0, u + idx(k0, k0 + kb, ldu), ldu, u + idx(k0, k0 + kb, ldu), ldu, 1.0,
                   a + idx(k0 + kb, k0 + kb, lda), lda);
    }
  }
  // Off-diagonal blocks of X = U^{-1}, block column by block column.
  for (int j0 = NB; j0 < n; j0 += NB) {
    const int jb = std::min(NB, n - j0);
    // tmp (j0 x jb) = U(0:j0, J) X_JJ
    gemm<double>(s, 'N', 'N', j0, jb, jb, 1.0, u + idx(0, j0, ldu), ldu, x + idx(j0, j0, ldx), ldx, 0.0, tmp, j0);
    // X(0:j0, J) = -X(0:j0, 0:j0) tmp
    gemm<double>(s, 'N', 'N', j0, jb, j0, -1.0, x, ldx, tmp, j0, 0.0, x + idx(0, j0, ldx), ldx);
  }
}

}  // namespace xtbeig
