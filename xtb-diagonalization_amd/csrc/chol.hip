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
// (k0, k0); the lower parts of those blocks zero). Right-looking, two barriers per column; the trailing update maps
// column q to threads q, q + 64, q + 128, q + 192 (rows interleaved by four). The inverse solves U Y = I with four
// lanes of one wavefront per column of Y, partial sums combined by shuffles (no barrier between rows).
__global__ __launch_bounds__(256) void potrf_diag_kernel(int n, int k0, double* __restrict__ a, int lda,
                                                         double* __restrict__ u, int ldu, double* __restrict__ x,
                                                         int ldx, int* info) {
  __shared__ double L[NB][NB + 1];  // L[i][j] = A(k0 + i, k0 + j)
  __shared__ double Y[NB][NB + 1];
  const int kb = min(NB, n - k0), t = threadIdx.x;
  if (*info != 0) return;
  for (int e = t; e < NB * NB; e += 256) {
    const int i = e & (NB - 1), j = e / NB;
    L[i][j] = (i < kb && j < kb) ? a[idx(k0 + i, k0 + j, lda)] : 0.0;
  }
  __syncthreads();
  int bad = 0;
  for (int j = 0; j < kb; ++j) {
    const double p = L[j][j];
    if (!(p > 0.0)) { bad = j + 1; break; }
    const double r = sqrt(p);
    __syncthreads();  // every thread has read the pivot
    if (t == 0) L[j][j] = r;
    for (int c = j + 1 + t; c < kb; c += 256) L[j][c] /= r;
    __syncthreads();
    const int q = t & (NB - 1);
    if (q > j && q < kb) {
      const double ljq = L[j][q];
      for (int pp = j + 1 + (t >> 6); pp <= q; pp += 4) L[pp][q] -= L[j][pp] * ljq;
    }
    __syncthreads();
  }
  if (bad) {
    if (t == 0) atomicCAS(info, 0, k0 + bad);
    return;
  }
  {
    const int c = t >> 2, part = t & 3;
    for (int i = kb - 1; i >= 0; --i) {
      double sum = 0.0;
      for (int l = i + 1 + part; l <= c && l < kb; l += 4) sum += L[i][l] * Y[l][c];
      sum += __shfl_xor(sum, 1, 64);
      sum += __shfl_xor(sum, 2, 64);
      if (part == 0 && c < kb) Y[i][c] = (i <= c) ? ((i == c ? 1.0 : 0.0) - sum) / L[i][i] : 0.0;
      __builtin_amdgcn_wave_barrier();
    }
  }
  __syncthreads();
  for (int e = t; e < kb * kb; e += 256) {
    const int i = e % kb, j = e / kb;
    u[idx(k0 + i, k0 + j, ldu)] = i <= j ? L[i][j] : 0.0;
    x[idx(k0 + i, k0 + j, ldx)] = i <= j ? Y[i][j] : 0.0;
  }
}

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
