// Cholesky factor of the overlap S = U^T U and its inverse X = U^{-1}, FP64 (the generalized solve keeps S's factor
// in FP64 in both precisions). Blocked right-looking potrf('U') with 64-column blocks: the diagonal block is factored
// and inverted by one wavefront in registers, the row panel is U_kk^{-T} A_k* as a product with the inverted block,
// and the trailing update is a GEMM on the matrix cores. X is then assembled by recursive halving,
// X12 = -X11 U12 X22 (two GEMMs per split). X is what the congruence H' = X^T H X and the back-solve C = X C' use
// (xTB's fact_solve path; scripts/kernels/congruence.c:1-8 of the reference).
#include "common.hpp"
#include "kernels.hpp"

namespace xtbeig {

namespace {

constexpr int NB = 64;

__device__ inline double readlane_d(double v, int l) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)bits, l);
  const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// One wavefront: factor A(k0:k0+kb, k0:k0+kb) = U^T U and invert U, lane j owning column j of the block (padded with
// the identity past kb). Right-looking Cholesky with the pivot row read across lanes (readlane), column j's
// entries in registers; then U goes to LDS and X = U^{-1} is solved row by row from the bottom, lane c holding
// column c of X (X(i, c) = (delta_ic - sum_{l > i} U(i, l) X(l, c)) / U(i, i), U's rows read as LDS broadcasts).
// Writes U and X at (k0, k0) (zero below the diagonal); info = k0 + j + 1 for a non-positive pivot j.
__global__ __launch_bounds__(64) void potrf_diag_kernel(int n, int k0, double* __restrict__ a, int lda,
                                                        double* __restrict__ u, int ldu, double* __restrict__ x,
                                                        int ldx, int* info) {
  __shared__ double Us[NB][NB + 1];
  if (*info != 0) return;
  const int kb = min(NB, n - k0), j = threadIdx.x;
  double c[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) c[i] = (i < kb && j < kb) ? a[idx(k0 + i, k0 + j, lda)] : (i == j ? 1.0 : 0.0);
  int bad = 0;
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    const double p = readlane_d(c[k], k);
    if (!(p > 0.0)) {
      bad = k + 1;
      break;
    }
    const double r = sqrt(p);
    c[k] = (j > k) ? c[k] / r : (j == k ? r : c[k]);  // row k of U: U(k, j) = A(k, j) / r
#pragma unroll
    for (int i = k + 1; i < NB; ++i) {
      const double uki = readlane_d(c[k], i);
      const double v = c[i] - uki * c[k];
      c[i] = (j >= i) ? v : c[i];
    }
  }
  if (bad) {
    if (j == 0) atomicCAS(info, 0, k0 + bad);
    return;
  }
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const double v = (i <= j) ? c[i] : 0.0;
    Us[i][j] = v;
    if (i < kb && j < kb) u[idx(k0 + i, k0 + j, ldu)] = v;
  }
  __syncthreads();
  double y[NB];
#pragma unroll
  for (int i = NB - 1; i >= 0; --i) {
    double s0 = (i == j) ? 1.0 : 0.0, s1 = 0.0;
#pragma unroll
    for (int l = i + 1; l < NB; l += 2) {
      s0 -= Us[i][l] * y[l];
      if (l + 1 < NB) s1 -= Us[i][l + 1] * y[l + 1];
    }
    y[i] = (i <= j) ? (s0 + s1) / Us[i][i] : 0.0;
  }
#pragma unroll
  for (int i = 0; i < NB; ++i)
    if (i < kb && j < kb) x[idx(k0 + i, k0 + j, ldx)] = y[i];
}

}  // namespace

namespace {

// X(lo:hi, lo:hi) = U(lo:hi, lo:hi)^{-1} from its diagonal NB-blocks (already in X): split on a block boundary m,
// invert both halves, then X12 = -X11 (U12 X22), two GEMMs whose shapes halve with the depth (the largest,
// n/2 x n/2 x n/2, first fills the machine).
void inverse_blocks(hipStream_t s, int lo, int hi, const double* u, int ldu, double* x, int ldx, double* tmp) {
  if (hi - lo <= NB) return;
  const int nblk = (hi - lo + NB - 1) / NB;
  const int m = lo + (nblk / 2) * NB;
  inverse_blocks(s, lo, m, u, ldu, x, ldx, tmp);
  inverse_blocks(s, m, hi, u, ldu, x, ldx, tmp);
  gemm<double>(s, 'N', 'N', m - lo, hi - m, hi - m, 1.0, u + idx(lo, m, ldu), ldu, x + idx(m, m, ldx), ldx, 0.0, tmp,
               m - lo);
  gemm<double>(s, 'N', 'N', m - lo, hi - m, m - lo, -1.0, x + idx(lo, lo, ldx), ldx, tmp, m - lo, 0.0,
               x + idx(lo, m, ldx), ldx);
}

}  // namespace

size_t potrf_workspace(int n) {
  const size_t h = (size_t)(n / 2 + NB);
  return h * h + (size_t)n * NB + 64;
}

void potrf_inverse(hipStream_t s, int n, double* a, int lda, double* u, int ldu, double* x, int ldx, double* tmp,
                   int* info) {
  if (n <= 0) return;
  set_zero(s, n, n, u, ldu);
  set_zero(s, n, n, x, ldx);
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int kb = std::min(NB, n - k0), r = n - k0 - kb;
    hipLaunchKernelGGL(potrf_diag_kernel, dim3(1), dim3(64), 0, s, n, k0, a, lda, u, ldu, x, ldx, info);
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
  inverse_blocks(s, 0, n, u, ldu, x, ldx, tmp);
}

}  // namespace xtbeig
