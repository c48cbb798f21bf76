// Cholesky factor of the overlap S = U^T U and its inverse X = U^{-1}, FP64 (the generalized solve keeps S's factor
// in FP64 in both precisions). Blocked right-looking potrf('U') with 64-column blocks: the diagonal block is factored
// and inverted by one workgroup (4 x 4 sub-blocks per thread in registers), the row panel is U_kk^{-T} A_k* as a product with the inverted block,
// and the trailing update is a GEMM on the matrix cores. X is then assembled by recursive halving,
// X12 = -X11 U12 X22 (two GEMMs per split). X is what the congruence H' = X^T H X and the back-solve C = X C' use
// (xTB's fact_solve path; scripts/kernels/congruence.c:1-8 of the reference).
#include "common.hpp"
#include "kernels.hpp"

namespace xtbeig {

namespace {

constexpr int NB = 64;

// Factor A(k0:k0+kb, k0:k0+kb) = U^T U and invert U, by one workgroup of 256 threads (padded with the identity past
// kb): thread (bi, bj) = (t % 16, t / 16) owns the 4 x 4 sub-block of rows 4 bi .., columns 4 bj .. in registers.
// Right-looking on unnormalized rows: at step k the owners of row k publish it through LDS (double-buffered by
// parity, so one barrier per step), every thread subtracts A(k, i) A(k, j) / A(k, k) from its elements with i, j > k,
// and U(i, j) = A(i, j) / sqrt(A(i, i)) at the end (A's rows as they stand after their own step).
__global__ __launch_bounds__(256) void potrf_diag_kernel(int n, int k0, double* __restrict__ a, int lda,
                                                         double* __restrict__ u, int ldu, double* __restrict__ x,
                                                         int ldx, int* info) {
  __shared__ double rowb[2][NB];
  __shared__ double piv[NB];
  __shared__ double Us[NB][NB + 1];
  __shared__ double Xs[NB][NB + 1];
  __shared__ double Ts[NB][NB + 1];
  if (*info != 0) return;
  const int kb = min(NB, n - k0), t = threadIdx.x;
  const int bi = t % 16, bj = t / 16, i0 = 4 * bi, j0 = 4 * bj;
  double c[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = i0 + r, j = j0 + q;
      c[r][q] = (i < kb && j < kb) ? a[idx(k0 + i, k0 + j, lda)] : (i == j ? 1.0 : 0.0);
    }
  int bad = 0;
  // right-looking on unnormalized rows: after step k, row k holds U(k, k) U(k, j); every thread owns its elements
  for (int k = 0; k < NB; ++k) {
    const int par = k & 1;
    if (bi == (k >> 2)) {
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (i0 + r == k) {
#pragma unroll
          for (int q = 0; q < 4; ++q) rowb[par][j0 + q] = c[r][q];
        }
    }
    __syncthreads();
    const double p = rowb[par][k];
    if (!(p > 0.0)) {
      bad = k0 + k + 1;
      break;
    }
    if (t == 0) piv[k] = p;
    const double rp = 1.0 / p;
    double ri[4], rj[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      ri[r] = rowb[par][i0 + r] * rp;
      rj[r] = rowb[par][j0 + r];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (i0 + r > k && j0 + q > k) c[r][q] -= ri[r] * rj[q];
  }
  if (bad) {
    if (t == 0) atomicCAS(info, 0, bad);
    return;
  }
  __syncthreads();
  // U(i, j) = c(i, j) / sqrt(piv[i]) for i <= j
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const double s = 1.0 / sqrt(piv[i0 + r]);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = i0 + r, j = j0 + q;
      Us[i][j] = (i <= j) ? c[r][q] * s : 0.0;
    }
  }
  __syncthreads();
  // X = U^{-1}: the 4 x 4 diagonal blocks by their owners (back substitution in registers), then for block sizes
  // s = 4, 8, 16, 32 the off-diagonal block of every sibling pair A = [c, c + s), B = [c + s, c + 2s):
  // X_AB = -X_AA (U_AB X_BB), through T = U_AB X_BB in LDS
  if (bi == bj) {
    double xb[4][4];
#pragma unroll
    for (int cc = 0; cc < 4; ++cc)
#pragma unroll
      for (int r = 3; r >= 0; --r) {
        double sum = (r == cc) ? 1.0 : 0.0;
#pragma unroll
        for (int l = r + 1; l < 4; ++l) sum -= Us[i0 + r][i0 + l] * xb[l][cc];
        xb[r][cc] = (r <= cc) ? sum / Us[i0 + r][i0 + r] : 0.0;
      }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int q = 0; q < 4; ++q) Xs[i0 + r][i0 + q] = xb[r][q];
  }
  if (bi > bj) {
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int q = 0; q < 4; ++q) Xs[i0 + r][j0 + q] = 0.0;
  }
  __syncthreads();
  for (int s = 4; s < NB; s *= 2) {
    // this thread's sub-block lies in some pair's X_AB: rows in an A half, columns in the B half of the same pair
    const bool mine = ((i0 / s) % 2 == 0) && (j0 / (2 * s) == i0 / (2 * s)) && ((j0 / s) % 2 == 1);
    double tv[4][4];
    if (mine) {
      const int cB = (j0 / s) * s;
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          double sum = 0;
          for (int l = cB; l < cB + s; ++l) sum += Us[i0 + r][l] * Xs[l][j0 + q];
          tv[r][q] = sum;
        }
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) Ts[i0 + r][j0 + q] = tv[r][q];
    }
    __syncthreads();
    if (mine) {
      const int cA = (i0 / s) * s;
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          double sum = 0;
          for (int l = cA; l < cA + s; ++l) sum += Xs[i0 + r][l] * Ts[l][j0 + q];
          tv[r][q] = -sum;
        }
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) Xs[i0 + r][j0 + q] = tv[r][q];
    }
    __syncthreads();
  }
  // write U and X (zero below the diagonal), lanes along rows
  for (int e = t; e < NB * NB; e += 256) {
    const int i = e % NB, j = e / NB;
    if (i < kb && j < kb) {
      u[idx(k0 + i, k0 + j, ldu)] = Us[i][j];
      x[idx(k0 + i, k0 + j, ldx)] = Xs[i][j];
    }
  }
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
  inverse_blocks(s, 0, n, u, ldu, x, ldx, tmp);
}

}  // namespace xtbeig
