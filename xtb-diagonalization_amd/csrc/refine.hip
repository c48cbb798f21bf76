// FP64 refinement of approximate generalized eigenpairs (H C = S C Lambda, C^T S C = I), Ogita-Aishima style with
// the S inner product, and the eigenpair sort. The reference refines FP32 eigenpairs in FP64 (scripts/kernels/
// fp32_error.c:95-158; the S-inner-product variant in scripts/kernels/refine.f90:254-325):
//     R = I - C^T S C,  A = C^T H C,  lambda_i = A_ii / (1 - R_ii),
//     E_ij = (A_ij + lambda_j R_ij) / (lambda_j - lambda_i) for resolved pairs i != j,  E_ii = R_ii / 2,
//     C <- C + C E.
// The four products and the update run as FP64 GEMMs on the matrix cores. A pair is treated as a cluster (E_ij =
// R_ij / 2: its subspace converges, the basis inside it is not rotated) when its Newton correction would not be
// small, |E_ij| > 1/4, which covers exactly and nearly degenerate pairs (the gap below what the current accuracy
// resolves); as the residual shrinks, near-degenerate pairs separate. The diagonal carries the second-order terms of
// the new Gram matrix, diag((I + E)^T (I - R) (I + E)) = 1 - R_ii + 2 E_ii + sum_k E_ki^2 - 2 sum_k R_ik E_ki + O(3):
//     E_ii = (R_ii - sum_{k != i} E_ki^2 + 2 sum_{k != i} R_ik E_ki) / 2,
// so that a column rotated by angle theta within a close pair keeps its S-norm to O(theta^3) instead of O(theta^2)
// (with E_ii = R_ii / 2 alone, two steps from FP32 leave the norms of such pairs off by theta^4, about 1e-9 for a gap
// of 2e-5).
#include "common.hpp"
#include "kernels.hpp"

namespace xtbeig {

namespace {

// R <- I - R (R = C^T S C on entry), lambda_i = A_ii / (C^T S C)_ii. Only the upper triangles of R and A are formed
// (the products are symmetric), so each lower element (j, i) is set from (i, j). One workgroup per 64 x 64 tile
// (I, J), I <= J, of the upper triangle: the tile is read and written down its columns, and its mirror image, staged
// transposed in LDS, is written down the columns of tile (J, I).
__global__ __launch_bounds__(256) void oa_prepare_kernel(int n, double* __restrict__ r, int ldr,
                                                         double* __restrict__ a, int lda, double* __restrict__ lam) {
  const int I = blockIdx.x, J = blockIdx.y;
  if (I > J) return;
  __shared__ double tr[64][65], ta[64][65];
  const int t = threadIdx.x, i0 = I * 64, j0 = J * 64;
#pragma unroll 4
  for (int e = t; e < 64 * 64; e += 256) {
    const int ii = e % 64, jj = e / 64, i = i0 + ii, j = j0 + jj;
    if (i >= n || j >= n || i > j) continue;
    const double v = r[idx(i, j, ldr)];
    if (i == j) {
      lam[i] = a[idx(i, i, lda)] / v;
      r[idx(i, i, ldr)] = 1.0 - v;
    } else {
      r[idx(i, j, ldr)] = -v;
      tr[jj][ii] = -v;
      ta[jj][ii] = a[idx(i, j, lda)];
    }
  }
  __syncthreads();
#pragma unroll 4
  for (int e = t; e < 64 * 64; e += 256) {
    const int jj = e % 64, ii = e / 64, i = i0 + ii, j = j0 + jj;  // lower element (j, i) = upper (i, j)
    if (i >= n || j >= n || i >= j) continue;
    r[idx(j, i, ldr)] = tr[jj][ii];
    a[idx(j, i, lda)] = ta[jj][ii];
  }
}

__global__ void oa_update_kernel(int n, const double* __restrict__ r, int ldr, const double* __restrict__ a, int lda,
                                 const double* __restrict__ lam, double* __restrict__ e, int lde) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= n) return;
  const double rij = r[idx(i, j, ldr)];
  const double gap = lam[j] - lam[i];
  double v = (a[idx(i, j, lda)] + lam[j] * rij) / gap;
  if (i == j || !(fabs(v) <= 0.25)) v = 0.5 * rij;
  e[idx(i, j, lde)] = v;
}

// 1 + E_ii from column i of E and R (one workgroup per column; R symmetric).
__global__ __launch_bounds__(256) void oa_diag_kernel(int n, const double* __restrict__ r, int ldr,
                                                      double* __restrict__ e, int lde) {
  __shared__ double red[16];
  const int i = blockIdx.x, t = threadIdx.x;
  const double* rc = r + (int64_t)i * ldr;
  const double* ec = e + (int64_t)i * lde;
  double s1 = 0, s2 = 0;
  for (int k = t; k < n; k += blockDim.x) {
    if (k == i) continue;
    const double ek = ec[k];
    s1 += ek * ek;
    s2 += rc[k] * ek;
  }
  s1 = block_sum(s1, red);
  s2 = block_sum(s2, red);
  if (t == 0) e[idx(i, i, lde)] = 1.0 + 0.5 * (rc[i] - s1 + 2.0 * s2);  // I + E: the update is one product C (I + E)
}

// rank of lambda_i in ascending order (ties by index): RANK_G lanes per i, each counting every RANK_G-th j
constexpr int RANK_G = 16;
__global__ __launch_bounds__(256) void rank_kernel(int n, const double* __restrict__ lam, int* __restrict__ rank) {
  const int g = (blockIdx.x * blockDim.x + threadIdx.x) / RANK_G, gl = threadIdx.x % RANK_G;
  const int i = min(g, n - 1);  // lanes past n count for a real i and store nothing
  const double li = lam[i];
  int r = 0;
  for (int j = gl; j < n; j += RANK_G) {
    const double lj = lam[j];
    r += (lj < li) || (lj == li && j < i);
  }
#pragma unroll
  for (int m = RANK_G / 2; m > 0; m >>= 1) r += __shfl_xor(r, m, RANK_G);
  if (g < n && gl == 0) rank[g] = r;
}

// dst(:, rank[j]) = src(:, j), w[rank[j]] = lam[j]
template <typename T>
__global__ void scatter_columns_kernel(int n, int m, const int* __restrict__ rank, const T* __restrict__ src, int lds,
                                       T* __restrict__ dst, int ldd, const double* __restrict__ lam,
                                       double* __restrict__ w) {
  const int j = blockIdx.y;
  const int r = rank[j];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    dst[idx(i, r, ldd)] = src[idx(i, j, lds)];
  if (blockIdx.x == 0 && threadIdx.x == 0 && lam != nullptr) w[r] = lam[j];
  (void)m;
}

}  // namespace

size_t refine_workspace(int n) { return 8 * (size_t)n * n + 2 * (size_t)n + 16; }

// One or more Ogita-Aishima steps on C (n x n, ldc), then eigenvalues w (ascending) and C's columns sorted to match.
void refine_oa(hipStream_t s, int n, const double* h, int ldh, const double* sm, int lds, double* c, int ldc,
               double* w, int iterations, double* work, int* iwork) {
  // no step: the eigenpairs stay as the FP32 solve left them (already in ascending order)
  if (iterations <= 0) return;
  const size_t nn = (size_t)n * n;
  double* sc = work;
  double* hc = sc + nn;
  double* r = hc + nn;
  double* a = r + nn;
  double* e = a + nn;
  double* cnew = e + nn;
  double* shm = cnew + nn;  // [S | H], each n x n (ld n): the paired products run as one batched launch
  double* lam = shm + 2 * nn;
  const dim3 eg(cdiv(n, 256), n);
  const dim3 tg(cdiv(n, 64), cdiv(n, 64));
  copy_matrix<double, double>(s, n, n, sm, lds, shm, n);
  copy_matrix<double, double>(s, n, n, h, ldh, shm + nn, n);
  // the current C alternates between c (ldc) and cnew (ld n)
  double* cur = c;
  int ldcur = ldc;
  for (int it = 0; it < iterations; ++it) {
    gemm<double>(s, 'N', 'N', n, n, n, 1.0, shm, n, cur, ldcur, 0.0, sc, n, 2, nn, 0, nn);  // S C, H C
    gemm_upper_tn<double>(s, n, n, 1.0, cur, ldcur, sc, n, 0.0, r, n, 2, 0, nn, nn);  // C^T S C, C^T H C (upper)
    hipLaunchKernelGGL(oa_prepare_kernel, tg, dim3(256), 0, s, n, r, n, a, n, lam);
    XTB_LAUNCH_CHECK();
    hipLaunchKernelGGL(oa_update_kernel, eg, dim3(256), 0, s, n, r, n, a, n, lam, e, n);
    XTB_LAUNCH_CHECK();
    hipLaunchKernelGGL(oa_diag_kernel, dim3(n), dim3(256), 0, s, n, r, n, e, n);
    XTB_LAUNCH_CHECK();
    double* nxt = (cur == c) ? cnew : c;
    const int ldnxt = (cur == c) ? n : ldc;
    gemm<double>(s, 'N', 'N', n, n, n, 1.0, cur, ldcur, e, n, 0.0, nxt, ldnxt);  // C (I + E)
    cur = nxt;
    ldcur = ldnxt;
  }
  // sort: columns of the current C scattered into the other buffer in ascending order of lambda, ending in c
  hipLaunchKernelGGL(rank_kernel, dim3(cdiv(n * RANK_G, 256)), dim3(256), 0, s, n, lam, iwork);
  XTB_LAUNCH_CHECK();
  if (cur == cnew) {
    hipLaunchKernelGGL(scatter_columns_kernel<double>, dim3(cdiv(n, 256), n), dim3(256), 0, s, n, n, iwork, cnew, n,
                       c, ldc, lam, w);
    XTB_LAUNCH_CHECK();
  } else {
    hipLaunchKernelGGL(scatter_columns_kernel<double>, dim3(cdiv(n, 256), n), dim3(256), 0, s, n, n, iwork, c, ldc,
                       cnew, n, lam, w);
    XTB_LAUNCH_CHECK();
    copy_matrix<double, double>(s, n, n, cnew, n, c, ldc);
  }
}

}  // namespace xtbeig
