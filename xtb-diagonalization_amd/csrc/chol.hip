// Cholesky factor of the overlap S = U^T U and its inverse X = U^{-1}, FP64 (the generalized solve keeps S's factor
// in FP64 in both precisions). Blocked right-looking potrf('U') with 64-column blocks: the diagonal block is factored
// and inverted by one workgroup (16-column panels in one wavefront's registers, MFMA block updates), the row panel is
// U_kk^{-T} A_k* as a product with the inverted block,
// and the trailing update is a GEMM on the matrix cores. X is then assembled by recursive halving,
// X12 = -X11 U12 X22 (two GEMMs per split). X is what the congruence H' = X^T H X and the back-solve C = X C' use
// (xTB's fact_solve path; scripts/kernels/congruence.c:1-8 of the reference).
#include "common.hpp"
#include "gemm_tile.hpp"
#include "kernels.hpp"

namespace xtbeig {

namespace {

constexpr int NB = 64;

// 64-bit value of lane l (l the same in every lane).
__device__ inline double readlane_d(double v, int l) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)bits, l);
  const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// 1 / sqrt(p) to a few ulps: the hardware estimate and two Newton steps.
__device__ inline double rsqrt_nr(double p) {
  double y = __builtin_amdgcn_rsq(p);
  y = y * (1.5 - 0.5 * p * y * y);
  y = y * (1.5 - 0.5 * p * y * y);
  return y;
}

// acc(row(l, r), l % 16) += sum over K = 4 ks of P(i, k) Q(k, j) for a 16 x 16 block: the FP64 MFMA on operands read
// from LDS as P(i, k) = pa[i * lp + k] and Q(k, j) = qb[k * lq + j].
__device__ inline gemm_detail::Mfma<double>::acc_t mfma_block(const double* pa, int lp, const double* qb, int lq,
                                                               int ksteps) {
  using M = gemm_detail::Mfma<double>;
  const int lane = threadIdx.x & 63, c = lane & 15, q = lane >> 4;
  M::acc_t acc = {0, 0, 0, 0};
  for (int ks = 0; ks < ksteps; ++ks) acc = M::op(pa[c * lp + 4 * ks + q], qb[(4 * ks + q) * lq + c], acc);
  return acc;
}

// Factor A(k0:k0+kb, k0:k0+kb) = L L^T (U = L^T) and invert L, by one workgroup of 256 threads, with the block in LDS
// (padded with the identity past kb) and lanes along rows:
// - four 16-column panels, right-looking: wavefront 0 factors the panel with row c0 + l in lane l's registers (the
//   pivot and the panel rows' entries reach the other lanes by readlane, so a panel needs no barrier), then the four
//   wavefronts subtract L21 L21^T from the trailing lower blocks with FP64 MFMA;
// - Y = L^{-1}: each wavefront inverts one 16 x 16 diagonal block by forward substitution (lane = column), then the
//   sibling blocks at sizes 16 and 32, Y_BA = -Y_BB (L_BA Y_AA), as MFMA block products.
// Writes U = L^T and X = U^{-1} = Y^T at (k0, k0) (zero below the diagonal); info = k0 + k + 1 at a non-positive
// pivot k (then nothing is written). sm: DIAG_LDS doubles of LDS.
constexpr int DIAG_LD = NB + 1;
constexpr int DIAG_LDS = 2 * NB * DIAG_LD + 32 * 33;
__device__ __forceinline__ void diag_block(int n, int k0, const double* __restrict__ a, int lda,
                                           double* __restrict__ u, int ldu, double* __restrict__ x, int ldx, int* info,
                                           double* sm) {
  constexpr int LD = DIAG_LD;
  double(*Am)[LD] = reinterpret_cast<double(*)[LD]>(sm);
  double(*Ym)[LD] = reinterpret_cast<double(*)[LD]>(sm + NB * LD);
  double(*Tm)[33] = reinterpret_cast<double(*)[33]>(sm + 2 * NB * LD);
  __shared__ int bad_s;
  if (*info != 0) return;
  const int kb = min(NB, n - k0), t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int lc = lane & 15, lq = lane >> 4;
  for (int e = t; e < NB * NB; e += 256) {
    const int i = e % NB, j = e / NB;
    Am[i][j] = (i < kb && j < kb) ? a[idx(k0 + i, k0 + j, lda)] : (i == j ? 1.0 : 0.0);
    Ym[i][j] = 0.0;  // Y's upper blocks are read (as zeros) by the size-32 merge
  }
  if (t == 0) bad_s = 0;
  __syncthreads();
  for (int p = 0; p < NB / 16; ++p) {
    const int c0 = 16 * p;
    if (wid == 0) {
      const int i = c0 + lane;
      const bool live = i < NB;
      double xr[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) xr[q] = live ? Am[i][c0 + q] : 0.0;
      int bad = 0;
      // no early exit (it would keep the loop rolled and index xr at run time): after a non-positive pivot the steps
      // go on with pivot 1, and nothing of this panel is stored
#pragma unroll
      for (int kk = 0; kk < 16; ++kk) {
        double piv = readlane_d(xr[kk], kk);
        if (!(piv > 0.0)) {
          if (bad == 0) bad = k0 + c0 + kk + 1;
          piv = 1.0;
        }
        const double y = rsqrt_nr(piv);
        xr[kk] = (lane == kk) ? piv * y : (lane > kk ? xr[kk] * y : xr[kk]);
        // lanes <= kk change only entries above the diagonal, which nothing reads
#pragma unroll
        for (int jj = kk + 1; jj < 16; ++jj) xr[jj] -= xr[kk] * readlane_d(xr[kk], jj);
      }
      if (bad) {
        if (lane == 0) bad_s = bad;
      } else if (live) {
#pragma unroll
        for (int q = 0; q < 16; ++q) Am[i][c0 + q] = xr[q];
      }
    }
    __syncthreads();
    if (*(volatile int*)&bad_s) {
      if (t == 0) atomicCAS(info, 0, bad_s);
      return;
    }
    // A22 -= L21 L21^T over the lower 16 x 16 blocks (I >= J) of rows and columns c0 + 16 ..
    const int nb2 = NB / 16 - 1 - p;
    for (int blk = wid; blk < nb2 * (nb2 + 1) / 2; blk += 4) {
      int I = 0;
      while ((I + 1) * (I + 2) / 2 <= blk) ++I;
      const int J = blk - I * (I + 1) / 2;
      const int rI = c0 + 16 + 16 * I, rJ = c0 + 16 + 16 * J;
      // P(i, k) = L(rI + i, c0 + k), Q(k, j) = L(rJ + j, c0 + k)
      using M = gemm_detail::Mfma<double>;
      M::acc_t acc = {0, 0, 0, 0};
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) acc = M::op(Am[rI + lc][c0 + 4 * ks + lq], Am[rJ + lc][c0 + 4 * ks + lq], acc);
#pragma unroll
      for (int r = 0; r < 4; ++r) Am[rI + M::row(lane, r)][rJ + lc] -= acc[r];
    }
    __syncthreads();
  }
  // diagonal blocks of Y = L^{-1}: wavefront w, lane column c, Y(i, c) = (delta_ic - sum_{c <= l < i} L(i, l) Y(l, c))
  // / L(i, i) for i >= c
  {
    const int d0 = 16 * wid;
    double y[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      double s0 = (i == lc) ? 1.0 : 0.0, s1 = 0.0;
#pragma unroll
      for (int l = 0; l + 1 < i; l += 2) {
        s0 -= Am[d0 + i][d0 + l] * y[l];
        s1 -= Am[d0 + i][d0 + l + 1] * y[l + 1];
      }
      if (i % 2 == 1) s0 -= Am[d0 + i][d0 + i - 1] * y[i - 1];
      y[i] = (i >= lc) ? (s0 + s1) / Am[d0 + i][d0 + i] : 0.0;
    }
    if (lane < 16) {
#pragma unroll
      for (int i = 0; i < 16; ++i) Ym[d0 + i][d0 + lc] = y[i];
    }
  }
  __syncthreads();
  using M = gemm_detail::Mfma<double>;
  // size 16: pairs (A, B) = (0, 1) and (2, 3) by wavefronts 0 and 1: T = L_BA Y_AA, then Y_BA = -Y_BB T
  if (wid < 2) {
    const int A0 = 32 * wid, B0 = A0 + 16;
    M::acc_t acc = mfma_block(&Am[B0][A0], LD, &Ym[A0][A0], LD, 4);
#pragma unroll
    for (int r = 0; r < 4; ++r) Tm[16 * wid + M::row(lane, r)][lc] = acc[r];
    __builtin_amdgcn_wave_barrier();
    acc = mfma_block(&Ym[B0][B0], LD, &Tm[16 * wid][0], 33, 4);
#pragma unroll
    for (int r = 0; r < 4; ++r) Ym[B0 + M::row(lane, r)][A0 + lc] = -acc[r];
  }
  __syncthreads();
  // size 32: A = [0, 32), B = [32, 64); wavefront w takes the output block (w / 2, w % 2) of T = L_BA Y_AA, then of
  // Y_BA = -Y_BB T
  {
    const int bi = wid >> 1, bj = wid & 1;
    M::acc_t acc = mfma_block(&Am[32 + 16 * bi][0], LD, &Ym[0][16 * bj], LD, 8);
#pragma unroll
    for (int r = 0; r < 4; ++r) Tm[16 * bi + M::row(lane, r)][16 * bj + lc] = acc[r];
    __syncthreads();
    acc = mfma_block(&Ym[32 + 16 * bi][32], LD, &Tm[0][16 * bj], 33, 8);
#pragma unroll
    for (int r = 0; r < 4; ++r) Ym[32 + 16 * bi + M::row(lane, r)][16 * bj + lc] = -acc[r];
  }
  __syncthreads();
  for (int e = t; e < NB * NB; e += 256) {
    const int i = e % NB, j = e / NB;
    if (i < kb && j < kb) {
      u[idx(k0 + i, k0 + j, ldu)] = (i <= j) ? Am[j][i] : 0.0;
      x[idx(k0 + i, k0 + j, ldx)] = (i <= j) ? Ym[j][i] : 0.0;
    }
  }
}

__global__ __launch_bounds__(256) void potrf_diag_kernel(int n, int k0, double* __restrict__ a, int lda,
                                                         double* __restrict__ u, int ldu, double* __restrict__ x,
                                                         int ldx, int* info, Zb zb) {
  __shared__ double sm[DIAG_LDS];
  diag_block(n, k0, zb.at(a), lda, zb.at(u), ldu, zb.at(x), ldx, zb.at(info), sm);
}

// One step's trailing update fused with the next step's diagonal block: A(rest, rest) -= P^T P on the tiles on and
// above the diagonal (P = U(k0 block, rest), the row panel; tile q = J (J + 1) / 2 + I as in the plain upper-tile
// product, so every tile sums the same products in the same order), and workgroup 0, after its tile (0, 0), which is
// the whole of the next diagonal block, factors and inverts that block (diag_block at k1 = k0 + kb). The next
// diagonal block is thus ready when the update is, instead of one launch later.
__global__ __launch_bounds__(256) void potrf_update_diag_kernel(int n, int k1, int kb, const double* __restrict__ p,
                                                                int ldu, double* __restrict__ a, int lda,
                                                                double* __restrict__ u, double* __restrict__ x,
                                                                int ldx, int* info, Zb zb) {
  constexpr int LDS_ELEMS = gemm_detail::tile_lds_elems<double>() > DIAG_LDS ? gemm_detail::tile_lds_elems<double>()
                                                                              : DIAG_LDS;
  __shared__ __attribute__((aligned(16))) double lds[LDS_ELEMS];
  p = zb.at(p);
  a = zb.at(a);
  u = zb.at(u);
  x = zb.at(x);
  info = zb.at(info);
  const int r = n - k1, q = blockIdx.x;
  int J = (int)((sqrtf(8.0f * (float)q + 1.0f) - 1.0f) * 0.5f);
  while (J * (J + 1) / 2 > q) --J;
  while ((J + 1) * (J + 2) / 2 <= q) ++J;
  const int I = q - J * (J + 1) / 2;
  gemm_detail::gemm_tile<double, true, false>(r, r, kb, -1.0, p, ldu, p, ldu, 1.0, a + idx(k1, k1, lda), lda,
                                              I * gemm_detail::BM, J * gemm_detail::BN, lds);
  if (q != 0) return;
  __syncthreads();  // the tile's stores of A(k1 block, k1 block) before the block's loads, and its LDS reads
  diag_block(n, k1, a, lda, u, ldu, x, ldx, info, lds);
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
                   int* info, hipStream_t s2, hipEvent_t ev, hipEvent_t ev_join) {
  if (n <= 0) return;
  set_zero(s, n, n, u, ldu);
  set_zero(s, n, n, x, ldx);
  const bool beside = s2 != nullptr && ev != nullptr && ev_join != nullptr;
  // the first diagonal block alone; every later one inside the previous step's trailing update
  hipLaunchKernelGGL(potrf_diag_kernel, dim3(1, 1, zgrid()), dim3(256), 0, s, n, 0, a, lda, u, ldu, x, ldx, info,
                     zb());
  XTB_LAUNCH_CHECK();
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int kb = std::min(NB, n - k0), r = n - k0 - kb;
    if (beside && k0 > 0) {
      // X(0:k0, block) = -(X(0:k0, 0:k0) U(0:k0, block)) X_kk on s2: U's block column came from the row panels of
      // the earlier steps, X_kk from this step's diagonal kernel, X(0:k0, 0:k0) from s2's earlier steps
      XTB_HIP(hipEventRecord(ev, s));
      XTB_HIP(hipStreamWaitEvent(s2, ev, 0));
      gemm<double>(s2, 'N', 'N', k0, kb, k0, 1.0, x, ldx, u + idx(0, k0, ldu), ldu, 0.0, tmp, k0);
      gemm<double>(s2, 'N', 'N', k0, kb, kb, -1.0, tmp, k0, x + idx(k0, k0, ldx), ldx, 0.0, x + idx(0, k0, ldx), ldx);
    }
    if (r > 0) {
      // U(k0 block, rest) = U_kk^{-T} A(k0 block, rest)
      gemm<double>(s, 'T', 'N', kb, r, kb, 1.0, x + idx(k0, k0, ldx), ldx, a + idx(k0, k0 + kb, lda), lda, 0.0,
                   u + idx(k0, k0 + kb, ldu), ldu);
      // A(rest, rest) -= U(k0 block, rest)^T U(k0 block, rest), on the tiles the later steps read: the diagonal
      // blocks and everything above them (the diagonal block reads its whole 64 x 64 block, the row panels the part
      // right of it); then the next diagonal block
      const int nt = cdiv(r, NB);
      hipLaunchKernelGGL(potrf_update_diag_kernel, dim3(nt * (nt + 1) / 2, 1, zgrid()), dim3(256), 0, s, n, k0 + kb,
                         kb, u + idx(k0, k0 + kb, ldu), ldu, a, lda, u, x, ldx, info, zb());
      XTB_LAUNCH_CHECK();
    }
  }
  if (beside) {
    XTB_HIP(hipEventRecord(ev_join, s2));
    XTB_HIP(hipStreamWaitEvent(s, ev_join, 0));
  } else {
    inverse_blocks(s, 0, n, u, ldu, x, ldx, tmp);
  }
}

}  // namespace xtbeig
