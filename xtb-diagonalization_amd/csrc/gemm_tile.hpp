// The MFMA GEMM tile shared by every matrix-product kernel (gemm.hip, and the divide-and-conquer merge products
// whose sizes are only known on the device). 64 x 64 tile of C per workgroup of 256 threads: 4 wavefronts in a 2 x 2
// grid, 32 x 32 each = 2 x 2 MFMA blocks of 16 x 16; K staged through LDS 16 at a time with the next slab prefetched
// into registers. MFMA operand layout (16x16x4): lane l supplies A(l % 16, l / 16) and B(l / 16, l % 16); the
// accumulator layout differs between FP32 and FP64 (Mfma<T>::row).
#pragma once
#include "common.hpp"

namespace xtbeig {
namespace gemm_detail {

constexpr int BM = 64, BN = 64, BK = 16, PAD = 4;

template <typename T> struct Mfma;
// Accumulator layout: register r of lane l holds C(row(l, r), l % 16). The FP32 16x16x4 instruction places rows
// 4 (l / 16) + r; the FP64 one places rows l / 16 + 4 r (each of its four 64-bit accumulators is a 4 x 16 slab).
template <> struct Mfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  static __device__ inline acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct Mfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  static __device__ inline acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// Loads of one K slab: 4 elements of op(A) (64 x 16) and 4 of op(B) (16 x 64) per thread.
template <typename T, bool TA, bool TB>
struct Slab {
  T a[4], b[4];
  __device__ inline void load(const T* __restrict__ A, int lda, const T* __restrict__ B, int ldb, int M, int N,
                              int K, int m0, int n0, int k0, int t) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int i, k;
      if (!TA) { i = t & 63; k = (t >> 6) + 4 * r; } else { k = t & 15; i = (t >> 4) + 16 * r; }
      const int gi = m0 + i, gk = k0 + k;
      T v = T(0);
      if (gi < M && gk < K) v = TA ? A[idx(gk, gi, lda)] : A[idx(gi, gk, lda)];
      a[r] = v;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int j, k;
      if (!TB) { k = t & 15; j = (t >> 4) + 16 * r; } else { j = t & 63; k = (t >> 6) + 4 * r; }
      const int gj = n0 + j, gk = k0 + k;
      T v = T(0);
      if (gj < N && gk < K) v = TB ? B[idx(gj, gk, ldb)] : B[idx(gk, gj, ldb)];
      b[r] = v;
    }
  }
  __device__ inline void store(T (*As)[BM + PAD], T (*Bs)[BN + PAD], int t) const {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int i, k;
      if (!TA) { i = t & 63; k = (t >> 6) + 4 * r; } else { k = t & 15; i = (t >> 4) + 16 * r; }
      As[k][i] = a[r];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int j, k;
      if (!TB) { k = t & 15; j = (t >> 4) + 16 * r; } else { j = t & 63; k = (t >> 6) + 4 * r; }
      Bs[k][j] = b[r];
    }
  }
};

// One 64 x 64 tile of C = alpha op(A) op(B) + beta C at (m0, n0), by the calling workgroup of 256 threads.
template <typename T, bool TA, bool TB>
__device__ inline void gemm_tile(int M, int N, int K, T alpha, const T* __restrict__ A, int lda,
                                 const T* __restrict__ B, int ldb, T beta, T* __restrict__ C, int ldc, int m0,
                                 int n0, T (*As)[BM + PAD], T (*Bs)[BN + PAD]) {
  using acc_t = typename Mfma<T>::acc_t;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int wm = wid & 1, wn = wid >> 1;
  acc_t acc[2][2];
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = acc_t{0, 0, 0, 0};

  Slab<T, TA, TB> slab;
  slab.load(A, lda, B, ldb, M, N, K, m0, n0, 0, t);
  for (int k0 = 0; k0 < K; k0 += BK) {
    __syncthreads();
    slab.store(As, Bs, t);
    __syncthreads();
    if (k0 + BK < K) slab.load(A, lda, B, ldb, M, N, K, m0, n0, k0 + BK, t);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kr = kk + (lane >> 4), c = lane & 15;
      T av[2], bv[2];
#pragma unroll
      for (int bi = 0; bi < 2; ++bi) av[bi] = As[kr][wm * 32 + bi * 16 + c];
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) bv[bj] = Bs[kr][wn * 32 + bj * 16 + c];
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = Mfma<T>::op(av[bi], bv[bj], acc[bi][bj]);
    }
  }
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) {
      const int col = n0 + wn * 32 + bj * 16 + (lane & 15);
      if (col >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * 32 + bi * 16 + Mfma<T>::row(lane, r);
        if (row < M) {
          T* p = C + idx(row, col, ldc);
          const T v = alpha * acc[bi][bj][r];
          *p = (beta == T(0)) ? v : v + beta * *p;
        }
      }
    }
}

}  // namespace gemm_detail
}  // namespace xtbeig
