// The MFMA GEMM tile shared by every matrix-product kernel (gemm.hip, and the divide-and-conquer merge products
// whose sizes are only known on the device). 64 x 64 tile of C per workgroup of 256 threads: 4 wavefronts in a 2 x 2
// grid, 32 x 32 each = 2 x 2 MFMA blocks of 16 x 16; K staged through LDS 32 at a time, with the next two slabs
// prefetched into registers (a slab is 2048 FP64 MFMA cycles per wavefront, less than the latency of its loads). MFMA operand layout (16x16x4): lane l supplies A(l % 16, l / 16) and B(l / 16, l % 16); the
// accumulator layout differs between FP32 and FP64 (Mfma<T>::row).
#pragma once
#include "common.hpp"

namespace xtbeig {
namespace gemm_detail {

constexpr int BM = 64, BN = 64, BK = 32, PAD = 4;

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

// Loads of one K slab by NT threads: BK * 64 / NT elements of op(A) (64 x BK) and as many of op(B) (BK x 64) per
// thread, along the contiguous dimension of each operand.
template <typename T, bool TA, bool TB, int NT = 256>
struct Slab {
  static constexpr int E = BK * 64 / NT;  // elements of each operand per thread
  static constexpr int RS = NT / 64;      // K step between a thread's elements along a column of 64
  static constexpr int CS = NT / BK;      // column step between a thread's elements along K-contiguous columns
  T a[E], b[E];
  // One base pointer per thread and operand, then uniform strides between its elements (so that few addresses stay
  // live across the two slabs in flight).
  __device__ inline void load(const T* __restrict__ A, int lda, const T* __restrict__ B, int ldb, int M, int N,
                              int K, int m0, int n0, int k0, int t) {
    if (!TA) {
      const int i = t & 63, kq = t >> 6;  // k = kq + RS r
      const bool oki = m0 + i < M;
      const T* pa = A + (m0 + i) + (int64_t)(k0 + kq) * lda;
#pragma unroll
      for (int r = 0; r < E; ++r) a[r] = (oki && k0 + kq + RS * r < K) ? pa[(int64_t)(RS * r) * lda] : T(0);
    } else {
      const int k = t % BK, iq = t / BK;  // i = iq + CS r
      const bool okk = k0 + k < K;
      const T* pa = A + (k0 + k) + (int64_t)(m0 + iq) * lda;
#pragma unroll
      for (int r = 0; r < E; ++r) a[r] = (okk && m0 + iq + CS * r < M) ? pa[(int64_t)(CS * r) * lda] : T(0);
    }
    if (!TB) {
      const int k = t % BK, jq = t / BK;  // j = jq + CS r
      const bool okk = k0 + k < K;
      const T* pb = B + (k0 + k) + (int64_t)(n0 + jq) * ldb;
#pragma unroll
      for (int r = 0; r < E; ++r) b[r] = (okk && n0 + jq + CS * r < N) ? pb[(int64_t)(CS * r) * ldb] : T(0);
    } else {
      const int j = t & 63, kq = t >> 6;  // k = kq + RS r
      const bool okj = n0 + j < N;
      const T* pb = B + (n0 + j) + (int64_t)(k0 + kq) * ldb;
#pragma unroll
      for (int r = 0; r < E; ++r) b[r] = (okj && k0 + kq + RS * r < K) ? pb[(int64_t)(RS * r) * ldb] : T(0);
    }
  }
  __device__ inline void store(T (*As)[BM + PAD], T (*Bs)[BN + PAD], int t) const {
#pragma unroll
    for (int r = 0; r < E; ++r) {
      int i, k;
      if (!TA) { i = t & 63; k = (t >> 6) + RS * r; } else { k = t % BK; i = t / BK + CS * r; }
      As[k][i] = a[r];
    }
#pragma unroll
    for (int r = 0; r < E; ++r) {
      int j, k;
      if (!TB) { k = t % BK; j = t / BK + CS * r; } else { j = t & 63; k = (t >> 6) + RS * r; }
      Bs[k][j] = b[r];
    }
  }
};

// One 64 x 64 tile of C = alpha op(A) op(B) + beta C at (m0, n0), by the calling workgroup of NT threads (256: four
// wavefronts of 32 x 32; 512: eight of 32 x 16, two per SIMD).
template <typename T, bool TA, bool TB, int NT = 256>
__device__ __forceinline__ void gemm_tile(int M, int N, int K, T alpha, const T* __restrict__ A, int lda,
                                          const T* __restrict__ B, int ldb, T beta, T* __restrict__ C, int ldc,
                                          int m0, int n0, T (*As)[BM + PAD], T (*Bs)[BN + PAD]) {
  using acc_t = typename Mfma<T>::acc_t;
  constexpr int WN = NT / 128;        // wavefront columns
  constexpr int BJ = 64 / WN / 16;    // 16-column MFMA blocks per wavefront
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int wm = wid & 1, wn = wid >> 1;
  acc_t acc[2][BJ];
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < BJ; ++bj) acc[bi][bj] = acc_t{0, 0, 0, 0};

  auto compute = [&]() {
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kr = kk + (lane >> 4), c = lane & 15;
      T av[2], bv[BJ];
#pragma unroll
      for (int bi = 0; bi < 2; ++bi) av[bi] = As[kr][wm * 32 + bi * 16 + c];
#pragma unroll
      for (int bj = 0; bj < BJ; ++bj) bv[bj] = Bs[kr][wn * (64 / WN) + bj * 16 + c];
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < BJ; ++bj) acc[bi][bj] = Mfma<T>::op(av[bi], bv[bj], acc[bi][bj]);
    }
  };
  // two slabs in flight: the loads of slab s + 2 are issued as slab s goes to LDS, so each has two slabs' compute
  // to arrive
  Slab<T, TA, TB, NT> s0, s1;
  s0.load(A, lda, B, ldb, M, N, K, m0, n0, 0, t);
  if (BK < K) s1.load(A, lda, B, ldb, M, N, K, m0, n0, BK, t);
  for (int k0 = 0; k0 < K; k0 += 2 * BK) {
    __syncthreads();
    s0.store(As, Bs, t);
    __syncthreads();
    if (k0 + 2 * BK < K) s0.load(A, lda, B, ldb, M, N, K, m0, n0, k0 + 2 * BK, t);
    compute();
    if (k0 + BK >= K) break;
    __syncthreads();
    s1.store(As, Bs, t);
    __syncthreads();
    if (k0 + 3 * BK < K) s1.load(A, lda, B, ldb, M, N, K, m0, n0, k0 + 3 * BK, t);
    compute();
  }
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < BJ; ++bj) {
      const int col = n0 + wn * (64 / WN) + bj * 16 + (lane & 15);
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
