// The MFMA GEMM tile shared by every matrix-product kernel (gemm.hip, and the divide-and-conquer merge products
// whose sizes are only known on the device). A 64 x TBN tile of C per workgroup (TBN 64 or 128), its wavefronts
// in a 2 x (threads / 128) grid of 16 x 16 MFMA blocks; K staged through two LDS buffers 32 at a time, with the next
// two slabs prefetched into registers (a slab is 2048 FP64 MFMA cycles per wavefront, less than the latency of its
// loads).
// MFMA operand layout (16x16x4): lane l supplies A(l % 16, l / 16) and B(l / 16, l % 16); the accumulator layout
// differs between FP32 and FP64 (Mfma<T>::row).
#pragma once
#include "common.hpp"

namespace xtbeig {
namespace gemm_detail {

constexpr int BM = 64, BN = 64, BK = 32, PAD = 4;
// LDS slab buffers per operand: slab s is staged in buffer s % 2 while the other is read, one barrier per slab.
constexpr int NBUF = 2;
// Slabs are stored k-contiguous (row i of op(A), column j of op(B), BK + KPAD elements each) and read 16 bytes at a
// time, one read feeding 16 / sizeof(T) consecutive MFMAs (a [k][i] layout read one element per MFMA was slower).

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

// LDS geometry of one tile: element (r, k) of op(A)'s slab buffer b and (c, k) of op(B)'s.
template <typename T, int TBN>
struct TileLds {
  static constexpr int P = 16 / sizeof(T);                        // elements per 16-byte read
  static constexpr int KPAD = P;                                  // keeps rows 16-byte aligned
  static constexpr int KS = BK + KPAD;                            // k-contiguous row length
  static constexpr int A_ELEMS = BM * KS;
  static constexpr int B_ELEMS = TBN * KS;
  static constexpr int ELEMS = NBUF * (A_ELEMS + B_ELEMS);
  static __device__ inline int a_off(int buf, int r, int k) {
    return buf * A_ELEMS + r * KS + k;
  }
  static __device__ inline int b_off(int buf, int c, int k) {
    return NBUF * A_ELEMS + buf * B_ELEMS + c * KS + k;
  }
};
// LDS elements a caller declares for a tile of width TBN: __shared__ T lds[tile_lds_elems<T, TBN>()].
template <typename T, int TBN = 64>
constexpr int tile_lds_elems() { return TileLds<T, TBN>::ELEMS; }

// Loads of one K slab by NT threads: op(A) (64 x BK) and op(B) (BK x TBN), each numbered along its operand's
// contiguous dimension in memory first (element f = t + NT e of thread t), so that consecutive lanes read
// consecutive addresses. When NT is a multiple of that dimension's length every element of a thread shares one base
// pointer and uniform strides (few addresses stay live across the two slabs in flight).
template <typename T, bool TA, bool TB, int NT = 256, int TBN = 64>
struct Slab {
  using L = TileLds<T, TBN>;
  static constexpr int EA = BK * BM / NT;   // elements of op(A) per thread
  static constexpr int EB = BK * TBN / NT;  // elements of op(B) per thread
  static_assert(EA * NT == BK * BM && EB * NT == BK * TBN, "slab does not divide over the threads");
  static constexpr int CA = TA ? BK : BM;   // contiguous length of A's slab
  static constexpr int CB = TB ? TBN : BK;  // contiguous length of B's slab
  T a[EA], b[EB];
  // (row of op(A) or column of op(B), k) of element e of thread t
  static __device__ inline void a_at(int t, int e, int& i, int& k) {
    const int f = t + NT * e;
    if (!TA) { i = f % CA; k = f / CA; } else { k = f % CA; i = f / CA; }
  }
  static __device__ inline void b_at(int t, int e, int& j, int& k) {
    const int f = t + NT * e;
    if (!TB) { k = f % CB; j = f / CB; } else { j = f % CB; k = f / CB; }
  }
  __device__ inline void load(const T* __restrict__ A, int lda, const T* __restrict__ B, int ldb, int M, int N,
                              int K, int m0, int n0, int k0, int t) {
#pragma unroll
    for (int e = 0; e < EA; ++e) {
      int i, k;
      a_at(t, e, i, k);
      const bool ok = m0 + i < M && k0 + k < K;
      const int64_t off =
          TA ? (int64_t)(k0 + k) + (int64_t)(m0 + i) * lda : (int64_t)(m0 + i) + (int64_t)(k0 + k) * lda;
      a[e] = ok ? A[off] : T(0);
    }
#pragma unroll
    for (int e = 0; e < EB; ++e) {
      int j, k;
      b_at(t, e, j, k);
      const bool ok = n0 + j < N && k0 + k < K;
      const int64_t off =
          TB ? (int64_t)(n0 + j) + (int64_t)(k0 + k) * ldb : (int64_t)(k0 + k) + (int64_t)(n0 + j) * ldb;
      b[e] = ok ? B[off] : T(0);
    }
  }
  __device__ inline void store(T* lds, int buf, int t) const {
#pragma unroll
    for (int e = 0; e < EA; ++e) {
      int i, k;
      a_at(t, e, i, k);
      lds[L::a_off(buf, i, k)] = a[e];
    }
#pragma unroll
    for (int e = 0; e < EB; ++e) {
      int j, k;
      b_at(t, e, j, k);
      lds[L::b_off(buf, j, k)] = b[e];
    }
  }
};

// One 64 x TBN tile of C = alpha op(A) op(B) + beta C at (m0, n0), by the calling workgroup of NT threads: the
// wavefronts in a 2 x (NT / 128) grid, each 32 rows by TBN / (NT / 128) columns (256 threads, TBN 64: four of
// 32 x 32; 512 threads: eight of 32 x 16 (TBN 64) or 32 x 32 (TBN 128), two per SIMD). lds: a
// tile_lds_elems<T, TBN>() array. The k order inside a slab is permuted (MFMA s of 16-byte read group q takes
// k = 4 P q + P g + s from lane group g, the same k for op(A) and op(B)); every element of C sums the same products in
// the same order whatever the tile width.
template <typename T, bool TA, bool TB, int NT = 256, int TBN = 64>
__device__ __forceinline__ void gemm_tile(int M, int N, int K, T alpha, const T* __restrict__ A, int lda,
                                          const T* __restrict__ B, int ldb, T beta, T* __restrict__ C, int ldc,
                                          int m0, int n0, T* lds) {
  using acc_t = typename Mfma<T>::acc_t;
  using L = TileLds<T, TBN>;
  constexpr int WN = NT / 128;        // wavefront columns
  constexpr int WC = TBN / WN;        // columns per wavefront
  constexpr int BJ = WC / 16;         // 16-column MFMA blocks per wavefront
  static_assert(BJ * 16 * WN == TBN, "tile width does not split into 16-column blocks per wavefront");
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int wm = wid & 1, wn = wid >> 1;
  const int g = lane >> 4, c = lane & 15;
  acc_t acc[2][BJ];
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < BJ; ++bj) acc[bi][bj] = acc_t{0, 0, 0, 0};

  // one slab from LDS buffer buf; mid() runs after the first group of MFMAs has been issued
  auto compute = [&](int buf, auto&& mid) {
    constexpr int P = L::P;
    typedef T vec_t __attribute__((ext_vector_type(P)));
#pragma unroll
    for (int q = 0; q < BK / (4 * P); ++q) {
      const int k = 4 * P * q + P * g;
      vec_t av[2], bv[BJ];
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
        av[bi] = *reinterpret_cast<const vec_t*>(lds + L::a_off(buf, wm * 32 + bi * 16 + c, k));
#pragma unroll
      for (int bj = 0; bj < BJ; ++bj)
        bv[bj] = *reinterpret_cast<const vec_t*>(lds + L::b_off(buf, wn * WC + bj * 16 + c, k));
#pragma unroll
      for (int s2 = 0; s2 < P; ++s2)
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
          for (int bj = 0; bj < BJ; ++bj) acc[bi][bj] = Mfma<T>::op(av[bi][s2], bv[bj][s2], acc[bi][bj]);
      if (q == 0) mid();
    }
  };
  // Two register slabs and two LDS buffers. Slab s + 1 (in registers since the previous slab) goes into the other
  // buffer while slab s is multiplied, after the first MFMA group, so that the stores overlap the matrix cores
  // instead of forming a phase of their own; the loads of slab s + 2 are issued before. Buffer (s + 1) % 2 was last
  // read by slab s - 1's compute, finished by every wavefront before the barrier that ended it; the barrier after
  // slab s orders the stores before slab s + 1's reads. The barrier on entry covers an earlier tile of the same
  // workgroup still reading the buffers.
  Slab<T, TA, TB, NT, TBN> s0, s1;
  s0.load(A, lda, B, ldb, M, N, K, m0, n0, 0, t);
  if (BK < K) s1.load(A, lda, B, ldb, M, N, K, m0, n0, BK, t);
  __syncthreads();
  s0.store(lds, 0, t);
  __syncthreads();
  for (int k0 = 0; k0 < K; k0 += 2 * BK) {
    if (k0 + 2 * BK < K) s0.load(A, lda, B, ldb, M, N, K, m0, n0, k0 + 2 * BK, t);
    compute(0, [&] {
      if (k0 + BK < K) s1.store(lds, 1, t);
    });
    __syncthreads();
    if (k0 + BK >= K) break;
    if (k0 + 3 * BK < K) s1.load(A, lda, B, ldb, M, N, K, m0, n0, k0 + 3 * BK, t);
    compute(1, [&] {
      if (k0 + 2 * BK < K) s0.store(lds, 0, t);
    });
    __syncthreads();
  }
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < BJ; ++bj) {
      const int col = n0 + wn * WC + bj * 16 + c;
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
