// Back-transformations of the two-stage reduction, on zt = Z^T: row i of the eigenvector matrix Z is the contiguous
// column i of zt, so the 64 lanes of a wavefront hold 64 eigenvectors.
//
// Bulge reflectors (the reference's backTransformBulges / backTransformBulgesBlocked,
// Stages/BulgeBackTransformation.swift:11-85): Z <- Q2 Z, Q2 the chase-order product of H(s, k) = I - w w^T, the
// reflector of sweep s at bulge k acting on rows rs = s + 1 + k b ..< rs + b (w = sqrt(tau) v as the chase stores
// it). The reflectors of g consecutive sweeps s0 ..< s0 + g at one k form a block on rows row0 ..< row0 + g + b - 1,
// applied last sweep first. Blocks in the order k ascending, s0 descending are a valid application order: relative
// to the reverse chase order only pairs on disjoint rows change places ((s, k) and (s', k') with s < s', k < k' start
// at least b + 1 rows apart). The host levels that sequence: a block's stage is one more than the latest stage of an
// earlier block sharing a row with it, so the blocks of a stage act on disjoint rows and need only earlier stages.
// One workgroup of BULGE_WAVES wavefronts owns 64 columns of Z (lane = column) and walks the stages with a barrier
// between them; its wavefronts take a stage's blocks round robin. A block's g + b - 1 rows sit in registers and its
// reflector entries in the wavefront's LDS slot (read as broadcasts), so a reflector costs 2b FMAs per lane and no
// shuffles.
//
// Band panels (the reference's backTransformBand, Stages/BandFormBackTransformation.swift:10-21): Z <- P_0 ... P_p Z
// with P_j = I - V_j T_j V_j^T. ppg consecutive panels merge into one block I - V T V^T, V = [V_a V_b ...] and T
// upper block triangular (T(0:c, a) = -T(0:c, 0:c) (V^T V)(0:c, a) T_aa, the merge of two compact-WY factors),
// applied as three products on the matrix cores: W = zt(:, r0:) V (split over rows), W2 = W T^T,
// zt(:, r0:) -= W2 V^T.
#include "common.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <vector>

namespace xtbeig {

namespace {

constexpr int BULGE_G = 16;     // sweeps per bulge block
constexpr int BULGE_WAVES = 8;  // wavefronts per workgroup
constexpr int GROUP_K = 64;     // reflectors per merged panel block
constexpr int SPLIT_KC = 128;   // rows per chunk of the split product W = zt V
constexpr size_t LDS_BYTES = 160 * 1024;  // LDS per workgroup on gfx950

// REV = false: Z <- Q2 Z (zt <- zt Q2^T), the sequence described above. REV = true: Z <- Q2^T Z (zt <- zt Q2), the
// same sequence run backwards: stages from the last to the first, each block's reflectors from its first sweep to
// its last.
template <typename T, int B, bool REV>
__global__ __launch_bounds__(64 * BULGE_WAVES) void bulge_apply_kernel(int n, int b, int m,
                                                                       const BulgeBlock* __restrict__ blocks,
                                                                       const int* __restrict__ stage_off,
                                                                       const int* __restrict__ order, int nstages,
                                                                       const T* __restrict__ v2, int ldv2,
                                                                       T* __restrict__ zt, int ldzt, Zb zb) {
  v2 = zb.at(v2);
  zt = zb.at(zt);
  constexpr int H = BULGE_G + B - 1;
  constexpr int WE = BULGE_G * B;  // reflector entries of one block
  // reflector j at ws[j WS ..< j WS + B]; the gap keeps the compiler from merging reads across reflectors
  constexpr int WS = B + 16 / (int)sizeof(T);
  __shared__ T wsh[BULGE_WAVES][BULGE_G * WS];
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int col = blockIdx.x * 64 + lane;  // < ldzt; columns past m are padding
  T* ws = wsh[wid];
  for (int si = 0; si < nstages; ++si) {
    const int st = REV ? nstages - 1 - si : si;
    const int beg = stage_off[st], end = stage_off[st + 1];
    for (int q = beg + wid; q < end; q += BULGE_WAVES) {
      const BulgeBlock bk = blocks[order[q]];
      const int r0 = bk.row0, cnt = bk.count, s0 = bk.s0;
      // the block's reflectors into this wavefront's LDS slot, entry r of sweep s0 + j at ws[j WS + r], zero past
      // the reflector's length and for absent sweeps (j >= cnt)
#pragma unroll
      for (int u = 0; u < WE / 64; ++u) {
        const int e = lane + 64 * u, j = e / B, r = e % B;
        const int rs = r0 + j;
        T x = 0;
        if (j < cnt && r < min(b, n - rs)) x = v2[(int64_t)(rs + r) + (int64_t)ldv2 * (s0 + j)];
        ws[j * WS + r] = x;
      }
      __builtin_amdgcn_wave_barrier();
      // the window: rows r0 ..< r0 + H (the stage's footprint of this block; rows past n are zero padding)
      T z[H];
#pragma unroll
      for (int i = 0; i < H; ++i) z[i] = zt[(int64_t)(r0 + i) * ldzt + col];
#pragma unroll
      for (int jj = 0; jj < BULGE_G; ++jj) {
        const int j = REV ? jj : BULGE_G - 1 - jj;
        // the reflector in registers for B <= 32; wider ones are read from LDS twice (a 64-entry copy beside the
        // 79-row window would not fit the register file)
        constexpr int BR = B <= 32 ? B : 1;
        T wr[BR];
        if constexpr (B <= 32) {
#pragma unroll
          for (int r = 0; r < B; ++r) wr[r] = ws[j * WS + r];
        }
        auto w_at = [&](int r) { if constexpr (B <= 32) return wr[r]; else return ws[j * WS + r]; };
        T d0 = 0, d1 = 0, d2 = 0, d3 = 0;
#pragma unroll
        for (int r = 0; r < B; r += 4) {
          d0 += w_at(r) * z[j + r];
          d1 += w_at(r + 1) * z[j + r + 1];
          d2 += w_at(r + 2) * z[j + r + 2];
          d3 += w_at(r + 3) * z[j + r + 3];
        }
        const T dot = (d0 + d1) + (d2 + d3);
#pragma unroll
        for (int r = 0; r < B; ++r) z[j + r] -= w_at(r) * dot;
        // compiler-only fence: the next reflector's LDS reads are not hoisted above this one (register pressure)
        __asm__ volatile("" ::: "memory");
      }
#pragma unroll
      for (int i = 0; i < H; ++i) zt[(int64_t)(r0 + i) * ldzt + col] = z[i];
      __builtin_amdgcn_wave_barrier();  // ws is rewritten by the next block
    }
    __syncthreads();
  }
}

// The same sequence with the strip zt(0 ..< rows, C columns) resident in LDS, for sizes where it fits: one workgroup
// per C columns (lane = column within a group of C lanes, 64 / C groups per wavefront), so a 1200-column zt spreads
// over 150 workgroups instead of 19. Each lane group is a slot that takes one block per round of a stage; its window
// moves between LDS and registers, so a stage costs the block's reflector chain and one barrier, not a round trip to
// memory. The reflector entries of the slot's next P blocks are in flight in registers while the current block
// computes (their descriptors one block further ahead), and staged into the slot's LDS area when that block starts:
// with one or two wavefronts per SIMD a block's compute is shorter than the latency of its loads.
template <typename T, int B, bool REV, int C, int W, int P>
__global__ __launch_bounds__(64 * W) void bulge_strip_kernel(int n, int b, int rows, int rs, int nstages,
                                                             const BulgeBlock* __restrict__ sorted,
                                                             const int* __restrict__ stage_off,
                                                             const T* __restrict__ v2, int ldv2,
                                                             T* __restrict__ zt, int ldzt, Zb zb) {
  v2 = zb.at(v2);
  zt = zb.at(zt);
  constexpr int G = BULGE_G, H = G + B - 1, WE = G * B;
  constexpr int GPW = 64 / C, SLOTS = W * GPW;
  constexpr int SS = WE + 16 / (int)sizeof(T);  // slot stride: the slots of one wavefront start 4 banks apart
  constexpr int PE = WE / C;                     // reflector entries a lane stages per block
  extern __shared__ __align__(16) char smem_raw[];
  T* slots = reinterpret_cast<T*>(smem_raw);
  T* strip = slots + SLOTS * SS;  // column c of the strip at strip[c rs + row]
  int* soff = reinterpret_cast<int*>(strip + C * rs);
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int grp = lane / C, cl = lane % C;
  const int slot = wid * GPW + grp;
  const int col0 = blockIdx.x * C;
  T* ws = slots + slot * SS;
  T* zs = strip + cl * rs;
  for (int e = t; e < rows * C; e += 64 * W) {
    const int r = e / C, c = e % C;
    strip[c * rs + r] = zt[(int64_t)r * ldzt + col0 + c];
  }
  for (int e = t; e <= nstages; e += 64 * W) soff[e] = stage_off[e];
  __syncthreads();

  // Tasks: (stage si in processing order, round rd); the slot's block of a task is q = beg + rd SLOTS + slot when
  // q < end. si == nstages marks the end of the sequence.
  auto advance = [&](int& si, int& rd) {
    if (si >= nstages) return;
    const int st = REV ? nstages - 1 - si : si;
    if ((rd + 1) * SLOTS < soff[st + 1] - soff[st]) {
      ++rd;
    } else {
      ++si;
      rd = 0;
    }
  };
  auto block_of = [&](int si, int rd) {
    BulgeBlock bk{0, 0, 0, 0};
    if (si < nstages) {
      const int st = REV ? nstages - 1 - si : si;
      const int q = soff[st] + rd * SLOTS + slot;
      if (q < soff[st + 1]) bk = sorted[q];
    }
    return bk;
  };
  auto fetch = [&](const BulgeBlock& bk, T (&pre)[PE]) {
#pragma unroll
    for (int u = 0; u < PE; ++u) {
      const int e = cl + C * u, j = e / B, r = e % B;
      const int rsw = bk.row0 + j;
      T x = 0;
      if (j < bk.count && r < min(b, n - rsw)) x = v2[(int64_t)(rsw + r) + (int64_t)ldv2 * (bk.s0 + j)];
      pre[u] = x;
    }
  };

  // the task queue: qsi / qrd / qb[k] the k-th task from the current one (k = 0 ..< P + 2); pre[k] the reflector
  // entries of task k (k < P)
  constexpr int Q = P + 2;
  int qsi[Q], qrd[Q];
  BulgeBlock qb[Q];
  qsi[0] = 0;
  qrd[0] = 0;
#pragma unroll
  for (int k = 1; k < Q; ++k) {
    qsi[k] = qsi[k - 1];
    qrd[k] = qrd[k - 1];
    advance(qsi[k], qrd[k]);
  }
#pragma unroll
  for (int k = 0; k < Q; ++k) qb[k] = block_of(qsi[k], qrd[k]);
  T pre[P][PE];
#pragma unroll
  for (int k = 0; k < P; ++k) fetch(qb[k], pre[k]);
  while (qsi[0] < nstages) {
    // stage the current block's reflectors (zero for absent sweeps and idle slots)
#pragma unroll
    for (int u = 0; u < PE; ++u) ws[cl + C * u] = pre[0][u];
    __builtin_amdgcn_wave_barrier();
    const BulgeBlock bk = qb[0];
    // the queue moves up one task; the entries of task P (its descriptor loaded a task ago) start loading
#pragma unroll
    for (int k = 0; k + 1 < P; ++k)
#pragma unroll
      for (int u = 0; u < PE; ++u) pre[k][u] = pre[k + 1][u];
    fetch(qb[P], pre[P - 1]);
    if (bk.count > 0) {
      const int r0 = bk.row0;
      T z[H];
#pragma unroll
      for (int i = 0; i < H; ++i) z[i] = zs[r0 + i];
      T wa[B];
      {
        const int j = REV ? 0 : G - 1;
#pragma unroll
        for (int r = 0; r < B; ++r) wa[r] = ws[j * B + r];
      }
#pragma unroll
      for (int jj = 0; jj < G; ++jj) {
        const int j = REV ? jj : G - 1 - jj;
        T wb[B];
        if (jj + 1 < G) {
          const int jn = REV ? j + 1 : j - 1;
#pragma unroll
          for (int r = 0; r < B; ++r) wb[r] = ws[jn * B + r];
        }
        T d0 = 0, d1 = 0, d2 = 0, d3 = 0;
#pragma unroll
        for (int r = 0; r < B; r += 4) {
          d0 += wa[r] * z[j + r];
          d1 += wa[r + 1] * z[j + r + 1];
          d2 += wa[r + 2] * z[j + r + 2];
          d3 += wa[r + 3] * z[j + r + 3];
        }
        const T dot = (d0 + d1) + (d2 + d3);
#pragma unroll
        for (int r = 0; r < B; ++r) z[j + r] -= wa[r] * dot;
        // compiler-only fence: reflector reads stay one reflector ahead of the arithmetic
        __asm__ volatile("" ::: "memory");
        if (jj + 1 < G) {
#pragma unroll
          for (int r = 0; r < B; ++r) wa[r] = wb[r];
        }
      }
#pragma unroll
      for (int i = 0; i < H; ++i) zs[r0 + i] = z[i];
    }
    __builtin_amdgcn_wave_barrier();  // the slot's reflectors are rewritten by the next block
    const bool stage_end = qsi[1] != qsi[0];
#pragma unroll
    for (int k = 0; k + 1 < Q; ++k) {
      qsi[k] = qsi[k + 1];
      qrd[k] = qrd[k + 1];
      qb[k] = qb[k + 1];
    }
    advance(qsi[Q - 1], qrd[Q - 1]);
    qb[Q - 1] = block_of(qsi[Q - 1], qrd[Q - 1]);
    if (stage_end) __syncthreads();
  }
  for (int e = t; e < n * C; e += 64 * W) {
    const int r = e / C, c = e % C;
    zt[(int64_t)r * ldzt + col0 + c] = strip[c * rs + r];
  }
}

// Merged T of panel group blockIdx.x (panels j0 + p b, p < ppg, j <= last): the Gram V^T V over rows r0 ..< n,
// streamed in 32-row chunks (each thread a 4 x 4 tile), then the block recurrence. Output tg (64 x 64, zero outside
// the group's K columns).
template <typename T>
__global__ __launch_bounds__(256) void group_t_kernel(int n, int b, int ppg, int last, const T* __restrict__ v1,
                                                      const T* __restrict__ t1, T* __restrict__ tg, Zb zb) {
  v1 = zb.at(v1);
  t1 = zb.at(t1);
  tg = zb.at(tg);
  extern __shared__ char smem_raw[];  // three GROUP_K x (GROUP_K + 1) arrays
  using Row = T[GROUP_K + 1];
  Row* G = reinterpret_cast<Row*>(smem_raw);
  Row* Tm = G + GROUP_K;
  Row* X = Tm + GROUP_K;
  const int g = blockIdx.x, t = threadIdx.x;
  const int j0 = g * ppg * b, r0 = j0 + b;
  int q = 0;
  while (q < ppg && j0 + q * b <= last) ++q;
  const int K = (q - 1) * b + min(b, n - (j0 + (q - 1) * b + b));
  {
    T* Vs = &X[0][0];  // 32 x 64 chunk, ld GROUP_K + 1
    const int ta = (t & 15) * 4, tc = (t >> 4) * 4;
    T acc[4][4] = {};
    for (int i0 = r0; i0 < n; i0 += 32) {
      for (int e = t; e < 32 * GROUP_K; e += 256) {
        const int i = e & 31, c = e >> 5;
        Vs[i * (GROUP_K + 1) + c] = (i0 + i < n && c < K) ? v1[(int64_t)(i0 + i) + (int64_t)(j0 + c) * n] : T(0);
      }
      __syncthreads();
#pragma unroll 4
      for (int i = 0; i < 32; ++i) {
        T va[4], vc[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          va[u] = Vs[i * (GROUP_K + 1) + ta + u];
          vc[u] = Vs[i * (GROUP_K + 1) + tc + u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[u][v] += va[u] * vc[v];
      }
      __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) G[ta + u][tc + v] = acc[u][v];
  }
  for (int e = t; e < GROUP_K * GROUP_K; e += 256) {
    const int r = e & (GROUP_K - 1), c = e / GROUP_K;
    const int pr = r / b, pc = c / b;
    T val = 0;
    if (pr == pc && r < K && c < K && r <= c)
      val = t1[(int64_t)(j0 + pr * b) * b + (r - pr * b) + (int64_t)(c - pc * b) * b];
    Tm[r][c] = val;
  }
  __syncthreads();
  for (int p = 1; p < q; ++p) {
    const int c0 = p * b, kp = min(b, K - c0);
    // X(0:c0, 0:kp) = G(0:c0, c0:c0+kp) T_pp
    for (int e = t; e < c0 * kp; e += 256) {
      const int r = e % c0, c = e / c0;
      T sum = 0;
      for (int l = 0; l <= c; ++l) sum += G[r][c0 + l] * Tm[c0 + l][c0 + c];
      X[r][c] = sum;
    }
    __syncthreads();
    // T(0:c0, c0:c0+kp) = -T(0:c0, 0:c0) X
    for (int e = t; e < c0 * kp; e += 256) {
      const int r = e % c0, c = e / c0;
      T sum = 0;
      for (int l = r; l < c0; ++l) sum += Tm[r][l] * X[l][c];
      Tm[r][c0 + c] = -sum;
    }
    __syncthreads();
  }
  T* out = tg + (int64_t)g * GROUP_K * GROUP_K;
  for (int e = t; e < GROUP_K * GROUP_K; e += 256) {
    const int r = e & (GROUP_K - 1), c = e / GROUP_K;
    out[e] = Tm[r][c];
  }
}

// W2 (m x K, ld m) = (sum_z P_z) T^T for COMBINE_ROWS rows per workgroup; P_z (m x K, ld m) at P + z m K.
constexpr int COMBINE_ROWS = 32;
template <typename T>
__global__ __launch_bounds__(256) void group_combine_kernel(int m, int K, int splits, const T* __restrict__ P,
                                                            const T* __restrict__ tg, T* __restrict__ w2, Zb zb) {
  P = zb.at(P);
  tg = zb.at(tg);
  w2 = zb.at(w2);
  __shared__ T Ts[GROUP_K][GROUP_K + 1];
  __shared__ T Ws[COMBINE_ROWS][GROUP_K + 1];
  const int i0 = blockIdx.x * COMBINE_ROWS, t = threadIdx.x;
  for (int e = t; e < GROUP_K * GROUP_K; e += 256) Ts[e & (GROUP_K - 1)][e / GROUP_K] = tg[e];
  const int64_t mk = (int64_t)m * K;
  for (int e = t; e < COMBINE_ROWS * K; e += 256) {
    const int i = e % COMBINE_ROWS, l = e / COMBINE_ROWS;
    T sum = 0;
    if (i0 + i < m)
      for (int z = 0; z < splits; ++z) sum += P[z * mk + (i0 + i) + (int64_t)l * m];
    Ws[i][l] = sum;
  }
  __syncthreads();
  for (int e = t; e < COMBINE_ROWS * K; e += 256) {
    const int i = e % COMBINE_ROWS, c = e / COMBINE_ROWS;
    if (i0 + i >= m) continue;
    T s0 = 0, s1 = 0;
    int l = c;
    for (; l + 2 <= K; l += 2) {
      s0 += Ws[i][l] * Ts[c][l];
      s1 += Ws[i][l + 1] * Ts[c][l + 1];
    }
    for (; l < K; ++l) s0 += Ws[i][l] * Ts[c][l];
    w2[(i0 + i) + (int64_t)c * m] = s0 + s1;
  }
}

// The LDS-resident variant with C columns per workgroup and W wavefronts, when its LDS fits (false: not launched).
template <typename T, int B, bool REV, int C, int W, int P>
bool launch_strip(hipStream_t s, const BackPlan& p, const T* v2, T* zt, int ldzt, int m) {
  constexpr int SS = BULGE_G * B + 16 / (int)sizeof(T);
  const int rs = p.rows | 1;  // odd column stride: the C lanes of a group read distinct banks
  const size_t lds = sizeof(T) * ((size_t)W * (64 / C) * SS + (size_t)C * rs) + sizeof(int) * (p.nstages + 1);
  if (lds > LDS_BYTES) return false;
  allow_lds(bulge_strip_kernel<T, B, REV, C, W, P>, lds);
  hipLaunchKernelGGL((bulge_strip_kernel<T, B, REV, C, W, P>), dim3(cdiv(m, C), 1, zgrid()), dim3(64 * W), lds, s, p.n,
                     p.b, p.rows, rs, p.nstages, p.sorted_dev, p.stage_off_dev, v2, p.n, zt, ldzt, zb());
  XTB_LAUNCH_CHECK();
  return true;
}

template <typename T, int B, bool REV>
bool launch_strips(hipStream_t s, const BackPlan& p, const T* v2, T* zt, int ldzt, int m) {
  if constexpr (B > 16) {
    return false;
  } else {
    // A single solve runs this beside the divide and conquer on few workgroups: four wavefronts, loads one block
    // ahead. A batch fills the GPU with strips: eight wavefronts and loads two blocks ahead hide more latency there
    // (measured: 32 solves of n = 1204, FP64 505 -> 531 solves/s).
    if (current_batch().count > 1 &&
        (launch_strip<T, B, REV, 16, 8, 2>(s, p, v2, zt, ldzt, m) ||
         launch_strip<T, B, REV, 8, 8, 2>(s, p, v2, zt, ldzt, m) ||
         launch_strip<T, B, REV, 8, 4, 2>(s, p, v2, zt, ldzt, m)))
      return true;
    return launch_strip<T, B, REV, 8, 4, 1>(s, p, v2, zt, ldzt, m) ||
           launch_strip<T, B, REV, 8, 2, 1>(s, p, v2, zt, ldzt, m);
  }
}

template <typename T, int B>
void launch_bulges(hipStream_t s, const BackPlan& p, const T* v2, T* zt, int ldzt, int m, bool transposed) {
  if (transposed ? launch_strips<T, B, true>(s, p, v2, zt, ldzt, m) : launch_strips<T, B, false>(s, p, v2, zt, ldzt, m))
    return;
  if (transposed)
    hipLaunchKernelGGL((bulge_apply_kernel<T, B, true>), dim3(cdiv(m, 64), 1, zgrid()), dim3(64 * BULGE_WAVES), 0, s,
                       p.n, p.b, m, p.blocks_dev, p.stage_off_dev, p.order_dev, p.nstages, v2, p.n, zt, ldzt, zb());
  else
    hipLaunchKernelGGL((bulge_apply_kernel<T, B, false>), dim3(cdiv(m, 64), 1, zgrid()), dim3(64 * BULGE_WAVES), 0, s,
                       p.n, p.b, m, p.blocks_dev, p.stage_off_dev, p.order_dev, p.nstages, v2, p.n, zt, ldzt, zb());
  XTB_LAUNCH_CHECK();
}

}  // namespace

int bulge_width_class(int b) { return b <= 4 ? 4 : b <= 8 ? 8 : b <= 16 ? 16 : b <= 32 ? 32 : 64; }

BackPlan::BackPlan(int n_, int b_) : n(n_), b(b_), g(BULGE_G) {
  if (b > 64) throw std::runtime_error("back-transform: band width above 64");
  // bulge blocks in a valid application order: k ascending, s0 descending. A block's footprint is the kernel's
  // window, g + B - 1 rows (B = b rounded up to 4, 8, 16, 32 or 64), running into the zero padding below row n.
  std::vector<BulgeBlock> bl;
  const int sweeps = n - 2;
  const int foot = g + bulge_width_class(b) - 1;
  if (b > 1 && sweeps > 0) {
    for (int k = 0; 1 + k * b < n; ++k) {
      for (int s0 = ((sweeps - 1) / g) * g; s0 >= 0; s0 -= g) {
        const int r0 = s0 + 1 + k * b;
        if (r0 >= n) continue;
        BulgeBlock w{};
        w.row0 = r0;
        w.height = foot;
        w.count = std::min(std::min(g, sweeps - s0), n - r0);
        w.s0 = s0;
        bl.push_back(w);
      }
    }
  }
  nbulge = (int)bl.size();
  if (nbulge) {
    std::vector<int> rowstage(n + foot, 0), stage(nbulge);
    for (int q = 0; q < nbulge; ++q) {
      int st = 0;
      for (int r = bl[q].row0; r < bl[q].row0 + bl[q].height; ++r) st = std::max(st, rowstage[r]);
      stage[q] = st + 1;
      for (int r = bl[q].row0; r < bl[q].row0 + bl[q].height; ++r) rowstage[r] = st + 1;
      nstages = std::max(nstages, st + 1);
    }
    std::vector<int> off(nstages + 1, 0), order(nbulge);
    for (int q = 0; q < nbulge; ++q) off[stage[q]]++;  // stage is 1-based: counts land at off[1..]
    for (int st = 0; st < nstages; ++st) {
      max_stage = std::max(max_stage, off[st + 1]);
      off[st + 1] += off[st];
    }
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int q = 0; q < nbulge; ++q) order[fill[stage[q] - 1]++] = q;
    std::vector<BulgeBlock> sorted(nbulge);
    for (int q = 0; q < nbulge; ++q) {
      sorted[q] = bl[order[q]];
      rows = std::max(rows, bl[q].row0 + bl[q].height);
    }
    XTB_HIP(hipMalloc(&blocks_dev, sizeof(BulgeBlock) * nbulge));
    XTB_HIP(hipMalloc(&sorted_dev, sizeof(BulgeBlock) * nbulge));
    XTB_HIP(hipMalloc(&stage_off_dev, sizeof(int) * (nstages + 1)));
    XTB_HIP(hipMalloc(&order_dev, sizeof(int) * nbulge));
    XTB_HIP(hipMemcpy(blocks_dev, bl.data(), sizeof(BulgeBlock) * nbulge, hipMemcpyHostToDevice));
    XTB_HIP(hipMemcpy(sorted_dev, sorted.data(), sizeof(BulgeBlock) * nbulge, hipMemcpyHostToDevice));
    XTB_HIP(hipMemcpy(stage_off_dev, off.data(), sizeof(int) * (nstages + 1), hipMemcpyHostToDevice));
    XTB_HIP(hipMemcpy(order_dev, order.data(), sizeof(int) * nbulge, hipMemcpyHostToDevice));
  }
  for (int j = 0; j < n - b - 1; j += b) last = j;
  ppg = std::max(1, GROUP_K / std::max(b, 1));
  ngroups = last < 0 ? 0 : last / (ppg * b) + 1;
}

BackPlan::~BackPlan() {
  if (blocks_dev) (void)hipFree(blocks_dev);
  if (sorted_dev) (void)hipFree(sorted_dev);
  if (stage_off_dev) (void)hipFree(stage_off_dev);
  if (order_dev) (void)hipFree(order_dev);
}

int zt_ld(int m) { return cdiv(std::max(m, 1), 64) * 64; }
// the widest bulge window, BULGE_G + 64 - 1 rows from a block's first row (< n), stays inside the zero rows
int zt_rows(int n) { return n + BULGE_G + 64; }

template <typename T>
size_t back_band_workspace(int n, int b, int m) {
  const int ppg = std::max(1, GROUP_K / std::max(b, 1));
  const size_t groups = (size_t)cdiv(std::max(n, 1), ppg * std::max(b, 1)) + 1;
  return groups * GROUP_K * GROUP_K + (size_t)m * GROUP_K * (cdiv(std::max(n, 1), SPLIT_KC) + 1) + 256;
}

template <typename T>
void back_transform_bulges(hipStream_t s, const BackPlan& p, const T* v2, T* zt, int ldzt, int m, bool transposed) {
  if (p.nbulge == 0 || m <= 0) return;
  if (ldzt < zt_ld(m)) throw std::runtime_error("bulge back-transform: zt leading dimension below zt_ld(m)");
  switch (bulge_width_class(p.b)) {
    case 4: launch_bulges<T, 4>(s, p, v2, zt, ldzt, m, transposed); break;
    case 8: launch_bulges<T, 8>(s, p, v2, zt, ldzt, m, transposed); break;
    case 16: launch_bulges<T, 16>(s, p, v2, zt, ldzt, m, transposed); break;
    case 32: launch_bulges<T, 32>(s, p, v2, zt, ldzt, m, transposed); break;
    default: launch_bulges<T, 64>(s, p, v2, zt, ldzt, m, transposed); break;
  }
}

template <typename T>
void back_transform_band(hipStream_t s, const BackPlan& p, const T* v1, const T* t1, T* zt, int ldzt, int m,
                         T* work) {
  if (p.ngroups == 0 || m <= 0) return;
  const int n = p.n, b = p.b;
  T* tg = work;
  T* w2 = tg + (size_t)(p.ngroups + 1) * GROUP_K * GROUP_K;
  T* parts = w2 + (size_t)m * GROUP_K;
  const size_t lds = 3 * sizeof(T) * GROUP_K * (GROUP_K + 1);
  allow_lds(group_t_kernel<T>, lds);
  hipLaunchKernelGGL(group_t_kernel<T>, dim3(p.ngroups, 1, zgrid()), dim3(256), lds, s, n, b, p.ppg, p.last, v1, t1,
                     tg, zb());
  XTB_LAUNCH_CHECK();
  for (int g = p.ngroups - 1; g >= 0; --g) {
    const int j0 = g * p.ppg * b, r0 = j0 + b, h = n - r0;
    int q = 0;
    while (q < p.ppg && j0 + q * b <= p.last) ++q;
    const int K = (q - 1) * b + std::min(b, n - (j0 + (q - 1) * b + b));
    const T* V = v1 + (int64_t)r0 + (int64_t)j0 * n;
    const int splits = gemm_splitk<T>(s, m, K, h, zt + (int64_t)r0 * ldzt, ldzt, V, n, parts, SPLIT_KC);
    hipLaunchKernelGGL(group_combine_kernel<T>, dim3(cdiv(m, COMBINE_ROWS), 1, zgrid()), dim3(256), 0, s, m, K, splits,
                       parts, tg + (size_t)g * GROUP_K * GROUP_K, w2, zb());
    XTB_LAUNCH_CHECK();
    gemm<T>(s, 'N', 'T', m, h, K, T(-1), w2, m, V, n, T(1), zt + (int64_t)r0 * ldzt, ldzt);
  }
}

template size_t back_band_workspace<double>(int, int, int);
template size_t back_band_workspace<float>(int, int, int);
template void back_transform_bulges<double>(hipStream_t, const BackPlan&, const double*, double*, int, int, bool);
template void back_transform_bulges<float>(hipStream_t, const BackPlan&, const float*, float*, int, int, bool);
template void back_transform_band<double>(hipStream_t, const BackPlan&, const double*, const double*, double*, int,
                                          int, double*);
template void back_transform_band<float>(hipStream_t, const BackPlan&, const float*, const float*, float*, int, int,
                                         float*);

}  // namespace xtbeig
