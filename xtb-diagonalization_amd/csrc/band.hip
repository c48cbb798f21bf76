// Stage 1 of the two-stage reduction: dense symmetric -> lower band of width b (the reference's reduceToBand with the
// combined sy2sb update, Stages/BandForm.swift:28-142); the back-transformation through its panels is in
// backtransform.hip.
//
// Per panel j (columns j ..< j + b, rows r = j + b ..< n, m = n - r, k = min(b, m)):
//   1. panel_qr_kernel: Householder QR of the m x b panel by one workgroup of 256 threads. The panel is processed
//      in sub-panels of SB columns held in registers (R rows per thread), earlier reflectors applied left-looking.
//      Writes R into the band (zeros below), V explicitly (unit diagonal, zeros above) into the back-transform store
//      v1 and into the update buffer [V | . | V], and tau.
//   2. panel_gram_kernel + panel_tvt_kernel: G = V^T V, the forward columnwise T (I - V T V^T = H_1 ... H_k) and
//      VT = V T, over row blocks of 64 on many workgroups.
//   3. X = A2 VT (split-K GEMM, partials summed in fixed order); panel_xy_kernel + panel_w_kernel: Y = VT^T X,
//      W = X - 1/2 V Y, again over row blocks;
//      A2 -= [V W] [W V]^T, one rank-2k GEMM on the matrix cores (full storage: both triangles of the trailing block
//      stay current).
// The fused path for band widths 8 and 16 (panel_factor_kernel, panel_xpart / xfin, panel_update_factor_kernel)
// keeps only A2's lower triangle current: its updates write the 64 x 64 tiles on and below the diagonal, and its
// product X = A2 V reads an element above the diagonal from its mirror, so a panel moves half the update bytes. (The
// panels' tile grids shift by b from one panel to the next, so it is the elementwise lower triangle that is current,
// not a fixed set of tiles.) The general path, which runs first when a panel is too tall for the fused one, updates
// both triangles.
// After the last panel the lower band is copied into the chase's band storage (element (i, j), 0 <= i - j <= 2b - 1,
// at i + (2b - 1) j; see chase.hip).
#include "common.hpp"
#include "gemm_tile.hpp"
#include "kernels.hpp"

namespace xtbeig {

namespace {

constexpr int QR_THREADS = 256;

// Block-wide sums of NV values per thread (4 wavefronts); the totals land in tot[0 .. NV). red: 4 * NV scratch.
template <typename T, int NV>
__device__ inline void block_sums(const T (&v)[NV], T* red, T* tot) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    const T s = wave_sum(v[c]);
    if (lane == 0) red[wid * NV + c] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) tot[threadIdx.x] = (red[threadIdx.x] + red[NV + threadIdx.x]) +
                                           (red[2 * NV + threadIdx.x] + red[3 * NV + threadIdx.x]);
  __syncthreads();
}

// Householder QR of the panel p (m x nb, ld lda; nb <= 64) by one workgroup of 256 threads: row i = t + 256 r of
// the current sub-panel of SB columns lives in registers of thread t (R rows per thread, m <= 256 R). Earlier
// sub-panels' reflectors are applied left-looking (re-read from v1). R overwrites the upper triangle of p's first
// min(m, nb) rows, zeros below. V (explicit: unit diagonal, zeros above) goes to v1 (ld ldv1) and to buf (ld m) at
// column offsets 0 and 2k; tau to tau.
template <typename T, int SB, int R>
__global__ __launch_bounds__(QR_THREADS) void panel_qr_kernel(int m, int nb, T* __restrict__ p, int lda,
                                                              T* __restrict__ v1, int ldv1, T* __restrict__ buf,
                                                              T* __restrict__ tau, Zb zb) {
  p = zb.at(p);
  v1 = zb.at(v1);
  buf = zb.at(buf);
  tau = zb.at(tau);
  __shared__ T red[4 * SB];
  __shared__ T tot[SB];
  __shared__ T sh_alpha;
  __shared__ double redd[4], totd[1];
  const int t = threadIdx.x;
  const int k = min(m, nb);
  T x[R][SB];
  for (int c0 = 0; c0 < nb; c0 += SB) {
    const int sb = min(SB, nb - c0);
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int c = 0; c < SB; ++c) {
        const int i = t + r * QR_THREADS;
        x[r][c] = (i < m && c < sb) ? p[idx(i, c0 + c, lda)] : T(0);
      }
    for (int q = 0; q < min(c0, k); ++q) {
      const T tq = tau[q];
      T vq[R];
      T part[SB];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = t + r * QR_THREADS;
        vq[r] = (i < m) ? v1[idx(i, q, ldv1)] : T(0);
      }
#pragma unroll
      for (int c = 0; c < SB; ++c) {
        T s = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) s += vq[r] * x[r][c];
        part[c] = s;
      }
      block_sums<T, SB>(part, red, tot);
#pragma unroll
      for (int c = 0; c < SB; ++c) {
        const T f = tq * tot[c];
#pragma unroll
        for (int r = 0; r < R; ++r) x[r][c] -= vq[r] * f;
      }
    }
#pragma unroll
    for (int c = 0; c < SB; ++c) {
      const int col = c0 + c;
      if (col < k) {
        // the column's tail norm and the reflector's scalars in FP64 (for FP32 panels too: a square of an FP32 entry
        // never underflows in FP64, so that the reflector stays orthogonal however small the column is); the entries
        // v_i = x_i / (alpha - beta) are rounded once to T
        double part1[1];
        part1[0] = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int i = t + r * QR_THREADS;
          if (i == col) sh_alpha = x[r][c];
          if (i > col && i < m) part1[0] += (double)x[r][c] * (double)x[r][c];
        }
        block_sums<double, 1>(part1, redd, totd);
        const double alpha = sh_alpha, xn2 = totd[0];
        double beta = alpha, tq = 0.0, scal = 0.0;
        if (xn2 != 0.0) {
          beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
          tq = (beta - alpha) / beta;
          scal = 1.0 / (alpha - beta);
        }
        T v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int i = t + r * QR_THREADS;
          v[r] = (i == col) ? T(1) : ((i > col && i < m) ? T((double)x[r][c] * scal) : T(0));
          if (i == col) x[r][c] = T(beta);
          else if (i > col) x[r][c] = 0;
        }
        if (t == 0) tau[col] = T(tq);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int i = t + r * QR_THREADS;
          if (i < m) {
            v1[idx(i, col, ldv1)] = v[r];
            buf[idx(i, col, m)] = v[r];
            buf[idx(i, 2 * k + col, m)] = v[r];
          }
        }
        if (c + 1 < SB) {
          T part[SB];
#pragma unroll
          for (int cc = 0; cc < SB; ++cc) {
            T s = 0;
            if (cc > c) {
#pragma unroll
              for (int r = 0; r < R; ++r) s += v[r] * x[r][cc];
            }
            part[cc] = s;
          }
          block_sums<T, SB>(part, red, tot);
#pragma unroll
          for (int cc = 0; cc < SB; ++cc) {
            if (cc > c) {
              const T f = T(tq) * tot[cc];
#pragma unroll
              for (int r = 0; r < R; ++r) x[r][cc] -= v[r] * f;
            }
          }
        }
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int c = 0; c < SB; ++c) {
        const int i = t + r * QR_THREADS;
        if (i < m && c < sb) p[idx(i, c0 + c, lda)] = x[r][c];
      }
    __syncthreads();  // v1 and tau of this sub-panel visible to the next one
  }
}

// Butterfly reduce-scatter of NB values per lane (NB = 8 or 16) over the 64 lanes of a wavefront: each exchange
// step halves the values a lane keeps (the half selected by one lane bit) and adds the partner's copy of that half;
// the last steps sum the one remaining value over the lanes that hold the same index. Lane l returns the wavefront
// total of value index (l >> 2) & 15 (NB = 16) or (l >> 3) & 7 (NB = 8). Pairings: lane bits 5 and 4 by xor, bit 3
// by rotating a row of 16 by 8, bit 2 by the mirror within 8 lanes (it also flips bits 1 and 0, which are not yet
// decided), bits 1 and 0 by quad permutes.
template <typename T, int H>
__device__ inline void bfly_split(const T (&v)[2 * H], bool up, T (&out)[H], int step) {
#pragma unroll
  for (int j = 0; j < H; ++j) {
    const T send = up ? v[j] : v[j + H];
    T recv;
    if (step == 32) recv = xor_lanes<32>(send);
    else if (step == 16) recv = xor_lanes<16>(send);
    else if (step == 8) recv = dpp_mov<DPP_ROR8>(send);
    else recv = dpp_mov<DPP_HALF_MIRROR>(send);
    out[j] = (up ? v[j + H] : v[j]) + recv;
  }
}

template <typename T, int NB>
__device__ inline T bfly_reduce(const T (&v)[NB]) {
  const int lane = threadIdx.x & 63;
  static_assert(NB == 8 || NB == 16, "bfly_reduce: NB must be 8 or 16");
  T a[NB / 2], b[NB / 4], c[NB / 8];
  bfly_split<T, NB / 2>(v, lane & 32, a, 32);
  bfly_split<T, NB / 4>(a, lane & 16, b, 16);
  bfly_split<T, NB / 8>(b, lane & 8, c, 8);
  T r;
  if constexpr (NB == 16) {
    T d[1];
    bfly_split<T, 1>(c, lane & 4, d, 4);
    r = d[0];
  } else {
    r = c[0];
    r += dpp_mov<DPP_HALF_MIRROR>(r);
  }
  r += dpp_mov<DPP_XOR2>(r);
  r += dpp_mov<DPP_XOR1>(r);
  return r;
}

// Householder QR of the m x NB panel p (ld lda; NB = 8 or 16, m <= 256 R) by one workgroup of 256 threads, with the
// panel's block-reflector data: row i = t + 256 r lives in registers of thread t. The column loop is a real loop
// (the body is the same for every column, so the kernel's code stays small enough for the instruction cache): the
// register array rotates by one column per step, so that the current column is always at position 0, the columns
// still to be reduced follow it and the finished reflectors v_0 .. v_{c-1} sit at the end. Column c takes one block
// reduction of NB sums over the rows i > c of x_c[i] times position j: at j = 0 the reflector's tail norm, at an
// unfinished column's position the part of v^T x_j below row c (row c itself comes through LDS), at v_l's position
// sum_{i > c} v_l[i] x_c[i], so that G(l, c) = v_l^T v_c = v_l[c] + scal_c * that sum comes with the same reduction.
// T (I - V T V^T = H_1 ... H_k, forward columnwise: T(0:c, c) = -tau_c T(0:c, 0:c) G(0:c, c), T(c, c) = tau_c) is
// built in LDS one column per step. Outputs: R over p's upper triangle (the part below it is left as it was: only the
// band, diagonals 0 .. b, is read afterwards), V (unit diagonal, zeros above) into v1 (ld ldv1), tau, and T into tt
// (ld ldt, zeros below).
// The factorization proper, on the panel already in registers (x[r][c] = p(t + 256 r, c), zero past m); shared by
// panel_factor_kernel and panel_update_factor_kernel.
template <typename T, int NB, int R>
__device__ __forceinline__ void factor_panel(int m, T (&x)[R][NB], T* __restrict__ p, int lda, T* __restrict__ v1,
                                             int ldv1, T* __restrict__ tau, T* __restrict__ tt, int ldt) {
  constexpr int WAVES = 4, SPREAD = 64 / NB;  // lanes holding each reduced value
  __shared__ T red[2][WAVES][NB];
  __shared__ T rowb[2][NB];
  __shared__ T Ts[NB][NB + 1];
  __shared__ double redd[WAVES * NB], totd[NB];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int k = min(m, NB);
  // T's strictly lower part is read (times a zero mask) by the branch-free T update: it must hold zeros, not whatever
  // an earlier kernel left in LDS (0 * NaN is NaN). The first column's barrier orders these stores before any read.
  for (int e = t; e < NB * (NB + 1); e += 256) (&Ts[0][0])[e] = T(0);
  // unrolled by 4: the rotation's register moves are needed only where the unrolled body loops back
#pragma unroll 4
  for (int c = 0; c < NB; ++c) {
    T newcol[R];
    if (c < k) {
      const int par = c & 1;
      T part[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) part[j] = 0;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const T xc = (r > 0 || t > c) ? x[r][0] : T(0);  // rows past m hold zeros
#pragma unroll
        for (int j = 0; j < NB; ++j) part[j] += xc * x[r][j];
      }
      if (t == c) {
#pragma unroll
        for (int j = 0; j < NB; ++j) rowb[par][j] = x[0][j];
      }
      const T wsum = bfly_reduce<T, NB>(part);
      if (lane % SPREAD == 0) red[par][wid][lane / SPREAD] = wsum;
      __syncthreads();
      T tot[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) tot[j] = (red[par][0][j] + red[par][1][j]) + (red[par][2][j] + red[par][3][j]);
      // FP32: row min(t, NB - 1) of T's finished columns read with the sums, before the branch below, so that the
      // LDS reads still overlap (the entries past column c - 1 are masked below)
      T trow[NB];
      if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int j = 0; j < NB; ++j) trow[j] = Ts[min(t, NB - 1)][max(j - NB + c, 0)];
      }
      const T alpha = rowb[par][0], xn2 = tot[0];
      T beta = alpha, tq = 0, scal = 0;
      // g[j] = v_l^T v_c at a finished column's position j = NB - c + l, v^T x_j / ... at an unfinished one's
      T g[NB];
      // FP32 sum of squares near underflow (below 2^-100, where squares and products that underflow are no longer
      // negligible): every sum of the step again in FP64 (the same products), so that the reflector and its column
      // of T stay orthogonal however small the column is. The test is uniform (every thread read the same sum).
      bool wide = false;
      if constexpr (sizeof(T) == 4) wide = __builtin_amdgcn_readfirstlane(xn2 >= T(0x1p-100) ? 0 : 1) != 0;
      if (wide) {
        double partd[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) partd[j] = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const double xc = (r > 0 || t > c) ? (double)x[r][0] : 0.0;
#pragma unroll
          for (int j = 0; j < NB; ++j) partd[j] += xc * (double)x[r][j];
        }
        block_sums<double, NB>(partd, redd, totd);
        const double xd = totd[0];
        double scald = 0.0;
        if (xd != 0.0) {
          const double a = alpha, bd = -copysign(sqrt(a * a + xd), a);
          beta = T(bd);
          tq = T((bd - a) / bd);
          scald = 1.0 / (a - bd);
        }
#pragma unroll
        for (int j = 0; j < NB; ++j) g[j] = T((double)rowb[par][j] + scald * totd[j]);
        // the reflector's entries, each rounded once from FP64, replace the column: the update below scales by 1
#pragma unroll
        for (int r = 0; r < R; ++r)
          if (r > 0 || t > c) x[r][0] = T((double)x[r][0] * scald);
        scal = T(1);
      } else {
        if (xn2 != T(0)) {
          beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
          tq = (beta - alpha) / beta;
          scal = T(1) / (alpha - beta);
        }
#pragma unroll
        for (int j = 0; j < NB; ++j) g[j] = rowb[par][j] + scal * tot[j];
      }
      {
        // T(t, c) = -tau_c sum_{l = t}^{c - 1} T(t, l) G(l, c), branch-free so that the LDS reads overlap
        const int tr = min(t, NB - 1);
        T sum = 0;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
          const int l = j - NB + c;
          const T gm = (l >= t) ? g[j] : T(0);  // the mask on g keeps the LDS read unconditional
          sum += (sizeof(T) == 4 ? trow[j] : Ts[tr][max(l, 0)]) * gm;
        }
        if (t < c) Ts[t][c] = -tq * sum;
      }
      if (t == c) {
        Ts[c][c] = tq;
        tau[c] = tq;
      }
      T f[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) f[j] = (j >= 1 && j < NB - c) ? tq * g[j] : T(0);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (r == 0 && t == c) {
#pragma unroll
          for (int j = 1; j < NB; ++j) x[0][j] -= f[j];
          newcol[0] = beta;
        } else if (r > 0 || t > c) {
          const T vi = x[r][0] * scal;
#pragma unroll
          for (int j = 1; j < NB; ++j) x[r][j] -= vi * f[j];
          newcol[r] = vi;
        } else {
          newcol[r] = x[r][0];  // R(i, c) above the diagonal
        }
      }
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) newcol[r] = x[r][0];  // columns past k (m < NB) are not reflected
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
      for (int j = 0; j + 1 < NB; ++j) x[r][j] = x[r][j + 1];
      x[r][NB - 1] = newcol[r];
    }
  }
  __syncthreads();  // T complete
  if (t < k * k) {
    const int r = t % k, c = t / k;
    tt[r + (int64_t)c * ldt] = (r <= c) ? Ts[r][c] : T(0);
  }
  if (t < NB && t < m) {
#pragma unroll
    for (int c = 0; c < NB; ++c)
      if (t <= c) p[idx(t, c, lda)] = x[0][c];
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = t + 256 * r;
    if (i >= m) continue;
#pragma unroll
    for (int c = 0; c < NB; ++c)
      if (c < k) v1[idx(i, c, ldv1)] = (i > c) ? x[r][c] : (i == c ? T(1) : T(0));
  }
}

template <typename T, int NB, int R>
__global__ __launch_bounds__(256) void panel_factor_kernel(int m, T* __restrict__ p, int lda, T* __restrict__ v1,
                                                           int ldv1, T* __restrict__ tau, T* __restrict__ tt,
                                                           int ldt, Zb zb) {
  p = zb.at(p);
  v1 = zb.at(v1);
  tau = zb.at(tau);
  tt = zb.at(tt);
  const int t = threadIdx.x;
  T x[R][NB];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = t + 256 * r;
#pragma unroll
    for (int c = 0; c < NB; ++c) x[r][c] = (i < m) ? p[idx(i, c, lda)] : T(0);
  }
  factor_panel<T, NB, R>(m, x, p, lda, v1, ldv1, tau, tt, ldt);
}

// X = A2 V T for the panel's trailing block A2 (m x m, ld lda) and its reflectors V (m x k, ld ldv), in two launches
// without any inter-workgroup synchronization: panel_xpart_kernel forms the split-K partials A2(rows, K split)
// V(K split, :) on the matrix cores (grid: row blocks of 64 x K splits of XKC; A2's block and V's rows are staged in
// LDS with all loads in flight at once; the product is taken as X^T = V^T A2^T so that lanes run along rows), and
// panel_xfin_kernel (one workgroup per row block) sums them in split order, applies T, stores X and the block's
// part of V^T X. Y = (V T)^T X = T^T (V^T X) is completed by the update kernel.
template <typename T> struct XSplit;
template <> struct XSplit<float> { static constexpr int K = 128; };
template <> struct XSplit<double> { static constexpr int K = 64; };

template <typename T>
__global__ __launch_bounds__(256) void panel_xpart_kernel(int m, int k, const T* __restrict__ a2, int lda,
                                                          const T* __restrict__ v, int ldv, T* __restrict__ xpart,
                                                          Zb zb) {
  a2 = zb.at(a2);
  v = zb.at(v);
  xpart = zb.at(xpart);
  using namespace gemm_detail;
  using acc_t = typename Mfma<T>::acc_t;
  constexpr int KC = XSplit<T>::K;
  __shared__ T As[KC][64 + 1];
  __shared__ T Vs[KC][16 + 1];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int i0 = blockIdx.x * 64, ks = blockIdx.y, l0 = ks * KC, kl = min(KC, m - l0);
  constexpr int NA = 64 * KC / 256, NV = 16 * KC / 256;
  // Only A2's lower triangle is current (the updates write the tiles on and below the diagonal): an element above it
  // is read from its mirror. A 64-column sub-chunk entirely above the diagonal is read with the lanes along the
  // mirror's rows (kk fastest), so that those loads stay coalesced; element (i, kk) of a 64 x 64 sub-chunk is the
  // thread's f = t + 256 (q mod 16) with (i, kk) = (f mod 64, f / 64), or (f / 64, f mod 64) when mirrored.
  auto mirrored = [&](int q) { return i0 + 63 < l0 + 64 * (q / 16); };
  auto coords = [&](int q, int& i, int& kk) {
    const int f = t + 256 * (q % 16);
    if (mirrored(q)) {
      i = f / 64;
      kk = 64 * (q / 16) + f % 64;
    } else {
      i = f % 64;
      kk = 64 * (q / 16) + f / 64;
    }
  };
  T ra[NA], rv[NV];  // every load in flight before the first LDS store
#pragma unroll
  for (int q = 0; q < NA; ++q) {
    int i, kk;
    coords(q, i, kk);
    const int gi = i0 + i, gk = l0 + kk;
    ra[q] = (gi < m && kk < kl) ? (gi >= gk ? a2[idx(gi, gk, lda)] : a2[idx(gk, gi, lda)]) : T(0);
  }
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const int e = t + 256 * q, kk = e % KC, c = e / KC;
    rv[q] = (kk < kl && c < k) ? v[idx(l0 + kk, c, ldv)] : T(0);
  }
#pragma unroll
  for (int q = 0; q < NA; ++q) {
    int i, kk;
    coords(q, i, kk);
    As[kk][i] = ra[q];
  }
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const int e = t + 256 * q;
    Vs[e % KC][e / KC] = rv[q];
  }
  __syncthreads();
  acc_t acc = {0, 0, 0, 0};
  const int il = wid * 16 + (lane & 15);
#pragma unroll 8
  for (int s0 = 0; s0 < KC; s0 += 4) {
    const int kk = s0 + (lane >> 4);
    acc = Mfma<T>::op(Vs[kk][lane & 15], As[kk][il], acc);  // A(c, kk) = V(kk, c), B(kk, i) = A2(i, kk)
  }
  const int i = i0 + il;
  T* xp = xpart + (size_t)ks * m * k;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int c = Mfma<T>::row(lane, r);
    if (i < m && c < k) xp[i + (size_t)c * m] = acc[r];
  }
}

template <typename T>
__global__ __launch_bounds__(256) void panel_xfin_kernel(int m, int k, const T* __restrict__ xpart, int nks,
                                                         const T* __restrict__ tt, int ldt, const T* __restrict__ v,
                                                         int ldv, T* __restrict__ x, T* __restrict__ ypart, Zb zb) {
  xpart = zb.at(xpart);
  tt = zb.at(tt);
  v = zb.at(v);
  x = zb.at(x);
  ypart = zb.at(ypart);
  __shared__ T Xs[64][17];
  __shared__ T Vr[64][17];
  __shared__ T Ts[16][17];
  const int t = threadIdx.x, rb = blockIdx.x, i0 = rb * 64;
  // the thread's four elements: the partials of all four are loaded a chunk of splits at a time, every load of a
  // chunk in flight at once, and summed in split order
  constexpr int CH = 8;
  T sum[4];
  size_t off[4];
  bool ok[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = t + 256 * q, r = e % 64, c = e / 64;
    ok[q] = i0 + r < m && c < k;
    off[q] = (size_t)(i0 + r) + (size_t)c * m;
    sum[q] = 0;
  }
  const size_t split = (size_t)m * k;
  for (int s0 = 0; s0 < nks; s0 += CH) {
    T pv[4][CH];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int j = 0; j < CH; ++j) pv[q][j] = (ok[q] && s0 + j < nks) ? xpart[(size_t)(s0 + j) * split + off[q]] : T(0);
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int j = 0; j < CH; ++j)
        if (s0 + j < nks) sum[q] += pv[q][j];
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = t + 256 * q, r = e % 64, c = e / 64;
    Xs[r][c] = sum[q];
    Vr[r][c] = ok[q] ? v[idx(i0 + r, c, ldv)] : T(0);
  }
  {
    const int r = t % 16, c = t / 16;
    Ts[r][c] = (r < k && c < k && r <= c) ? tt[r + (size_t)c * ldt] : T(0);
  }
  __syncthreads();
  T xv[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = t + 256 * q, r = e % 64, c = e / 64;
    T sum = 0;
#pragma unroll
    for (int l = 0; l < 16; ++l) sum += Xs[r][l] * Ts[l][c];
    xv[q] = sum;
    if (i0 + r < m && c < k) x[(i0 + r) + (size_t)c * m] = sum;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = t + 256 * q;
    Xs[e % 64][e / 64] = xv[q];
  }
  __syncthreads();
  const int a = t % 16, c = t / 16;
  T vx = 0;
#pragma unroll 8
  for (int r = 0; r < 64; ++r) vx += Vr[r][a] * Xs[r][c];
  ypart[(size_t)rb * 256 + t] = vx;  // (V^T X)(a, c) over this block's rows
}

// A2 -= V W^T + W V^T with W = X - V Y / 2 and Y = T^T (V^T X) (the row blocks' parts of V^T X summed in block
// order), one 64 x 64 tile of A2 per workgroup on the matrix cores (computed as its transpose, so that lanes run
// along A2's rows). The tile's rows I and columns J stage [V(I) | W(I)] and [W(J) | V(J)] (zero past k) in LDS;
// K = 32.
template <typename T>
struct UpdateLds {
  T Ai[64][33];
  T Bj[64][33];
  T Ys[16][17];
  T Ss[16][17];
  T Ts[16][17];
};

// Y = T^T (V^T X) into L.Ys (k x k, zero elsewhere), from the row blocks' parts and T; ends with a barrier.
template <typename T>
__device__ __forceinline__ void form_y(int k, const T* __restrict__ ypart, int nrb, const T* __restrict__ tt,
                                       int ldt, UpdateLds<T>& L) {
  const int t = threadIdx.x;
  {
    const int r = t % 16, c = t / 16;
    // the row blocks' parts summed in block order, up to 32 loads in flight at once
    T sum = 0;
    for (int q0 = 0; q0 < nrb; q0 += 32) {
      T pv[32];
#pragma unroll
      for (int j = 0; j < 32; ++j) pv[j] = (q0 + j < nrb) ? ypart[(size_t)(q0 + j) * 256 + t] : T(0);
#pragma unroll
      for (int j = 0; j < 32; ++j)
        if (q0 + j < nrb) sum += pv[j];
    }
    L.Ss[r][c] = sum;
    L.Ts[r][c] = (r < k && c < k && r <= c) ? tt[r + (size_t)c * ldt] : T(0);
  }
  __syncthreads();
  {
    const int r = t % 16, c = t / 16;
    T sum = 0;
#pragma unroll
    for (int l = 0; l < 16; ++l) sum += L.Ts[l][r] * L.Ss[l][c];
    L.Ys[r][c] = (r < k && c < k) ? sum : T(0);
  }
  __syncthreads();
}

// Tile q of the lower-triangle enumeration (I >= J, row tile I of the q-th row I (I + 1) / 2 + J): A2 is symmetric and
// only its tiles on and below the diagonal are updated (and read, by the next panel's product and factorization).
__device__ inline void lower_tile(int q, int& I, int& J) {
  I = (int)((sqrtf(8.0f * (float)q + 1.0f) - 1.0f) * 0.5f);
  while (I * (I + 1) / 2 > q) --I;
  while ((I + 1) * (I + 2) / 2 <= q) ++I;
  J = q - I * (I + 1) / 2;
}

// One 64 x 64 tile (rows I0.., columns J0..) of the update. Elements (i, j) with i >= skip and j < skip are left as
// they are (the fused next-panel factorization takes those).
template <typename T>
__device__ __forceinline__ void update_tile(int m, int k, T* __restrict__ a2, int lda, const T* __restrict__ v,
                                            int ldv, const T* __restrict__ x, const T* __restrict__ ypart, int nrb,
                                            const T* __restrict__ tt, int ldt, int I0, int J0, int skip,
                                            UpdateLds<T>& L) {
  using namespace gemm_detail;
  using acc_t = typename Mfma<T>::acc_t;
  auto& Ai = L.Ai;
  auto& Bj = L.Bj;
  auto& Ys = L.Ys;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int wm = wid & 1, wn = wid >> 1;
  // this thread's 16 elements of the A2 tile, in the accumulator layout, loaded first so that their latency overlaps
  // everything below
  T old[2][2][4];
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) {
      const int i = I0 + wn * 32 + bj * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = J0 + wm * 32 + bi * 16 + Mfma<T>::row(lane, r);
        old[bi][bj][r] = (i < m && j < m) ? a2[idx(i, j, lda)] : T(0);
      }
    }
  form_y<T>(k, ypart, nrb, tt, ldt, L);
  for (int e = t; e < 64 * 16; e += 256) {
    const int r = e % 64, c = e / 64;
    const bool oi = I0 + r < m && c < k, oj = J0 + r < m && c < k;
    Ai[r][c] = oi ? v[idx(I0 + r, c, ldv)] : T(0);
    Ai[r][16 + c] = oi ? x[(I0 + r) + (size_t)c * m] : T(0);
    Bj[r][16 + c] = oj ? v[idx(J0 + r, c, ldv)] : T(0);
    Bj[r][c] = oj ? x[(J0 + r) + (size_t)c * m] : T(0);
  }
  __syncthreads();
  T wi[4], wj[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = t + 256 * q, r = e % 64, c = e / 64;
    T si = 0, sj = 0;
#pragma unroll
    for (int l = 0; l < 16; ++l) {
      si += Ai[r][l] * Ys[l][c];
      sj += Bj[r][16 + l] * Ys[l][c];
    }
    wi[q] = Ai[r][16 + c] - T(0.5) * si;
    wj[q] = Bj[r][c] - T(0.5) * sj;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = t + 256 * q, r = e % 64, c = e / 64;
    Ai[r][16 + c] = wi[q];
    Bj[r][c] = wj[q];
  }
  __syncthreads();
  // C^T(j, i) = sum_kk Bj[j][kk] Ai[i][kk]: wavefront (wm, wn) takes j in 32 wm .., i in 32 wn ..
  acc_t acc[2][2];
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = acc_t{0, 0, 0, 0};
#pragma unroll
  for (int kk = 0; kk < 32; kk += 4) {
    const int kr = kk + (lane >> 4), c = lane & 15;
    T av[2], bv[2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi) av[bi] = Bj[wm * 32 + bi * 16 + c][kr];
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) bv[bj] = Ai[wn * 32 + bj * 16 + c][kr];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = Mfma<T>::op(av[bi], bv[bj], acc[bi][bj]);
  }
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) {
      const int i = I0 + wn * 32 + bj * 16 + (lane & 15);
      if (i >= m) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = J0 + wm * 32 + bi * 16 + Mfma<T>::row(lane, r);
        if (j < m && (j >= skip || i < skip)) a2[idx(i, j, lda)] = old[bi][bj][r] - acc[bi][bj][r];
      }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void panel_update_kernel(int m, int k, T* __restrict__ a2, int lda,
                                                           const T* __restrict__ v, int ldv, const T* __restrict__ x,
                                                           const T* __restrict__ ypart, int nrb,
                                                           const T* __restrict__ tt, int ldt, Zb zb) {
  a2 = zb.at(a2);
  v = zb.at(v);
  x = zb.at(x);
  ypart = zb.at(ypart);
  tt = zb.at(tt);
  __shared__ UpdateLds<T> L;
  int I, J;
  lower_tile(blockIdx.x, I, J);
  update_tile<T>(m, k, a2, lda, v, ldv, x, ypart, nrb, tt, ldt, I * 64, J * 64, 0, L);
}

// The update of this panel (as panel_update_kernel, workgroups 1 ..) fused with the factorization of the next panel
// (workgroup 0): the next panel is A2(NB .., 0 ..< NB), and workgroup 0 applies this panel's update to just those
// m - NB rows and NB columns itself, A2(i, c) - sum_l V(i, l) W(c, l) + W(i, l) V(c, l) with W = X - V Y / 2, holding
// the result in registers for the factorization, while the other workgroups update the rest of A2 and leave those
// elements alone. No workgroup waits for another; workgroup 0 is dispatched first, so the factorization runs beside
// the bulk of the update instead of after it.
template <typename T, int NB, int R>
__global__ __launch_bounds__(256) void panel_update_factor_kernel(int m, T* __restrict__ a2, int lda,
                                                                  const T* __restrict__ v, int ldv,
                                                                  const T* __restrict__ x,
                                                                  const T* __restrict__ ypart, int nrb,
                                                                  const T* __restrict__ tt, int ldt,
                                                                  T* __restrict__ v1n, int ldv1,
                                                                  T* __restrict__ taun, T* __restrict__ ttn,
                                                                  int ldtn, Zb zb) {
  a2 = zb.at(a2);
  v = zb.at(v);
  x = zb.at(x);
  ypart = zb.at(ypart);
  tt = zb.at(tt);
  v1n = zb.at(v1n);
  taun = zb.at(taun);
  ttn = zb.at(ttn);
  __shared__ UpdateLds<T> L;
  const int t = threadIdx.x;
  if (blockIdx.x > 0) {
    int I, J;
    lower_tile(blockIdx.x - 1, I, J);
    update_tile<T>(m, NB, a2, lda, v, ldv, x, ypart, nrb, tt, ldt, I * 64, J * 64, NB, L);
    return;
  }
  const int mn = m - NB;  // the next panel's rows
  // Every load that does not wait on Y first: this thread's rows of A2(:, 0 ..< NB) (the factorization holds them
  // in registers anyway), of V and X where they fit (FP32, 8 columns: 64 more registers at 4 rows per thread), and
  // the panel columns' rows of V and X; their latency then overlaps the small products below.
  constexpr bool EARLY = sizeof(T) == 4 && NB == 8 && R <= 4;
  constexpr int RE = EARLY ? R : 1;
  T xr[R][NB], vr[RE][NB], xq[RE][NB];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = NB + t + 256 * r;
    const bool ok = t + 256 * r < mn;
#pragma unroll
    for (int c = 0; c < NB; ++c) xr[r][c] = ok ? a2[idx(i, c, lda)] : T(0);
  }
  if constexpr (EARLY) {
#pragma unroll
    for (int r = 0; r < RE; ++r) {
      const int i = NB + t + 256 * r;
      const bool ok = t + 256 * r < mn;
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        vr[r][c] = ok ? v[idx(i, c, ldv)] : T(0);
        xq[r][c] = ok ? x[i + (size_t)c * m] : T(0);
      }
    }
  }
  const int cc = t % NB, lc = (t / NB) % NB;  // (c, l) of the panel columns' part for t < NB * NB
  T vcl[NB];
#pragma unroll
  for (int l2 = 0; l2 < NB; ++l2) vcl[l2] = v[idx(cc, l2, ldv)];
  const T xcl = x[cc + (size_t)lc * m];
  form_y<T>(NB, ypart, nrb, tt, ldt, L);
  // V(c, :) and W(c, :) of the panel's columns c < NB, in Ai[c][0 ..< NB] and Ai[c][16 ..]
  if (t < NB * NB) {
    T w = xcl;
#pragma unroll
    for (int l2 = 0; l2 < NB; ++l2) w -= T(0.5) * vcl[l2] * L.Ys[l2][lc];
    L.Ai[cc][lc] = vcl[lc];
    L.Ai[cc][16 + lc] = w;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = NB + t + 256 * r;  // row of A2
    if (t + 256 * r < mn) {
      T vi[NB], wi[NB];
#pragma unroll
      for (int l = 0; l < NB; ++l) {
        if constexpr (EARLY) {
          vi[l] = vr[r][l];
          wi[l] = xq[r][l];
        } else {
          vi[l] = v[idx(i, l, ldv)];
          wi[l] = x[i + (size_t)l * m];
        }
      }
#pragma unroll
      for (int l = 0; l < NB; ++l) {
        T sy = 0;
#pragma unroll
        for (int l2 = 0; l2 < NB; ++l2) sy += vi[l2] * L.Ys[l2][l];
        wi[l] -= T(0.5) * sy;
      }
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        T u = 0;
#pragma unroll
        for (int l = 0; l < NB; ++l) u += vi[l] * L.Ai[c][16 + l] + wi[l] * L.Ai[c][l];
        xr[r][c] -= u;
      }
    }
  }
  factor_panel<T, NB, R>(mn, xr, a2 + idx(NB, 0, lda), lda, v1n, ldv1, taun, ttn, ldtn);
}

// The per-panel reductions over the m rows run on cdiv(m, PR) workgroups, PR rows each; k x k partial products per
// row block go to a partial buffer (ld KM, KM^2 elements per block) and are summed by the consumer in block order.
// KM: the widest panel of the general path (band widths up to 64).
constexpr int PR = 64;
constexpr int KM = 64;
constexpr int KM1 = KM + 1;

// Rows i0 ..< i0 + PR of the m x k (ld m) matrix a into sh (PR x KM1), zeros past m and k.
template <typename T>
__device__ inline void stage_rows(int m, int k, int i0, const T* __restrict__ a, T (*sh)[KM1]) {
  for (int e = threadIdx.x; e < PR * KM; e += blockDim.x) {
    const int i = e % PR, c = e / PR;
    sh[i][c] = (i0 + i < m && c < k) ? a[(int64_t)(i0 + i) + (int64_t)c * m] : T(0);
  }
}

// out[r + KM c] = sum_i A(i, r) B(i, c) over the staged rows (k x k; sym: r <= c only).
template <typename T>
__device__ inline void rows_atb(int k, const T (*A)[KM1], const T (*B)[KM1], T* __restrict__ out, bool sym) {
  for (int u = 0; u < KM * KM / 256; ++u) {
    const int o = threadIdx.x + 256 * u;
    if (o < k * k) {
      const int r = o % k, c = o / k;
      if (!sym || r <= c) {
        T s0 = 0, s1 = 0;
#pragma unroll 8
        for (int i = 0; i < PR; i += 2) {
          s0 += A[i][r] * B[i][c];
          s1 += A[i + 1][r] * B[i + 1][c];
        }
        out[r + KM * c] = s0 + s1;
      }
    }
  }
}

// k x k (ld KM) sum of the nblk partials into sh (KM x KM1).
template <typename T>
__device__ inline void sum_partials(int k, int nblk, const T* __restrict__ part, T (*sh)[KM1], bool sym) {
  for (int o = threadIdx.x; o < k * k; o += blockDim.x) {
    const int r = o % k, c = o / k;
    if (sym && r > c) continue;
    T sum = 0;
    for (int q = 0; q < nblk; ++q) sum += part[(int64_t)q * KM * KM + r + KM * c];
    sh[r][c] = sum;
  }
}

// Gp[blk] = V(rows)^T V(rows), upper triangle.
template <typename T>
__global__ __launch_bounds__(256) void panel_gram_kernel(int m, int k, const T* __restrict__ v, T* __restrict__ gp,
                                                         Zb zb) {
  v = zb.at(v);
  gp = zb.at(gp);
  __shared__ T Vs[PR][KM1];
  stage_rows<T>(m, k, blockIdx.x * PR, v, Vs);
  __syncthreads();
  rows_atb<T>(k, Vs, Vs, gp + (int64_t)blockIdx.x * KM * KM, true);
}

// G = sum of the Gram partials; the forward columnwise T (I - V T V^T = H_1 ... H_k; T(0:c, c) = -tau_c T(0:c, 0:c)
// G(0:c, c), T(c, c) = tau_c), written by workgroup 0 to tt (ld ldt); VT(rows) = V(rows) T.
template <typename T>
__global__ __launch_bounds__(256) void panel_tvt_kernel(int m, int k, const T* __restrict__ v,
                                                        const T* __restrict__ tau, const T* __restrict__ gp, int nblk,
                                                        T* __restrict__ tt, int ldt, T* __restrict__ vt, Zb zb) {
  v = zb.at(v);
  tau = zb.at(tau);
  gp = zb.at(gp);
  tt = zb.at(tt);
  vt = zb.at(vt);
  __shared__ T G[KM][KM1];
  __shared__ T Ts[KM][KM1];
  __shared__ T Vs[PR][KM1];
  const int t = threadIdx.x, i0 = blockIdx.x * PR;
  sum_partials<T>(k, nblk, gp, G, true);
  stage_rows<T>(m, k, i0, v, Vs);
  for (int e = t; e < KM * KM1; e += blockDim.x) (&Ts[0][0])[e] = T(0);
  __syncthreads();
  for (int c = 0; c < k; ++c) {
    if (t < c) {
      T sum = 0;
      for (int l = t; l < c; ++l) sum += Ts[t][l] * G[l][c];
      Ts[t][c] = -tau[c] * sum;
    }
    if (t == 0) Ts[c][c] = tau[c];
    __syncthreads();
  }
  if (blockIdx.x == 0)
    for (int e = t; e < k * k; e += blockDim.x) {
      const int r = e % k, c = e / k;
      tt[r + (int64_t)c * ldt] = (r <= c) ? Ts[r][c] : T(0);
    }
  for (int e = t; e < PR * k; e += blockDim.x) {
    const int i = e % PR, c = e / PR;
    if (i0 + i >= m) continue;
    T sum = 0;
    for (int l = 0; l <= c; ++l) sum += Vs[i][l] * Ts[l][c];
    vt[(int64_t)(i0 + i) + (int64_t)c * m] = sum;
  }
}

// X(rows) = sum of the nsplit split-K partials (each m x k) into x; Yp[blk] = VT(rows)^T X(rows).
template <typename T>
__global__ __launch_bounds__(256) void panel_xy_kernel(int m, int k, const T* __restrict__ vt,
                                                       const T* __restrict__ xpart, int nsplit, T* __restrict__ x,
                                                       T* __restrict__ yp, Zb zb) {
  vt = zb.at(vt);
  xpart = zb.at(xpart);
  x = zb.at(x);
  yp = zb.at(yp);
  __shared__ T Xs[PR][KM1];
  __shared__ T Ws[PR][KM1];
  const int t = threadIdx.x, i0 = blockIdx.x * PR;
  const int64_t mk = (int64_t)m * k;
  for (int e = t; e < PR * KM; e += blockDim.x) {
    const int i = e % PR, c = e / PR;
    T sum = 0;
    if (i0 + i < m && c < k) {
      const int64_t o = (int64_t)(i0 + i) + (int64_t)c * m;
      for (int q = 0; q < nsplit; ++q) sum += xpart[o + q * mk];
      x[o] = sum;
    }
    Xs[i][c] = sum;
  }
  stage_rows<T>(m, k, i0, vt, Ws);
  __syncthreads();
  rows_atb<T>(k, Ws, Xs, yp + (int64_t)blockIdx.x * KM * KM, false);
}

// Y = sum of the Yp partials; W(rows) = X(rows) - V(rows) Y / 2, in place of x.
template <typename T>
__global__ __launch_bounds__(256) void panel_w_kernel(int m, int k, const T* __restrict__ v, const T* __restrict__ yp,
                                                      int nblk, T* __restrict__ x, Zb zb) {
  v = zb.at(v);
  yp = zb.at(yp);
  x = zb.at(x);
  __shared__ T Y[KM][KM1];
  __shared__ T Vs[PR][KM1];
  const int t = threadIdx.x, i0 = blockIdx.x * PR;
  sum_partials<T>(k, nblk, yp, Y, false);
  stage_rows<T>(m, k, i0, v, Vs);
  __syncthreads();
  for (int e = t; e < PR * k; e += blockDim.x) {
    const int i = e % PR, c = e / PR;
    if (i0 + i >= m) continue;
    T sum = 0;
    for (int l = 0; l < k; ++l) sum += Vs[i][l] * Y[l][c];
    x[(int64_t)(i0 + i) + (int64_t)c * m] -= T(0.5) * sum;
  }
}

// Lower band of a (diagonals 0 .. b) into band storage ab, diagonals b + 1 .. 2b - 1 zero.
template <typename T>
__global__ void extract_band_kernel(int n, int b, const T* __restrict__ a, int lda, T* __restrict__ ab, Zb zb) {
  a = zb.at(a);
  ab = zb.at(ab);
  const int j = blockIdx.x, ld = band_ld(b);
  for (int dd = threadIdx.x; dd <= ld; dd += blockDim.x) {
    const int i = j + dd;
    if (i < n) ab[(int64_t)i + (int64_t)ld * j] = (dd <= b) ? a[idx(i, j, lda)] : T(0);
  }
}

template <typename T, int SB, int R>
void launch_qr(hipStream_t s, int m, int nb, T* p, int lda, T* v1, int ldv1, T* buf, T* tau) {
  hipLaunchKernelGGL((panel_qr_kernel<T, SB, R>), dim3(1, 1, zgrid()), dim3(QR_THREADS), 0, s, m, nb, p, lda, v1, ldv1,
                     buf, tau, zb());
}

template <typename T>
void launch_panel_qr(hipStream_t s, int m, int nb, T* p, int lda, T* v1, int ldv1, T* buf, T* tau) {
  constexpr int SB = 16;  // wider sub-panels exceed the unroller's budget and spill the register array
  switch ((m + QR_THREADS - 1) / QR_THREADS) {
    case 0: case 1: launch_qr<T, SB, 1>(s, m, nb, p, lda, v1, ldv1, buf, tau); break;
    case 2: launch_qr<T, SB, 2>(s, m, nb, p, lda, v1, ldv1, buf, tau); break;
    case 3: launch_qr<T, SB, 3>(s, m, nb, p, lda, v1, ldv1, buf, tau); break;
    case 4: launch_qr<T, SB, 4>(s, m, nb, p, lda, v1, ldv1, buf, tau); break;
    // FP64 with five or more rows per thread: 8-column sub-panels keep the register array out of scratch
    case 5: launch_qr<T, sizeof(T) == 8 ? 8 : SB, 5>(s, m, nb, p, lda, v1, ldv1, buf, tau); break;
    case 6: launch_qr<T, sizeof(T) == 8 ? 8 : SB, 6>(s, m, nb, p, lda, v1, ldv1, buf, tau); break;
    case 7: launch_qr<T, sizeof(T) == 8 ? 8 : SB, 7>(s, m, nb, p, lda, v1, ldv1, buf, tau); break;
    case 8: launch_qr<T, sizeof(T) == 8 ? 8 : SB, 8>(s, m, nb, p, lda, v1, ldv1, buf, tau); break;
    default: throw std::runtime_error("reduce_to_band: n > 2048 + b is not supported");
  }
  XTB_LAUNCH_CHECK();
}

}  // namespace

constexpr int SPLIT_ROWS = 128;  // K per split of the X = A2 VT product

template <typename T>
size_t band_workspace(int n, int b) {
  // tau (b) | buf [V W V] (3 n b) | VT (n b) | split-K partials of X (ceil(n / 64) n b: the fused path splits K by
  // 64 (FP64) or 128 (FP32), the general path by SPLIT_ROWS) | k x k partials of the row blocks (ceil(n / PR) KM^2)
  return (size_t)b + 4 * (size_t)n * b + (size_t)((n + 63) / 64) * n * b + (size_t)((n + PR - 1) / PR) * KM * KM +
         64;
}

template <typename T, int NB, int R>
void launch_factor(hipStream_t s, int m, T* p, int lda, T* v1, int ldv1, T* tau, T* tt, int ldt) {
  hipLaunchKernelGGL((panel_factor_kernel<T, NB, R>), dim3(1, 1, zgrid()), dim3(256), 0, s, m, p, lda, v1, ldv1, tau,
                     tt, ldt, zb());
  XTB_LAUNCH_CHECK();
}

// Fused panel factorization (QR, T and VT in one workgroup) for panel widths 8 and 16 and m <= 2048; false when the
// panel needs the general path.
template <typename T>
bool panel_factor(hipStream_t s, int m, int nb, T* p, int lda, T* v1, int ldv1, T* tau, T* tt, int ldt) {
  const int r = (m + 255) / 256;
  if (nb == 8) {
    switch (r) {
      case 0: case 1: launch_factor<T, 8, 1>(s, m, p, lda, v1, ldv1, tau, tt, ldt); return true;
      case 2: launch_factor<T, 8, 2>(s, m, p, lda, v1, ldv1, tau, tt, ldt); return true;
      case 3: launch_factor<T, 8, 3>(s, m, p, lda, v1, ldv1, tau, tt, ldt); return true;
      case 4: launch_factor<T, 8, 4>(s, m, p, lda, v1, ldv1, tau, tt, ldt); return true;
      case 5: launch_factor<T, 8, 5>(s, m, p, lda, v1, ldv1, tau, tt, ldt); return true;
      case 6: launch_factor<T, 8, 6>(s, m, p, lda, v1, ldv1, tau, tt, ldt); return true;
      case 7: launch_factor<T, 8, 7>(s, m, p, lda, v1, ldv1, tau, tt, ldt); return true;
      case 8: launch_factor<T, 8, 8>(s, m, p, lda, v1, ldv1, tau, tt, ldt); return true;
      default: return false;
    }
  }
  if (nb == 16) {
    switch (r) {
      case 0: case 1: launch_factor<T, 16, 1>(s, m, p, lda, v1, ldv1, tau, tt, ldt); return true;
      case 2: launch_factor<T, 16, 2>(s, m, p, lda, v1, ldv1, tau, tt, ldt); return true;
      case 3: launch_factor<T, 16, 3>(s, m, p, lda, v1, ldv1, tau, tt, ldt); return true;
      case 4: launch_factor<T, 16, 4>(s, m, p, lda, v1, ldv1, tau, tt, ldt); return true;
      case 5: launch_factor<T, 16, 5>(s, m, p, lda, v1, ldv1, tau, tt, ldt); return true;
      case 6: launch_factor<T, 16, 6>(s, m, p, lda, v1, ldv1, tau, tt, ldt); return true;
      // FP64 with more than 6 rows per thread: the register array no longer fits; the general path takes it
      case 7:
        if constexpr (sizeof(T) == 8) return false;
        else launch_factor<T, 16, 7>(s, m, p, lda, v1, ldv1, tau, tt, ldt);
        return true;
      case 8:
        if constexpr (sizeof(T) == 8) return false;
        else launch_factor<T, 16, 8>(s, m, p, lda, v1, ldv1, tau, tt, ldt);
        return true;
      default: return false;
    }
  }
  return false;
}

template <typename T, int NB, int R>
void launch_update_factor(hipStream_t s, int m, T* a2, int lda, const T* v, int ldv, const T* x, const T* kpart,
                          int nrb, const T* tt, int ldt, T* v1n, int ldv1, T* tau, T* ttn) {
  hipLaunchKernelGGL((panel_update_factor_kernel<T, NB, R>), dim3(1 + nrb * (nrb + 1) / 2, 1, zgrid()), dim3(256), 0,
                     s, m, a2, lda, v, ldv, x, kpart, nrb, tt, ldt, v1n, ldv1, tau, ttn, NB, zb());
  XTB_LAUNCH_CHECK();
}

// This panel's update fused with the next panel's factorization (band widths 8 and 16, the next panel's m - b rows
// within panel_factor's limits); false when the two must run separately.
template <typename T>
bool update_factor(hipStream_t s, int m, int b, T* a2, int lda, const T* v, int ldv, const T* x, const T* kpart,
                   int nrb, const T* tt, T* v1n, int ldv1, T* tau, T* ttn) {
  const int r = cdiv(m - b, 256);
#define XTB_UF(NB, R) launch_update_factor<T, NB, R>(s, m, a2, lda, v, ldv, x, kpart, nrb, tt, b, v1n, ldv1, tau, ttn)
  if (b == 8) {
    switch (r) {
      case 0: case 1: XTB_UF(8, 1); return true;
      case 2: XTB_UF(8, 2); return true;
      case 3: XTB_UF(8, 3); return true;
      case 4: XTB_UF(8, 4); return true;
      case 5: XTB_UF(8, 5); return true;
      case 6: XTB_UF(8, 6); return true;
      case 7: XTB_UF(8, 7); return true;
      case 8: XTB_UF(8, 8); return true;
      default: return false;
    }
  }
  if (b == 16) {
    switch (r) {
      case 0: case 1: XTB_UF(16, 1); return true;
      case 2: XTB_UF(16, 2); return true;
      case 3: XTB_UF(16, 3); return true;
      case 4: XTB_UF(16, 4); return true;
      case 5: XTB_UF(16, 5); return true;
      case 6: XTB_UF(16, 6); return true;
      case 7:
        if constexpr (sizeof(T) == 8) return false;
        else XTB_UF(16, 7);
        return true;
      case 8:
        if constexpr (sizeof(T) == 8) return false;
        else XTB_UF(16, 8);
        return true;
      default: return false;
    }
  }
#undef XTB_UF
  return false;
}

template <typename T>
void reduce_to_band(hipStream_t s, int n, int b, T* a, int lda, T* v1, T* t1, T* ab, T* work) {
  if (b > KM || b < 1) throw std::runtime_error("reduce_to_band: band width must be in [1, 64]");
  T* tau = work;
  T* buf = tau + b;
  T* vt = buf + 3 * (size_t)n * b;
  T* xpart = vt + (size_t)n * b;
  T* kpart = xpart + (size_t)((n + 63) / 64) * n * b;
  // v1 is zero outside the panels' V, so that consecutive panels read as one block [V_a V_b ...] from row r_a
  set_zero<T>(s, n, n, v1, n);
  bool factored = false;  // this panel was factored by the previous panel's fused update
  for (int j = 0; j < n - b - 1; j += b) {
    const int r = j + b, m = n - r, k = std::min(b, m);
    T* panel = a + idx(r, j, lda);
    T* v = v1 + idx(r, j, n);
    T* tt = t1 + idx(0, j, b);
    T* a2 = a + idx(r, r, lda);
    T* x = buf + (size_t)m * k;  // slot 1 of [V | W | V]
    const int nblk = cdiv(m, PR);
    if (factored || panel_factor<T>(s, m, b, panel, lda, v, n, tau, tt, b)) {
      // X = A2 V T (split-K partials, then their sum with T applied and the row blocks' parts of V^T X), then
      // A2 -= V W^T + W V^T with Y and W formed per tile, with the next panel's factorization when it can ride along
      T* xb = buf;
      const int nks = cdiv(m, XSplit<T>::K), nrb = cdiv(m, 64);
      hipLaunchKernelGGL(panel_xpart_kernel<T>, dim3(nrb, nks, zgrid()), dim3(256), 0, s, m, k, a2, lda, v, n,
                         xpart, zb());
      XTB_LAUNCH_CHECK();
      hipLaunchKernelGGL(panel_xfin_kernel<T>, dim3(nrb, 1, zgrid()), dim3(256), 0, s, m, k, xpart, nks, tt, b, v,
                         n, xb, kpart, zb());
      XTB_LAUNCH_CHECK();
      const int jn = j + b;
      factored = jn < n - b - 1 && k == b &&
                 update_factor<T>(s, m, b, a2, lda, v, n, xb, kpart, nrb, tt, v1 + idx(jn + b, jn, n), n, tau,
                                  t1 + idx(0, jn, b));
      if (!factored) {
        hipLaunchKernelGGL(panel_update_kernel<T>, dim3(nrb * (nrb + 1) / 2, 1, zgrid()), dim3(256), 0, s, m, k, a2,
                           lda, v, n, xb, kpart, nrb, tt, b, zb());
        XTB_LAUNCH_CHECK();
      }
      continue;
    }
    factored = false;
    {
      launch_panel_qr<T>(s, m, b, panel, lda, v, n, buf, tau);
      hipLaunchKernelGGL(panel_gram_kernel<T>, dim3(nblk, 1, zgrid()), dim3(256), 0, s, m, k, buf, kpart,
                         zb());  // V^T V parts
      XTB_LAUNCH_CHECK();
      hipLaunchKernelGGL(panel_tvt_kernel<T>, dim3(nblk, 1, zgrid()), dim3(256), 0, s, m, k, buf, tau, kpart, nblk,
                         tt, b, vt, zb());  // T, VT = V T
      XTB_LAUNCH_CHECK();
    }
    const int nsplit = gemm_splitk<T>(s, m, k, m, a2, lda, vt, m, xpart, SPLIT_ROWS);  // A2 VT in parts
    hipLaunchKernelGGL(panel_xy_kernel<T>, dim3(nblk, 1, zgrid()), dim3(256), 0, s, m, k, vt, xpart, nsplit, x,
                       kpart, zb());  // X, Y
    XTB_LAUNCH_CHECK();
    hipLaunchKernelGGL(panel_w_kernel<T>, dim3(nblk, 1, zgrid()), dim3(256), 0, s, m, k, buf, kpart, nblk, x,
                       zb());  // W
    XTB_LAUNCH_CHECK();
    gemm<T>(s, 'N', 'T', m, m, 2 * k, T(-1), buf, m, buf + (size_t)m * k, m, T(1), a2, lda);  // A2 -= [V W][W V]^T
  }
  hipLaunchKernelGGL(extract_band_kernel<T>, dim3(n, 1, zgrid()), dim3(128), 0, s, n, b, a, lda, ab, zb());
  XTB_LAUNCH_CHECK();
}

template size_t band_workspace<double>(int, int);
template size_t band_workspace<float>(int, int);
template void reduce_to_band<double>(hipStream_t, int, int, double*, int, double*, double*, double*, double*);
template void reduce_to_band<float>(hipStream_t, int, int, float*, int, float*, float*, float*, float*);
}  // namespace xtbeig
