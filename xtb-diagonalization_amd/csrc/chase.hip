// Stage 2 of the two-stage reduction: band (width b) -> tridiagonal by Schwarz bulge chasing, one Householder
// reflector per bulge (the reference's chaseBulgesLower / chaseBulgeLower, Stages/BulgeChasing.swift:123-321), and
// its reflectors for the back-transformation (backtransform.hip).
//
// Chase: one workgroup; each wavefront owns one sweep at a time (sweep s on wave s mod W) and chases its bulges down
// the band. Bulge k of sweep s waits until sweep s - 1 has finished bulges 0 .. k + 2, published through LDS
// progress counters. (The reference waits for 0 .. k + 3, BulgeChasing.swift:143-153, from the windows rows
// rs - b ..< rs + 2b; counting only the lower-triangle elements a bulge touches -- left block rows R x columns
// rs - b ..< rs, block R x R, block below rows re ..< re + b x columns R -- bulge (s - 1, k + 3) starts at row
// rs + 3b - 1 and shares none of them, while (s - 1, k + 2) shares element (rs + 2b - 1, rs + b - 1).) Every pair of
// bulges that touch a common element therefore runs in sequential order, and the result equals the
// one-sweep-at-a-time chase. Per bulge, with
// H = I - w w^T on rows R = rs ..< re, the lower triangle of the window splits into three blocks:
//   columns lo ..< rs, rows R:   A[R, j] -= w (w^T A[R, j])        lane = column
//   rows re ..< hi, columns R:   A[i, R] -= (A[i, R] w) w^T        lane = row
//   rows and columns R:          A_RR -= w q^T + q w^T, p = A_RR w, q = p - (w^T p) w / 2
// Band storage: element (i, j), 0 <= i - j <= 2b - 1, at i + (2b - 1) j. The reference's compact band uses leading
// dimension 2b for diagonals 0 ..< 2b + 1 (BulgeChasing.swift:8-12); the chase itself never reaches further than
// 2b - 1 below the diagonal (left block: rs + b - 1 - (rs - b); block below: re + b - 1 - rs), so one diagonal less
// is stored and no two stored elements share an address. The band lives in LDS when its 2b n elements fit (FP32 up
// to b = 16, FP64 up to b = 8 at n = 1204), else in global memory (L2-resident at these sizes).

#include "common.hpp"
#include "kernels.hpp"

namespace xtbeig {

namespace {

__device__ inline int ld_acquire(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline void st_release(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <typename T>
struct Band {
  T* ab;
  int ld;  // 2b
  __device__ inline T& operator()(int i, int j) const { return ab[i + ld * j]; }  // band_elems < 2^31
};

// One bulge (s, k) by one wavefront for 32 < b <= 64, the band in device memory or LDS. The reflector comes from the
// column spread over the lanes (lane i holds x_i); every block product is a per-lane sum in ascending order. Each
// block is held in registers, 64 elements per lane: the left block (lane = column), then the block below (lane = row),
// then the diagonal block (lane = row), each loaded, updated and stored before the next is loaded. The three blocks
// are disjoint (columns < rs; rows >= re; rows and columns in R), so storing one before loading the next changes
// nothing, and one block at a time fits a wavefront's registers. Returns false when the sweep has no such bulge.
template <typename T>
__device__ inline bool chase_bulge_blocks(int n, int b, const Band<T>& A, T* __restrict__ v2, int ldv2, int s, int k,
                                          T* wsh, T* qsh) {
  constexpr int MB = 64;  // widest band
  const int lane = threadIdx.x & 63;
  const int rs = s + 1 + k * b;
  if (rs >= n) return false;
  const int re = min(rs + b, n), len = re - rs;
  const int c = (k == 0) ? s : rs - b;
  // larfg on x = A[R, c], lane i holding x_i
  const T xi = (lane < len) ? A(rs + lane, c) : T(0);
  const T alpha = __shfl(xi, 0, 64);
  const T xn2 = wave_sum((lane >= 1 && lane < len) ? xi * xi : T(0));
  T tau = 0, beta = alpha, wi;
  if (sizeof(T) == 4 && !(xn2 >= T(0x1p-100))) {
    // FP32 sum of squares near underflow: the reflector from the entries in FP64 (see lane_reflector)
    const double xd = wave_sum((lane >= 1 && lane < len) ? (double)xi * (double)xi : 0.0);
    double sqd = 0.0, scald = 0.0;
    if (xd != 0.0) {
      const double a = alpha, bd = -copysign(sqrt(a * a + xd), a), taud = (bd - a) / bd;
      beta = T(bd);
      tau = T(taud);
      sqd = sqrt(taud);
      scald = 1.0 / (a - bd);
    }
    wi = (lane == 0) ? T(sqd) : ((lane < len) ? T(sqd * ((double)xi * scald)) : T(0));
  } else {
    T scal = 0;
    if (xn2 != T(0)) {
      beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
      tau = (beta - alpha) / beta;
      scal = T(1) / (alpha - beta);
    }
    const T sq = sqrt(tau);
    wi = (lane == 0) ? sq : ((lane < len) ? sq * (xi * scal) : T(0));
  }
  if (lane < len) {
    v2[(int64_t)(rs + lane) + (int64_t)ldv2 * s] = wi;
    wsh[lane] = wi;
  }
  __builtin_amdgcn_wave_barrier();
  if (tau != T(0)) {
    const int lo = max(rs - b, 0), hi = min(re + b, n);
    // left block: columns lo ..< rs, lane = column (column c is rewritten below)
    if (lane < rs - lo) {
      const int j = lo + lane;
      T L[MB];
#pragma unroll
      for (int i = 0; i < MB; ++i) L[i] = (i < len) ? A(rs + i, j) : T(0);
      T dot = 0;
#pragma unroll
      for (int i = 0; i < MB; ++i)
        if (i < len) dot += wsh[i] * L[i];
#pragma unroll
      for (int i = 0; i < MB; ++i)
        if (i < len) A(rs + i, j) = L[i] - wsh[i] * dot;
    }
    // block below: rows re ..< hi, lane = row
    if (lane < hi - re) {
      const int i = re + lane;
      T Bl[MB];
#pragma unroll
      for (int jj = 0; jj < MB; ++jj) Bl[jj] = (jj < len) ? A(i, rs + jj) : T(0);
      T y = 0;
#pragma unroll
      for (int jj = 0; jj < MB; ++jj)
        if (jj < len) y += Bl[jj] * wsh[jj];
#pragma unroll
      for (int jj = 0; jj < MB; ++jj)
        if (jj < len) A(i, rs + jj) = Bl[jj] - y * wsh[jj];
    }
    // diagonal block R x R (lower triangle stored): p = A_RR w, then A_RR -= w q^T + q w^T
    T D[MB];
    T p = 0;
    if (lane < len) {
      const int i = rs + lane;
#pragma unroll
      for (int jj = 0; jj < MB; ++jj) D[jj] = (jj < len) ? ((jj <= lane) ? A(i, rs + jj) : A(rs + jj, i)) : T(0);
#pragma unroll
      for (int jj = 0; jj < MB; ++jj)
        if (jj < len) p += D[jj] * wsh[jj];
    }
    const T half = wave_sum(lane < len ? wi * p : T(0)) * T(0.5);
    const T q = p - half * wi;
    if (lane < len) qsh[lane] = q;
    __builtin_amdgcn_wave_barrier();
    if (lane < len) {
      const int i = rs + lane;
#pragma unroll
      for (int jj = 0; jj < MB; ++jj)
        if (jj <= lane) A(i, rs + jj) = D[jj] - (wi * qsh[jj] + q * wsh[jj]);
    }
  }
  __threadfence_block();
  // exact zeros where the reflector annihilated
  if (lane < len) A(rs + lane, c) = (lane == 0) ? beta : T(0);
  __threadfence_block();
  return true;
}

// One bulge (s, k) by one wavefront with every load of the bulge issued up front. Band width b <= B (compile time);
// the 64 lanes split as B columns (or rows) x LPC = 64 / B lanes, each lane holding RPL = B / LPC elements of each of
// the three blocks in registers: lane = j * LPC + q handles column (row) j and the elements jj = q + LPC u, u < RPL.
// Sums over a column (row) run across its LPC lanes (lane-xor offsets below LPC).
template <typename T, int B>
__device__ inline T group_sum(T v) {
  return group_sum_lanes<64 / B>(v);
}

template <typename T, int B>
__device__ inline bool chase_bulge_reg(int n, int b, const Band<T>& A, T* __restrict__ v2, int ldv2, int s, int k,
                                       T* wsh, T* qsh) {
  constexpr int LPC = 64 / B, RPL = B / LPC;
  const int lane = threadIdx.x & 63;
  const int j = lane / LPC, q = lane % LPC;
  const int rs = s + 1 + k * b;
  if (rs >= n) return false;
  const int re = min(rs + b, n), len = re - rs;
  const int c = (k == 0) ? s : rs - b;
  const int lo = max(rs - b, 0), hi = min(re + b, n);
  const int nl = rs - lo, nb = hi - re;
  // ---- loads: left block (column lo + j, rows rs + jj), block below (row re + j, columns rs + jj), diagonal block
  // (row rs + j, columns rs + jj, read from the lower triangle)
  T L[RPL], Bl[RPL], D[RPL];
#pragma unroll
  for (int u = 0; u < RPL; ++u) {
    const int jj = q + LPC * u;
    L[u] = (j < nl && jj < len) ? A(rs + jj, lo + j) : T(0);
    Bl[u] = (j < nb && jj < len) ? A(re + j, rs + jj) : T(0);
    D[u] = (j < len && jj < len) ? ((jj <= j) ? A(rs + j, rs + jj) : A(rs + jj, rs + j)) : T(0);
  }
  // ---- larfg on x = A[R, c]: column c is left-block column c - lo
  const int jc = c - lo;
  T part = 0, x0 = 0;
#pragma unroll
  for (int u = 0; u < RPL; ++u) {
    const int jj = q + LPC * u;
    if (j == jc && jj >= 1 && jj < len) part += L[u] * L[u];
    if (j == jc && jj == 0) x0 = L[u];
  }
  const T xn2 = wave_sum(part);
  const T alpha = wave_sum(x0);
  T tau = 0, beta = alpha;
  // FP32 sum of squares near underflow: the reflector from the entries in FP64 (see lane_reflector)
  const bool wide = sizeof(T) == 4 && !(xn2 >= T(0x1p-100));
  double sqd = 0.0, scald = 0.0;
  T sq = 0, scal = 0;
  if (wide) {
    double pd = 0.0;
#pragma unroll
    for (int u = 0; u < RPL; ++u) {
      const int jj = q + LPC * u;
      if (j == jc && jj >= 1 && jj < len) pd += (double)L[u] * (double)L[u];
    }
    const double xd = wave_sum(pd);
    if (xd != 0.0) {
      const double a = alpha, bd = -copysign(sqrt(a * a + xd), a), taud = (bd - a) / bd;
      beta = T(bd);
      tau = T(taud);
      sqd = sqrt(taud);
      scald = 1.0 / (a - bd);
    }
  } else {
    if (xn2 != T(0)) {
      beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
      tau = (beta - alpha) / beta;
      scal = T(1) / (alpha - beta);
    }
    sq = sqrt(tau);
  }
  // w_i for i < len: lane (jc, q) owns x_i for i = q + LPC u; publish w through LDS
  if (j == jc) {
#pragma unroll
    for (int u = 0; u < RPL; ++u) {
      const int jj = q + LPC * u;
      if (jj < len) {
        const T wv = wide ? ((jj == 0) ? T(sqd) : T(sqd * ((double)L[u] * scald)))
                          : ((jj == 0) ? sq : sq * (L[u] * scal));
        wsh[jj] = wv;
        v2[(int64_t)(rs + jj) + (int64_t)ldv2 * s] = wv;
      }
    }
  }
  __builtin_amdgcn_wave_barrier();
  T w[RPL];
#pragma unroll
  for (int u = 0; u < RPL; ++u) {
    const int jj = q + LPC * u;
    w[u] = (jj < len) ? wsh[jj] : T(0);
  }
  if (tau != T(0)) {
    // left block: A[R, lo + j] -= w (w^T A[R, lo + j])
    T dl = 0;
#pragma unroll
    for (int u = 0; u < RPL; ++u) dl += w[u] * L[u];
    dl = group_sum<T, B>(dl);
#pragma unroll
    for (int u = 0; u < RPL; ++u) L[u] -= w[u] * dl;
    // block below: A[re + j, R] -= (A[re + j, R] w) w^T
    T y = 0;
#pragma unroll
    for (int u = 0; u < RPL; ++u) y += Bl[u] * w[u];
    y = group_sum<T, B>(y);
#pragma unroll
    for (int u = 0; u < RPL; ++u) Bl[u] -= y * w[u];
    // diagonal block: p = A_RR w, q = p - (w^T p) w / 2, A_RR -= w q^T + q w^T (lower triangle)
    T pj = 0;
#pragma unroll
    for (int u = 0; u < RPL; ++u) pj += D[u] * w[u];
    pj = group_sum<T, B>(pj);
    const T wj = (j < len) ? wsh[j] : T(0);
    const T half = wave_sum((q == 0 && j < len) ? wj * pj : T(0)) * T(0.5);
    const T qj = pj - half * wj;
    if (q == 0 && j < len) qsh[j] = qj;
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int u = 0; u < RPL; ++u) {
      const int jj = q + LPC * u;
      if (jj < len) D[u] -= wj * qsh[jj] + qj * w[u];
    }
  }
  // ---- stores (the column annihilated gets exact zeros)
#pragma unroll
  for (int u = 0; u < RPL; ++u) {
    const int jj = q + LPC * u;
    if (j < nl && jj < len) A(rs + jj, lo + j) = (j == jc) ? ((jj == 0) ? beta : T(0)) : L[u];
    if (tau != T(0)) {
      if (j < nb && jj < len) A(re + j, rs + jj) = Bl[u];
      if (j < len && jj <= j) A(rs + j, rs + jj) = D[u];
    }
  }
  __threadfence_block();
  return true;
}

// Householder scalars for x = (alpha, x_1 ..): beta, sq = sqrt(tau) and s2 = sqrt(tau) / (alpha - beta), so that
// w = (sq, s2 x_1, ..) = sqrt(tau) v. With d = beta - alpha (|d| = |alpha| + |beta|, d beta > 0): tau = d / beta,
// sqrt(tau) = |d| / sqrt(d beta) and sqrt(tau) / (alpha - beta) = -sign(d) / sqrt(d beta), one square root and one
// reciprocal square root and no division. The hardware FP32 forms (1 ulp) run while every square stays normal; a
// tail below that range takes the division form in FP64.
__device__ inline float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ inline float fast_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ inline double fast_sqrt(double x) { return sqrt(x); }
__device__ inline double fast_rsqrt(double x) { return rsqrt(x); }
template <typename T>
__device__ inline void larfg_scalars(T alpha, T xn2, T& beta, T& sq, T& s2) {
  constexpr T tiny = sizeof(T) == 4 ? T(1.17549435e-38f) : T(2.2250738585072014e-308);
  if (xn2 >= tiny) {
    beta = -copysign(fast_sqrt(alpha * alpha + xn2), alpha);
    const T d = beta - alpha;
    const T r = fast_rsqrt(d * beta);
    sq = fabs(d) * r;
    s2 = -copysign(r, d);
  } else if (xn2 != T(0)) {
    const double a = alpha, bd = -copysign(sqrt(a * a + (double)xn2), a);
    const double sqd = sqrt((bd - a) / bd);
    beta = T(bd);
    sq = T(sqd);
    s2 = T(sqd / (a - bd));
  } else {
    beta = alpha;
    sq = T(0);
    s2 = T(0);
  }
}

// Sum over the H lanes l ^ (m B), m < H (the parts of a column's rows held by one group).
template <int B, int H, typename T>
__device__ inline T part_sum(T v) {
  if constexpr (H >= 2) {
    if constexpr (B == 8) v += dpp_mov<DPP_ROR8>(v);
    else v += __shfl_xor(v, B, 64);
  }
  if constexpr (H >= 4) v += __shfl_xor(v, 2 * B, 64);
  return v;
}

// The reflector of x = (alpha, x_1 ..) for one lane group: beta, and w = sqrt(tau) v over the lane's U entries (in
// place of the x values w holds on entry) and its own element (wc, from xc). In FP32 the sum of squares is used while
// it is at least 2^-100: every square that underflow could drop (below 2^-126) is then under the sum's rounding.
// Below that the sum is taken again in FP64 from the entries and the reflector formed in FP64, so that ||w||^2 = 2
// (an orthogonal reflector) however small the entries are; a reflector formed from a sum that lost squares to
// underflow would not be orthogonal, and the back-transformation would lose orthogonality with it.
template <typename T, int B, int H, int U>
__device__ __forceinline__ void lane_reflector(T alpha, T xn2, int u0, int cl, T xc, T (&w)[U], T& beta, T& wc) {
  if (sizeof(T) == 4 && __builtin_expect(!(xn2 >= T(0x1p-100)), 0)) {
    double xd = (u0 == 0) ? 0.0 : (double)w[0] * (double)w[0];
#pragma unroll
    for (int j = 1; j < U; ++j) xd += (double)w[j] * (double)w[j];
    if (H > 1) xd = part_sum<B, H>(xd);
    double sqd = 0.0, s2d = 0.0;
    beta = alpha;
    if (xd != 0.0) {
      const double a = alpha, bd = -copysign(sqrt(a * a + xd), a), dd = bd - a;
      const double r = 1.0 / sqrt(dd * bd);
      beta = T(bd);
      sqd = fabs(dd) * r;
      s2d = -copysign(r, dd);
    }
#pragma unroll
    for (int j = 0; j < U; ++j) w[j] = (j == 0 && u0 == 0) ? T(sqd) : T(s2d * (double)w[j]);
    wc = (cl == 0) ? T(sqd) : T(s2d * (double)xc);
    return;
  }
  T sq, s2;
  larfg_scalars(alpha, xn2, beta, sq, s2);  // tau = 0 gives w = 0: every update below leaves its block unchanged
#pragma unroll
  for (int j = 0; j < U; ++j) w[j] = (j == 0 && u0 == 0) ? sq : s2 * w[j];
  wc = (cl == 0) ? sq : s2 * xc;
}

// Lane-per-column form for b = B (8, 16 or 32): H * B lanes carry one bulge. Lane (h, cl) owns column rs - b + cl of the
// left block, row re + cl of the block below and row rs + cl of the diagonal block, over the U = B / H rows (columns) u
// = h U ..< (h + 1) U, so the three block products are per-lane sums over U terms followed, for H > 1, by a sum over
// the group's H parts. Every lane recomputes the reflector scalars from column c (read as a broadcast), w^T p is a sum
// over the B lanes of one half, and q is exchanged through LDS. One wavefront advances 64 / (H B) sweeps in lockstep,
// each group on its own sweep and at its own bulge. The band is read and written as if the matrix continued past n with
// zeros (band_elems(n, b) covers b extra columns, and the kernel clears every stored element outside the matrix before
// the chase), so a bulge at the end of the band takes the same branch-free path as any other: the rows past n give x =
// 0, w = 0 there and leave the zeros in place. In band storage A(rs + a, rs + c) sits at rs (ld + 1) + a + ld c = rs 2b
// + a + ld c, so every element a bulge touches is at rs 2b plus a constant of the lane: one set for the first bulge of
// a sweep (whose left block is the single column s = rs - 1) and one for the others. The kernel forms them once.
struct LaneOffsets {
  int x0, x1;  // x = A[R, c]: first bulge, others
  int l0, l1;  // the lane's left block column
  int bl, dr, dc;
};

template <int B, int H>
__device__ inline LaneOffsets lane_offsets(int cl, int h) {
  constexpr int U = B / H, LD = 2 * B - 1;
  const int u0 = h * U;
  return {-LD, -LD * B, u0 - LD, u0 + LD * (cl - B), B + cl + LD * u0, cl + LD * u0, u0 + LD * cl};
}

template <typename T, int B, int H>
__device__ __forceinline__ bool chase_bulge_lanes(int n, T* __restrict__ ab, T* __restrict__ v2s, int rs, bool first,
                                                  int cl, int h, const LaneOffsets& o, T* __restrict__ qs) {
  constexpr int U = B / H;
  constexpr int LD = 2 * B - 1;
  if (rs >= n) return false;
  // left block: columns rs - B ..< rs; for the first bulge of a sweep only column s is nonzero (the columns before it
  // are already tridiagonal), and every lane reads that column
  const int u0 = h * U;
  T* const pos = ab + rs * (2 * B);        // A(rs, rs)
  T* const xp = pos + (first ? o.x0 : o.x1);  // x = A[R, c] at xp[u]
  T* const lp = pos + (first ? o.l0 : o.l1);  // left block column at lp[j]
  T* const bp = pos + o.bl;               // block-below row re + cl, columns rs + u, at bp[j LD]
  T* const dr = pos + o.dr;               // diagonal block row rs + cl, columns rs + u <= rs + cl, at dr[j LD]
  T* const dc = pos + o.dc;               // diagonal block column rs + cl, rows rs + u >= rs + cl, at dc[j]
  T L[U], Bl[U], D[U], w[U];
#pragma unroll
  for (int j = 0; j < U; ++j) {
    w[j] = xp[u0 + j];
    L[j] = lp[j];
    Bl[j] = bp[j * LD];
    // both reads stay inside the band storage (for u > cl, dr[j LD] is another stored element)
    const T r = dr[j * LD], q = dc[j];
    D[j] = (u0 + j <= cl) ? r : q;
  }
  const T alpha = xp[0];
  const T xc = xp[cl];  // this lane's element of x, for its own w
  T xn2 = (u0 == 0) ? T(0) : w[0] * w[0];
#pragma unroll
  for (int j = 1; j < U; ++j) xn2 += w[j] * w[j];
  if (H > 1) xn2 = part_sum<B, H>(xn2);
  T beta, wc;
  lane_reflector<T, B, H, U>(alpha, xn2, u0, cl, xc, w, beta, wc);
  if (h == 0 && rs + cl < n) v2s[rs + cl] = wc;
  T dl = 0, y = 0, p = 0;
#pragma unroll
  for (int j = 0; j < U; ++j) {
    dl += w[j] * L[j];  // left block: w^T A[R, lcol]
    y += Bl[j] * w[j];  // block below: A[re + cl, R] w
    p += D[j] * w[j];   // diagonal block: (A_RR w)[cl]
  }
  if (H > 1) {
    dl = part_sum<B, H>(dl);
    y = part_sum<B, H>(y);
    p = part_sum<B, H>(p);
  }
#pragma unroll
  for (int j = 0; j < U; ++j) {
    L[j] -= w[j] * dl;
    Bl[j] -= y * w[j];
  }
  const T half = group_sum_lanes<B>(wc * p) * T(0.5);
  const T qc = p - half * wc;
  qs[cl] = qc;  // every part stores the same value
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int j = 0; j < U; ++j) D[j] -= wc * qs[u0 + j] + qc * w[j];
  __builtin_amdgcn_wave_barrier();
  if (!first || cl == 0) {
#pragma unroll
    for (int j = 0; j < U; ++j) lp[j] = (cl == 0) ? ((j == 0 && u0 == 0) ? beta : T(0)) : L[j];
  }
#pragma unroll
  for (int j = 0; j < U; ++j) {
    bp[j * LD] = Bl[j];
    if (u0 + j <= cl) dr[j * LD] = D[j];  // the lower triangle of the diagonal block, one lane per row
  }
  __threadfence_block();
  return true;
}

// Band element access by flat index f = i + (2b - 1) j (column j in flat [2b j, 2b j + 2b)) for a band split at a
// column boundary fs between device memory (f < fs) and an LDS copy of the rest.
template <typename T>
struct PtrBand {  // the whole band behind one pointer (device memory or LDS, so each access keeps its address space)
  T* p;
  __device__ inline T& at(int f) const { return p[f]; }
};
template <typename T>
struct SplitBand {
  T* g;
  T* l;  // LDS element f - fs
  int fs;
  __device__ inline T& at(int f) const { return f < fs ? g[f] : l[f - fs]; }
};

// The same bulge through an accessor, for bulges whose columns straddle the split between device memory and LDS
// (pointer offsets from one base would cross the two address spaces).
template <typename T, int B, int H, typename Acc>
__device__ __forceinline__ bool chase_bulge_lanes_acc(int n, const Acc& ab, T* __restrict__ v2s, int rs, bool first,
                                                      int cl, int h, const LaneOffsets& o, T* __restrict__ qs) {
  constexpr int U = B / H;
  constexpr int LD = 2 * B - 1;
  if (rs >= n) return false;
  // left block: columns rs - B ..< rs; for the first bulge of a sweep only column s is nonzero (the columns before it
  // are already tridiagonal), and every lane reads that column
  const int u0 = h * U;
  const int pos = rs * (2 * B);               // A(rs, rs)
  const int xp = pos + (first ? o.x0 : o.x1);  // x = A[R, c] at xp + u
  const int lp = pos + (first ? o.l0 : o.l1);  // left block column at lp + j
  const int bp = pos + o.bl;                   // block-below row re + cl, columns rs + u, at bp + j LD
  const int dr = pos + o.dr;                   // diagonal block row rs + cl, columns rs + u <= rs + cl, at dr + j LD
  const int dc = pos + o.dc;                   // diagonal block column rs + cl, rows rs + u >= rs + cl, at dc + j
  T L[U], Bl[U], D[U], w[U];
#pragma unroll
  for (int j = 0; j < U; ++j) {
    w[j] = ab.at(xp + u0 + j);
    L[j] = ab.at(lp + j);
    Bl[j] = ab.at(bp + j * LD);
    // both reads stay inside the band storage (for u > cl, dr + j LD is another stored element)
    const T r = ab.at(dr + j * LD), q = ab.at(dc + j);
    D[j] = (u0 + j <= cl) ? r : q;
  }
  const T alpha = ab.at(xp);
  const T xc = ab.at(xp + cl);  // this lane's element of x, for its own w
  T xn2 = (u0 == 0) ? T(0) : w[0] * w[0];
#pragma unroll
  for (int j = 1; j < U; ++j) xn2 += w[j] * w[j];
  if (H > 1) xn2 = part_sum<B, H>(xn2);
  T beta, wc;
  lane_reflector<T, B, H, U>(alpha, xn2, u0, cl, xc, w, beta, wc);
  if (h == 0 && rs + cl < n) v2s[rs + cl] = wc;
  T dl = 0, y = 0, p = 0;
#pragma unroll
  for (int j = 0; j < U; ++j) {
    dl += w[j] * L[j];  // left block: w^T A[R, lcol]
    y += Bl[j] * w[j];  // block below: A[re + cl, R] w
    p += D[j] * w[j];   // diagonal block: (A_RR w)[cl]
  }
  if (H > 1) {
    dl = part_sum<B, H>(dl);
    y = part_sum<B, H>(y);
    p = part_sum<B, H>(p);
  }
#pragma unroll
  for (int j = 0; j < U; ++j) {
    L[j] -= w[j] * dl;
    Bl[j] -= y * w[j];
  }
  const T half = group_sum_lanes<B>(wc * p) * T(0.5);
  const T qc = p - half * wc;
  qs[cl] = qc;  // every part stores the same value
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int j = 0; j < U; ++j) D[j] -= wc * qs[u0 + j] + qc * w[j];
  __builtin_amdgcn_wave_barrier();
  if (!first || cl == 0) {
#pragma unroll
    for (int j = 0; j < U; ++j) ab.at(lp + j) = (cl == 0) ? ((j == 0 && u0 == 0) ? beta : T(0)) : L[j];
  }
#pragma unroll
  for (int j = 0; j < U; ++j) {
    ab.at(bp + j * LD) = Bl[j];
    if (u0 + j <= cl) ab.at(dr + j * LD) = D[j];  // the lower triangle of the diagonal block, one lane per row
  }
  __threadfence_block();
  return true;
}

// Band storage positions outside the matrix (rows >= n in the last columns' slots, the b extra columns) -> 0.
template <typename T, int B, typename Acc>
__device__ inline void clear_band_padding(int n, const Acc& ab, int64_t nel) {
  const int64_t first = (int64_t)2 * B * max(0, n - 2 * B);
  for (int64_t q = first + threadIdx.x; q < nel; q += blockDim.x) {
    const int64_t j = q / (2 * B), i = j + q % (2 * B);
    if (i >= n) ab.at((int)q) = T(0);
  }
}

// LDS bytes of the lane chase: the q slots, and the band from flat element fs on (fs = 0: all of it; none for a band
// left in device memory).
template <int B, int H, typename T>
inline size_t chase_lanes_lds(int n, int waves, bool band_in_lds, int64_t fs = 0) {
  const size_t groups = (size_t)waves * (64 / (H * B));
  return sizeof(T) * (groups * B + (band_in_lds ? (size_t)(band_elems(n, B) - fs) : 0));
}

// Where the band lives during the lane chase: all in device memory, all in LDS, or split at flat element fs (a
// column boundary) with the columns past it in LDS. The split keeps the part of the band where most bulge steps
// happen on chip: position p is passed by p sweeps, so with the last columns in LDS the first (n - jsplit)^2 / n^2 of
// the steps never leave the CU.
enum { BAND_DEVICE = 0, BAND_LDS = 1, BAND_SPLIT = 2 };

// dynamic LDS: the band's on-chip part (BAND_LDS: all of it; BAND_SPLIT: flat elements fs ..< nel) first, then one
// q slot of B elements per group
template <typename T, int B, int H, int MODE, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void chase_lanes_kernel(int n, T* __restrict__ ab_global, T* __restrict__ v2,
                                                                 int ldv2, T* __restrict__ d, T* __restrict__ e,
                                                                 int* __restrict__ status, int fs, Zb zb) {
  constexpr int GS = H * B;         // lanes per group
  ab_global = zb.at(ab_global);
  v2 = zb.at(v2);
  d = zb.at(d);
  e = zb.at(e);
  status = zb.at(status);
  constexpr int SPW = 64 / GS;      // sweeps per wavefront
  constexpr int NG = WAVES * SPW;   // sweep groups: group g takes sweeps g, g + NG, ...
  extern __shared__ char smem_raw[];
  __shared__ int prog[NG];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int gl = lane % GS, cl = gl % B, h = gl / B, g = wid * SPW + lane / GS;
  const int64_t nel = band_elems(n, B);
  if (MODE != BAND_SPLIT) fs = (MODE == BAND_LDS) ? 0 : (int)nel;
  const int64_t on_chip = nel - fs;  // band elements in LDS
  T* const sm = reinterpret_cast<T*>(smem_raw);
  T* const qs = sm + on_chip + g * B;  // this group's q exchange slot
  for (int64_t q = t; q < on_chip; q += blockDim.x) sm[q] = ab_global[fs + q];
  __syncthreads();
  const SplitBand<T> sab{ab_global, sm, fs};  // any element (BAND_SPLIT only: a generic pointer per access)
  T* const lab = sm - fs;                      // elements from fs on, on chip (lab[f], f >= fs)
  if constexpr (MODE == BAND_SPLIT) clear_band_padding<T, B>(n, sab, nel);
  else if constexpr (MODE == BAND_LDS) clear_band_padding<T, B>(n, PtrBand<T>{sm}, nel);
  else clear_band_padding<T, B>(n, PtrBand<T>{ab_global}, nel);
  if (t < NG) prog[t] = -1;
  __syncthreads();
  const int sweeps = max(0, n - 2);
  const int KB = (n + B - 1) / B + 8;  // progress code: s * KB + bulges done (monotonic per group)
  const int pg = (g + NG - 1) % NG;    // the group of sweep s - 1
  const LaneOffsets off = lane_offsets<B, H>(cl, h);
  const int js = fs / (2 * B);         // first column on chip
  bool failed = false;
  // per sweep: s, its first row rs = s + 1 + k B, the progress code s KB + k this group publishes after bulge k - 1,
  // the code (s - 1) KB + k + 3 it waits for before bulge k, and v2's column s
  int s = g, k = 0, rs = s + 1, code = s * KB, wait = code - KB + 3;
  T* v2s = v2 + (int64_t)ldv2 * s;
  long spins = 0;
  while (__any(s < sweeps)) {
    bool ready = s < sweeps;
    if (ready && s > 0) ready = ld_acquire(&prog[pg]) >= wait;
    if (ready) {
      spins = 0;
      // a bulge at rs touches columns rs - B ..< rs + B: on chip, in device memory, or across the split
      bool more;
      if (MODE == BAND_LDS || (MODE == BAND_SPLIT && rs - B >= js))
        more = chase_bulge_lanes<T, B, H>(n, lab, v2s, rs, k == 0, cl, h, off, qs);
      else if (MODE == BAND_DEVICE || rs + B <= js)
        more = chase_bulge_lanes<T, B, H>(n, ab_global, v2s, rs, k == 0, cl, h, off, qs);
      else
        more = chase_bulge_lanes_acc<T, B, H>(n, sab, v2s, rs, k == 0, cl, h, off, qs);
      if (more) {
        ++k;
        ++code;
        ++wait;
        rs += B;
        if (gl == 0) st_release(&prog[g], code);
      } else {
        if (gl == 0) st_release(&prog[g], code - k + (KB - 1));
        s += NG;
        k = 0;
        rs = s + 1;
        code = s * KB;
        wait = code - KB + 3;
        v2s = v2 + (int64_t)ldv2 * s;
      }
    } else if (s < sweeps && ++spins > (1L << 24)) {
      failed = true;
      s = sweeps;
    }
    // a wavefront none of whose groups may go on polls again at once: a sleep between polls only delays the step
    // that the flag releases (measured: 2.955 -> 2.93 ms at n = 1204, FP32); the spin count above still bounds the wait
  }
  if (failed && gl == 0) atomicExch(status, 1);
  __syncthreads();
  auto diag = [&](const auto& band) {
    for (int i = t; i < n; i += blockDim.x) {
      d[i] = band.at((2 * B) * i);
      if (i + 1 < n) e[i] = band.at((2 * B) * i + 1);
    }
  };
  if constexpr (MODE == BAND_SPLIT) diag(sab);
  else if constexpr (MODE == BAND_LDS) diag(PtrBand<T>{sm});
  else diag(PtrBand<T>{ab_global});
  __syncthreads();
  for (int64_t q = t; q < on_chip; q += blockDim.x) ab_global[fs + q] = sm[q];
}

template <typename T, int B, bool LDS, int CHASE_WAVES>
__global__ __launch_bounds__(64 * CHASE_WAVES) void chase_kernel(int n, int b, T* __restrict__ ab_global,
                                                                 T* __restrict__ v2, int ldv2, T* __restrict__ d,
                                                                 T* __restrict__ e, int* __restrict__ status,
                                                                 Zb zb) {
  // dynamic LDS: per-wave w and q vectors (b each), then the band when it is LDS-resident
  ab_global = zb.at(ab_global);
  v2 = zb.at(v2);
  d = zb.at(d);
  e = zb.at(e);
  status = zb.at(status);
  extern __shared__ char smem_raw[];
  __shared__ int prog[CHASE_WAVES];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int ld = band_ld(b);
  const int64_t nel = band_elems(n, b);
  T* wbuf = reinterpret_cast<T*>(smem_raw);
  T* qbuf = wbuf + CHASE_WAVES * b;
  T* ab = ab_global;
  if (LDS) {
    T* sm = qbuf + CHASE_WAVES * b;
    for (int64_t q = t; q < nel; q += blockDim.x) sm[q] = ab_global[q];
    ab = sm;
  }
  if (t < CHASE_WAVES) prog[t] = -1;
  __syncthreads();
  const Band<T> A{ab, ld};
  const int sweeps = max(0, n - 2);
  const int KB = (n + b - 1) / b + 8;  // progress code: s * KB + bulges done (monotonic per wave)
  bool failed = false;
  if (b > 1) {
    for (int s = wid; s < sweeps && !failed; s += CHASE_WAVES) {
      const int pw = (s + CHASE_WAVES - 1) % CHASE_WAVES;
      for (int k = 0;; ++k) {
        if (s > 0) {
          const int need = (s - 1) * KB + (k + 3);
          long spins = 0;
          while (ld_acquire(&prog[pw]) < need) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > (1L << 24)) { failed = true; break; }
          }
          if (failed) break;
        }
        const bool more = (B > 0) ? chase_bulge_reg<T, (B > 0 ? B : 8)>(n, b, A, v2, ldv2, s, k, wbuf + wid * b,
                                                                          qbuf + wid * b)
                                  : chase_bulge_blocks<T>(n, b, A, v2, ldv2, s, k, wbuf + wid * b, qbuf + wid * b);
        if (!more) break;
        if (lane == 0) st_release(&prog[wid], s * KB + (k + 1));
      }
      if (lane == 0) st_release(&prog[wid], s * KB + (KB - 1));
    }
  }
  if (failed && lane == 0) atomicExch(status, 1);
  __syncthreads();
  for (int i = t; i < n; i += blockDim.x) {
    d[i] = A(i, i);
    if (i + 1 < n) e[i] = A(i + 1, i);
  }
  if (LDS) {
    __syncthreads();
    for (int64_t q = t; q < nel; q += blockDim.x) ab_global[q] = ab[q];
  }
}

}  // namespace

template <typename T, int B>
void launch_chase(hipStream_t s, int n, int b, T* ab, T* v2, T* d, T* e, int* status) {
  // 16 wavefronts (sweeps in flight); 8 for b > 16, whose per-lane blocks need more than 128 registers
  constexpr int CHASE_WAVES = (B == 0 || B == 32) ? 8 : 16;
  const size_t vec_bytes = 2 * sizeof(T) * CHASE_WAVES * (size_t)b;
  const size_t band_bytes = sizeof(T) * (size_t)band_elems(n, b);
  if (vec_bytes + band_bytes + sizeof(int) * CHASE_WAVES <= 160 * 1024) {
    allow_lds(chase_kernel<T, B, true, CHASE_WAVES>, vec_bytes + band_bytes);
    hipLaunchKernelGGL((chase_kernel<T, B, true, CHASE_WAVES>), dim3(1, 1, zgrid()), dim3(64 * CHASE_WAVES),
                       vec_bytes + band_bytes, s, n, b, ab, v2, n, d, e, status, zb());
  } else {
    hipLaunchKernelGGL((chase_kernel<T, B, false, CHASE_WAVES>), dim3(1, 1, zgrid()), dim3(64 * CHASE_WAVES), vec_bytes,
                       s, n, b, ab, v2, n, d, e, status, zb());
  }
  XTB_LAUNCH_CHECK();
}

template <typename T, int B, int H, int WAVES = 8>
void launch_chase_lanes(hipStream_t s, int n, T* ab, T* v2, T* d, T* e, int* status) {
  constexpr int NG = WAVES * (64 / (H * B));
  constexpr size_t LDS_MAX = 160 * 1024 - sizeof(int) * NG;  // dynamic LDS beside the progress flags
  const int64_t nel = band_elems(n, B);
  const size_t all = chase_lanes_lds<B, H, T>(n, WAVES, true);
  if (all <= LDS_MAX) {
    allow_lds(chase_lanes_kernel<T, B, H, BAND_LDS, WAVES>, all);
    hipLaunchKernelGGL((chase_lanes_kernel<T, B, H, BAND_LDS, WAVES>), dim3(1, 1, zgrid()), dim3(64 * WAVES), all, s,
                       n, ab, v2, n, d, e, status, 0, zb());
  } else {
    // the last columns on chip: fs, a column boundary, leaves at most LDS_MAX bytes for them and the q slots
    const size_t qbytes = chase_lanes_lds<B, H, T>(n, WAVES, false);
    const int64_t cap = (int64_t)((LDS_MAX - qbytes) / sizeof(T));
    int64_t fs = nel - cap;
    fs = (fs + 2 * B - 1) / (2 * B) * (2 * B);
    if (fs < nel - 4 * B * (2 * B)) {  // worth it only with a few columns on chip
      const size_t lds = chase_lanes_lds<B, H, T>(n, WAVES, true, fs);
      allow_lds(chase_lanes_kernel<T, B, H, BAND_SPLIT, WAVES>, lds);
      hipLaunchKernelGGL((chase_lanes_kernel<T, B, H, BAND_SPLIT, WAVES>), dim3(1, 1, zgrid()), dim3(64 * WAVES), lds,
                         s, n, ab, v2, n, d, e, status, (int)fs, zb());
    } else {
      hipLaunchKernelGGL((chase_lanes_kernel<T, B, H, BAND_DEVICE, WAVES>), dim3(1, 1, zgrid()), dim3(64 * WAVES),
                         qbytes, s, n, ab, v2, n, d, e, status, (int)nel, zb());
    }
  }
  XTB_LAUNCH_CHECK();
}

template <typename T>
void chase_bulges(hipStream_t s, int n, int b, T* ab, T* v2, T* d, T* e, int* status) {
  if (b > 64) throw std::runtime_error("chase_bulges: band width must be <= 64");
  // lane-per-column bulges for b = 8, 16 and 32 (one wavefront per bulge at 32), register-resident wavefront bulges
  // for the other b <= 32, wavefront bulges one register-staged block at a time beyond
  if (b == 8 && n > 2) launch_chase_lanes<T, 8, 2>(s, n, ab, v2, d, e, status);
  else if (b == 16 && n > 2) launch_chase_lanes<T, 16, 2>(s, n, ab, v2, d, e, status);
  // FP64 at 32: four wavefronts, so that each has the whole register file of its SIMD (eight spill the split form)
  else if (b == 32 && n > 2) launch_chase_lanes<T, 32, 2, sizeof(T) == 8 ? 4 : 8>(s, n, ab, v2, d, e, status);
  else if (b <= 8) launch_chase<T, 8>(s, n, b, ab, v2, d, e, status);
  else if (b <= 16) launch_chase<T, 16>(s, n, b, ab, v2, d, e, status);
  else if (b <= 32) launch_chase<T, 32>(s, n, b, ab, v2, d, e, status);
  else launch_chase<T, 0>(s, n, b, ab, v2, d, e, status);
}

bool chase_in_lds(int n, int b, bool fp32) {
  constexpr int WAVES = 8;
  if (n > 2 && (b == 8 || b == 16 || b == 32)) {
    size_t need;
    if (b == 8)
      need = fp32 ? chase_lanes_lds<8, 2, float>(n, WAVES, true) : chase_lanes_lds<8, 2, double>(n, WAVES, true);
    else if (b == 16)
      need = fp32 ? chase_lanes_lds<16, 2, float>(n, WAVES, true) : chase_lanes_lds<16, 2, double>(n, WAVES, true);
    else
      need = fp32 ? chase_lanes_lds<32, 2, float>(n, WAVES, true) : chase_lanes_lds<32, 2, double>(n, 4, true);
    const int waves = (b == 32 && !fp32) ? 4 : WAVES;
    return need + sizeof(int) * waves * (64 / (2 * b)) <= 160 * 1024;  // as launch_chase_lanes decides
  }
  // as launch_chase decides
  const size_t el = fp32 ? sizeof(float) : sizeof(double);
  const int waves = b > 16 ? 8 : 16;
  return 2 * el * waves * (size_t)b + el * (size_t)band_elems(n, b) + sizeof(int) * waves <= 160 * 1024;
}

template void chase_bulges<double>(hipStream_t, int, int, double*, double*, double*, double*, int*);
template void chase_bulges<float>(hipStream_t, int, int, float*, float*, float*, float*, int*);

}  // namespace xtbeig
