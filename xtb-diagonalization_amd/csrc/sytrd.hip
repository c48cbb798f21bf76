// One-stage reduction to tridiagonal form in the LAPACK dsytrd('U') convention, and its back-transformation
// (dormtr 'L', 'U', 'N'): the sytrd.c entry points xtb_{d,s}sytrd_upper, xtb_dsytrd_full and xtb_{d,s}ormtr_upper.
//
// Reduction (the reference's tridiagonalize / latrdUpper / sytd2Upper, Sources/Diagonalization/
// Tridiagonalization.swift:510-618, which is LAPACK dsytrd with ilaenv's nb = 32, nx = 32): panels of 32 columns
// from the last one (dlatrd 'U'), each followed by the rank-64 update of the leading block, then dsytd2 on the
// remaining kk <= 32 columns. The matrix is kept in full storage on the GPU (both triangles current, the reference's
// symmetric = false variant): the symmetric matrix-vector product is a column-wise gemv 'T' (coalesced), and the
// trailing update is two GEMMs on the matrix cores. For the upper-triangle entry point only the upper triangle is
// copied back, so the caller sees the dsytrd('U') result.
//
// One column step of a panel (column c, 0-based; W column wc; the panel's later columns c + 1 ..< nc already done)
// is three launches, each over many workgroups with deterministic partial sums:
//   trd_colupd   the previous column's w += -tau/2 (w^T v) v (its alpha from the partials, in every workgroup),
//                then A(0:c+1, c) -= A(:, c+1:nc) W(c, wc+1:)^T + W(:, wc+1:) A(c, c+1:nc)^T; partial |x|^2. The
//                one value of that w a workgroup needs outside its own rows, W(c, wc+1), reaches every workgroup
//                as a scalar (before the update) and gets the same update there.
//   trd_symv     larfg from the partials (every workgroup, redundantly), v in LDS; y = A(0:c, 0:c) v, and as
//                2m more columns of the same product u = W(0:c, wc+1:)^T v and u2 = A(0:c, c+1:nc)^T v
//   trd_wrow     w = tau (y - A(:, c+1:nc) u - W(:, wc+1:) u2); A(0:c, c) = v; partial w^T v
// The last column of a panel gets its w update alone (trd_walpha) before the trailing update uses W.
// Back-transformation (TridiagonalBackTransformation.swift:16-45): 32-reflector blocks in increasing order, V
// explicit (unit, zeros below), T = larft('B', 'C') from the block's Gram, then W = V^T Z, W2 = T W, Z -= V W2.
#include "common.hpp"
#include "kernels.hpp"

#include <algorithm>

namespace xtbeig {

namespace {

constexpr int NB = 32, NX = 32;
constexpr int ROWS = 256;  // rows per workgroup of the row kernels
constexpr int SYMV_COLS = 8;  // columns of y per workgroup (two per wavefront)

// With prev: the previous column's (c + 1) w += alpha v first, alpha = -tau/2 (w^T v) from its partials pw (npw of
// them, summed in order), its e and tau written by workgroup 0; W(c, wc+1) from the scalar wnext plus the same
// update. Then A(0:c+1, c) -= A(0:c+1, c+1:nc) W(c, wc+1:wc+1+m)^T + W(0:c+1, wc+1:) A(c, c+1:nc)^T (m = nc - 1 - c
// may be 0); part[blockIdx.x] = sum of A(r, c)^2 over this block's rows r <= c - 2 (the larfg vector).
template <typename T>
__global__ __launch_bounds__(ROWS) void trd_colupd_kernel(int c, int m, T* __restrict__ a, int lda,
                                                          T* __restrict__ w, int ldw, int wc, T* __restrict__ part,
                                                          int prev, const T* __restrict__ vbuf,
                                                          const T* __restrict__ scal3, const T* __restrict__ pw,
                                                          int npw, T* __restrict__ e, T* __restrict__ tau) {
  __shared__ T wr[NB], ar[NB], red[ROWS / 64];
  const int t = threadIdx.x, r = blockIdx.x * ROWS + t;
  T alpha = 0;
  if (prev) {
    T dot = 0;
    for (int q = 0; q < npw; ++q) dot += pw[q];
    const T tv = scal3[0];
    alpha = T(-0.5) * tv * dot;
    if (r <= c) w[idx(r, wc + 1, ldw)] += alpha * vbuf[r];
    if (blockIdx.x == 0 && t == 0) {
      e[c] = scal3[1];
      tau[c] = tv;
    }
  }
  if (t < m) {
    wr[t] = (t == 0 && prev) ? scal3[2] + alpha * vbuf[c] : w[idx(c, wc + 1 + t, ldw)];
    ar[t] = a[idx(c, c + 1 + t, lda)];
  }
  __syncthreads();
  T x = 0;
  if (r <= c) {
    T v = a[idx(r, c, lda)];
    for (int q = 0; q < m; ++q) v -= a[idx(r, c + 1 + q, lda)] * wr[q] + w[idx(r, wc + 1 + q, ldw)] * ar[q];
    if (m > 0) a[idx(r, c, lda)] = v;
    if (r <= c - 2) x = v * v;
  }
  const T s = block_sum(x, red);
  if (t == 0) part[blockIdx.x] = s;
}

// larfg of (alpha = A(c-1, c), x = A(0:c-1, c)) from the norm partials, computed by every workgroup; v in LDS
// (v(c-1) = 1). Workgroup 0 publishes v, tau, beta. Columns j < c + 2m of the product: y(j) = A(0:c, j)^T v into
// W(j, wc) for j < c; then u(q) = W(0:c, wc+1+q)^T v into uu[q] and u2(q) = A(0:c, c+1+q)^T v into uu[NB + q] and
// W(c+1+q, wc) (where LAPACK's dlatrd leaves it), q < m.
template <typename T>
__global__ __launch_bounds__(256) void trd_symv_kernel(int c, int m, const T* __restrict__ a, int lda,
                                                       const T* __restrict__ part, int nparts, T* __restrict__ w,
                                                       int ldw, int wc, T* __restrict__ vbuf, T* __restrict__ scal3,
                                                       T* __restrict__ uu) {
  extern __shared__ char smem_raw[];
  T* v = reinterpret_cast<T*>(smem_raw);  // c entries
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  T xn = 0;
  for (int q = 0; q < nparts; ++q) xn += part[q];
  const T alpha = a[idx(c - 1, c, lda)];
  T tau = 0, beta = alpha, scal = 1;
  if (xn != T(0)) {
    beta = -copysign(sqrt(alpha * alpha + xn), alpha);
    tau = (beta - alpha) / beta;
    scal = T(1) / (alpha - beta);
  }
  for (int r = t; r < c; r += blockDim.x) v[r] = (r == c - 1) ? T(1) : a[idx(r, c, lda)] * scal;
  __syncthreads();
  if (blockIdx.x == 0) {
    for (int r = t; r < c; r += blockDim.x) vbuf[r] = v[r];
    if (t == 0) {
      scal3[0] = tau;
      scal3[1] = beta;
    }
  }
  for (int u = 0; u < SYMV_COLS / 4; ++u) {
    const int j = blockIdx.x * SYMV_COLS + wid * (SYMV_COLS / 4) + u;
    if (j >= c + 2 * m) break;
    const T* col = j < c ? a + idx(0, j, lda) : j < c + m ? w + idx(0, wc + 1 + (j - c), ldw)
                                                          : a + idx(0, c + 1 + (j - c - m), lda);
    T s0 = 0, s1 = 0;
    int r = lane;
    for (; r + 64 < c; r += 128) {
      s0 += col[r] * v[r];
      s1 += col[r + 64] * v[r + 64];
    }
    if (r < c) s0 += col[r] * v[r];
    const T s = wave_sum(s0 + s1);
    if (lane == 0) {
      if (j < c) {
        w[idx(j, wc, ldw)] = s;
      } else if (j < c + m) {
        uu[j - c] = s;
      } else {
        uu[NB + (j - c - m)] = s;
        w[idx(c + 1 + (j - c - m), wc, ldw)] = s;
      }
    }
  }
}

// w(r) = tau (y(r) - A(r, c+1:nc) u - W(r, wc+1:) u2) for r < c, into W(r, wc); A(r, c) = v(r); part = partial w^T v;
// scal3[2] = w(c - 1) (the next column's trd_colupd reads it in every workgroup).
template <typename T>
__global__ __launch_bounds__(ROWS) void trd_wrow_kernel(int c, int m, T* __restrict__ a, int lda, T* __restrict__ w,
                                                        int ldw, int wc, const T* __restrict__ vbuf,
                                                        const T* __restrict__ uu, T* __restrict__ scal3,
                                                        T* __restrict__ part) {
  __shared__ T us[2 * NB], red[ROWS / 64];
  const int t = threadIdx.x, r = blockIdx.x * ROWS + t;
  if (t < 2 * NB) us[t] = (t % NB < m) ? uu[t] : T(0);
  __syncthreads();
  const T tau = scal3[0];
  T x = 0;
  if (r < c) {
    T y = w[idx(r, wc, ldw)];
    for (int q = 0; q < m; ++q) y -= a[idx(r, c + 1 + q, lda)] * us[q] + w[idx(r, wc + 1 + q, ldw)] * us[NB + q];
    y *= tau;
    w[idx(r, wc, ldw)] = y;
    if (r == c - 1) scal3[2] = y;
    const T vr = vbuf[r];
    a[idx(r, c, lda)] = vr;
    x = y * vr;
  }
  const T s = block_sum(x, red);
  if (t == 0) part[blockIdx.x] = s;
}

template <typename T>
__global__ __launch_bounds__(ROWS) void trd_walpha_kernel(int c, T* __restrict__ w, int ldw, int wc,
                                                          const T* __restrict__ vbuf, const T* __restrict__ scal3,
                                                          const T* __restrict__ part, int nparts, T* __restrict__ e,
                                                          T* __restrict__ tau) {
  T dot = 0;
  for (int q = 0; q < nparts; ++q) dot += part[q];
  const T tv = scal3[0];
  const T alpha = T(-0.5) * tv * dot;
  const int r = blockIdx.x * ROWS + threadIdx.x;
  if (r < c) w[idx(r, wc, ldw)] += alpha * vbuf[r];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    e[c - 1] = scal3[1];
    tau[c - 1] = tv;
  }
}

// After a panel (columns L ..< L + NB): A(c - 1, c) = e(c - 1), d(c) = A(c, c).
template <typename T>
__global__ void trd_panel_fix_kernel(int L, T* __restrict__ a, int lda, const T* __restrict__ e, T* __restrict__ d) {
  const int c = L + threadIdx.x;
  if (threadIdx.x < NB) {
    a[idx(c - 1, c, lda)] = e[c - 1];
    d[c] = a[idx(c, c, lda)];
  }
}

// dsytd2 ('U') on the leading kk x kk block (kk <= NX + 1), full storage, one workgroup in LDS: for i = kk - 1 down
// to 1 (1-based): larfg(i, A(i, i+1), A(1:i-1, i+1)); w = tau A(1:i, 1:i) v; w += -tau/2 (w^T v) v;
// A(1:i, 1:i) -= v w^T + w v^T.
template <typename T>
__global__ __launch_bounds__(256) void trd_sytd2_kernel(int kk, T* __restrict__ a, int lda, T* __restrict__ d,
                                                        T* __restrict__ e, T* __restrict__ tau) {
  __shared__ T S[NX + 1][NX + 2];
  __shared__ T v[NX + 1], wv[NX + 1], red[4];
  __shared__ T sc[3];
  const int t = threadIdx.x;
  for (int q = t; q < kk * kk; q += blockDim.x) S[q % kk][q / kk] = a[idx(q % kk, q / kk, lda)];
  __syncthreads();
  for (int i = kk - 1; i >= 1; --i) {
    // column i (0-based) above the diagonal: alpha = S(i-1, i), x = S(0:i-1, i)
    T x = (t < i - 1) ? S[t][i] : T(0);
    const T xn = block_sum(x * x, red);
    if (t == 0) {
      const T alpha = S[i - 1][i];
      T tv = 0, beta = alpha, scal = 1;
      if (xn != T(0)) {
        beta = -copysign(sqrt(alpha * alpha + xn), alpha);
        tv = (beta - alpha) / beta;
        scal = T(1) / (alpha - beta);
      }
      sc[0] = tv;
      sc[1] = beta;
      sc[2] = scal;
    }
    __syncthreads();
    const T tv = sc[0], beta = sc[1], scal = sc[2];
    if (t < i) v[t] = (t == i - 1) ? T(1) : S[t][i] * scal;
    __syncthreads();
    if (tv != T(0)) {
      if (t < i) {
        T s = 0;
        for (int r = 0; r < i; ++r) s += S[r][t] * v[r];
        wv[t] = tv * s;
      }
      __syncthreads();
      const T dot = block_sum((t < i) ? wv[t] * v[t] : T(0), red);
      const T alpha2 = T(-0.5) * tv * dot;
      if (t < i) wv[t] += alpha2 * v[t];
      __syncthreads();
      for (int q = t; q < i * i; q += blockDim.x) {
        const int r = q % i, cc = q / i;
        S[r][cc] -= v[r] * wv[cc] + wv[r] * v[cc];
      }
    }
    __syncthreads();
    if (t < i - 1 && tv != T(0)) S[t][i] = v[t];
    if (t == 0) {
      S[i - 1][i] = beta;
      e[i - 1] = beta;
      tau[i - 1] = tv;
      d[i] = S[i][i];
    }
    __syncthreads();
  }
  if (t == 0) d[0] = S[0][0];
  __syncthreads();
  for (int q = t; q < kk * kk; q += blockDim.x) a[idx(q % kk, q / kk, lda)] = S[q % kk][q / kk];
}

// Block q of the back-transformation: reflectors i0 = 32 q ..< i0 + ib, rows = i0 + ib. V (rows x ib, ld n,
// at vall + q * n * NB) explicit: V(p, cc) = A(p, i0 + cc + 1) for p < i0 + cc, 1 at p = i0 + cc, 0 below. T lower
// (ib x ib, ld NB, at tall + q * NB * NB) by larft('B', 'C'): T(cc, cc) = tau, T(cc+1:, cc) = -tau T(cc+1:, cc+1:)
// V(:, cc+1:)^T V(:, cc).
template <typename T>
__global__ __launch_bounds__(256) void ormtr_blocks_kernel(int n, const T* __restrict__ a, int lda,
                                                           const T* __restrict__ tau, T* __restrict__ vall,
                                                           T* __restrict__ tall) {
  __shared__ T G[NB][NB + 1];
  __shared__ T Ts[NB][NB + 1];
  __shared__ T Vs[64][NB + 1];
  const int q = blockIdx.x, t = threadIdx.x, k = n - 1;
  const int i0 = q * NB, ib = min(NB, k - i0), rows = i0 + ib;
  T* v = vall + (int64_t)q * n * NB;
  for (int64_t e = t; e < (int64_t)rows * ib; e += blockDim.x) {
    const int p = (int)(e % rows), cc = (int)(e / rows), r = i0 + cc;
    v[p + (int64_t)cc * n] = (p < r) ? a[idx(p, r + 1, lda)] : (p == r ? T(1) : T(0));
  }
  // Gram over rows, staged 64 at a time
  const int gr = t % NB, gc0 = t / NB;  // 8 output columns per thread group: gc0, gc0 + 8, ...
  T acc[4] = {0, 0, 0, 0};
  for (int p0 = 0; p0 < rows; p0 += 64) {
    __syncthreads();
    for (int e = t; e < 64 * NB; e += blockDim.x) {
      const int p = e % 64, cc = e / 64;
      const int pr = p0 + p, r = i0 + cc;
      Vs[p][cc] = (pr < rows && cc < ib) ? ((pr < r) ? a[idx(pr, r + 1, lda)] : (pr == r ? T(1) : T(0))) : T(0);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int gc = gc0 + 8 * u;
      T s = 0;
      for (int p = 0; p < 64; ++p) s += Vs[p][gr] * Vs[p][gc];
      acc[u] += s;
    }
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) G[gr][gc0 + 8 * u] = acc[u];
  for (int e = t; e < NB * (NB + 1); e += blockDim.x) (&Ts[0][0])[e] = T(0);
  __syncthreads();
  for (int cc = ib - 1; cc >= 0; --cc) {
    const T tc = tau[i0 + cc];
    // Ts(cc+1:ib, cc) = -tc Ts(cc+1:ib, cc+1:ib) G(cc+1:ib, cc)
    T val = 0;
    if (t > cc && t < ib && tc != T(0)) {
      for (int l = cc + 1; l <= t; ++l) val += Ts[t][l] * G[l][cc];
      val *= -tc;
    }
    __syncthreads();
    if (t > cc && t < ib) Ts[t][cc] = val;
    if (t == cc) Ts[cc][cc] = tc;
    __syncthreads();
  }
  T* tt = tall + (int64_t)q * NB * NB;
  for (int e = t; e < NB * NB; e += blockDim.x) tt[e] = Ts[e % NB][e / NB];
}

}  // namespace

template <typename T>
size_t sytrd_workspace(int n) {
  // W (n x NB) | vbuf (n) | uu (2 NB) | scal3 (4) | norm partials, 16, w^T v partials (cdiv(n, ROWS) each)
  return (size_t)n * NB + n + 2 * NB + 4 + 2 * (size_t)cdiv(n, SYMV_COLS) + 2 * (size_t)cdiv(n, ROWS) + 256;
}

template <typename T>
void sytrd_upper(hipStream_t s, int n, T* a, int lda, T* d, T* e, T* tau, T* work, bool full_input) {
  if (n <= 0) return;
  if (!full_input) symmetrize<T>(s, 'U', n, a, lda);
  T* w = work;
  const int ldw = n;
  T* vbuf = w + (size_t)n * NB;
  T* uu = vbuf + n;
  T* scal3 = uu + 2 * NB;
  T* partn = scal3 + 4;                       // norm partials of the larfg vector
  T* partw = partn + cdiv(n, ROWS) + 16;      // partials of w^T v
  int kk = n;
  if (n > NX) {
    kk = n - ((n - NX + NB - 1) / NB) * NB;
    for (int i1 = n - NB + 1; i1 >= kk + 1; i1 -= NB) {
      const int nc = i1 + NB - 1;  // leading block of the panel
      const int L = nc - NB;       // columns L ..< nc, leading block of the update L x L
      int npw = 0;                 // partials of the previous column's w^T v
      for (int c = nc - 1; c >= L; --c) {
        const int wc = c - L, m = nc - 1 - c;
        const int nrow = cdiv(c + 1, ROWS);
        hipLaunchKernelGGL(trd_colupd_kernel<T>, dim3(nrow), dim3(ROWS), 0, s, c, m, a, lda, w, ldw, wc, partn,
                           c < nc - 1 ? 1 : 0, vbuf, scal3, partw, npw, e, tau);
        XTB_LAUNCH_CHECK();
        // c >= L >= 1 here (kk >= 1), so the column has a reflector
        const size_t lds = sizeof(T) * (size_t)c;
        allow_lds(trd_symv_kernel<T>, lds);
        hipLaunchKernelGGL(trd_symv_kernel<T>, dim3(cdiv(c + 2 * m, SYMV_COLS)), dim3(256), lds, s, c, m, a, lda,
                           partn, nrow, w, ldw, wc, vbuf, scal3, uu);
        XTB_LAUNCH_CHECK();
        const int nr2 = cdiv(c, ROWS);
        hipLaunchKernelGGL(trd_wrow_kernel<T>, dim3(nr2), dim3(ROWS), 0, s, c, m, a, lda, w, ldw, wc, vbuf, uu, scal3,
                           partw);
        XTB_LAUNCH_CHECK();
        npw = nr2;
      }
      // the panel's last column: its w update before the trailing update reads W
      hipLaunchKernelGGL(trd_walpha_kernel<T>, dim3(cdiv(L, ROWS)), dim3(ROWS), 0, s, L, w, ldw, 0, vbuf, scal3, partw,
                         npw, e, tau);
      XTB_LAUNCH_CHECK();
      // A(0:L, 0:L) -= V W^T + W V^T, V = A(0:L, L:L+NB), both triangles
      gemm<T>(s, 'N', 'T', L, L, NB, T(-1), a + idx(0, L, lda), lda, w, ldw, T(1), a, lda);
      gemm<T>(s, 'N', 'T', L, L, NB, T(-1), w, ldw, a + idx(0, L, lda), lda, T(1), a, lda);
      hipLaunchKernelGGL(trd_panel_fix_kernel<T>, dim3(1), dim3(64), 0, s, L, a, lda, e, d);
      XTB_LAUNCH_CHECK();
    }
  }
  hipLaunchKernelGGL(trd_sytd2_kernel<T>, dim3(1), dim3(256), 0, s, kk, a, lda, d, e, tau);
  XTB_LAUNCH_CHECK();
}

template <typename T>
size_t ormtr_workspace(int n, int m) {
  const int nblk = cdiv(std::max(n - 1, 1), NB);
  return (size_t)nblk * n * NB + (size_t)nblk * NB * NB + 2 * (size_t)NB * std::max(m, 1) + 256;
}

template <typename T>
void ormtr_upper(hipStream_t s, int n, int m, const T* a, int lda, const T* tau, T* z, int ldz, T* work) {
  const int k = n - 1;
  if (m <= 0 || k <= 0) return;
  const int nblk = cdiv(k, NB);
  T* vall = work;
  T* tall = vall + (size_t)nblk * n * NB;
  T* w1 = tall + (size_t)nblk * NB * NB;
  T* w2 = w1 + (size_t)NB * m;
  hipLaunchKernelGGL(ormtr_blocks_kernel<T>, dim3(nblk), dim3(256), 0, s, n, a, lda, tau, vall, tall);
  XTB_LAUNCH_CHECK();
  for (int q = 0; q < nblk; ++q) {
    const int i0 = q * NB, ib = std::min(NB, k - i0), rows = i0 + ib;
    const T* v = vall + (size_t)q * n * NB;
    const T* tt = tall + (size_t)q * NB * NB;
    gemm<T>(s, 'T', 'N', ib, m, rows, T(1), v, n, z, ldz, T(0), w1, ib);
    gemm<T>(s, 'N', 'N', ib, m, ib, T(1), tt, NB, w1, ib, T(0), w2, ib);
    gemm<T>(s, 'N', 'N', rows, m, ib, T(-1), v, n, w2, ib, T(1), z, ldz);
  }
}

template size_t sytrd_workspace<double>(int);
template size_t sytrd_workspace<float>(int);
template void sytrd_upper<double>(hipStream_t, int, double*, int, double*, double*, double*, double*, bool);
template void sytrd_upper<float>(hipStream_t, int, float*, int, float*, float*, float*, float*, bool);
template size_t ormtr_workspace<double>(int, int);
template size_t ormtr_workspace<float>(int, int);
template void ormtr_upper<double>(hipStream_t, int, int, const double*, int, const double*, double*, int, double*);
template void ormtr_upper<float>(hipStream_t, int, int, const float*, int, const float*, float*, int, float*);

}  // namespace xtbeig
