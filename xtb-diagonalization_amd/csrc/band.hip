// Stage 1 of the two-stage reduction: dense symmetric -> lower band of width b (the reference's reduceToBand with the
// combined sy2sb update, Stages/BandForm.swift:28-142), and its back-transformation (BandFormBackTransformation.swift
// :10-21).
//
// Per panel j (columns j ..< j + b, rows r = j + b ..< n, m = n - r, k = min(b, m)):
//   1. panel_qr_kernel: Householder QR of the m x b panel by one workgroup of 1024 threads. The panel is processed
//      in sub-panels of SB columns held in registers (R rows per thread), earlier reflectors applied left-looking.
//      Writes R into the band (zeros below), V explicitly (unit diagonal, zeros above) into the back-transform store
//      v1 and into the update buffer [V | . | V], and tau.
//   2. G = V^T V (GEMM), then larft_kernel builds the forward columnwise T (I - V T V^T = H_1 ... H_k).
//   3. VT = V T; X = A2 VT; Y = VT^T X; W = X - 1/2 V Y (GEMMs); A2 -= [V W] [W V]^T, one rank-2k GEMM on the
//      matrix cores (full storage: both triangles of the trailing block stay current).
// After the last panel the lower band is copied into the chase's band storage (element (i, j), 0 <= i - j <= 2b - 1,
// at i + (2b - 1) j; see chase.hip).
#include "common.hpp"
#include "kernels.hpp"

namespace xtbeig {

namespace {

constexpr int QR_THREADS = 1024;

// Block-wide sums of NV values per thread; the totals land in tot[0 .. NV). red: 16 * NV scratch.
template <typename T, int NV>
__device__ inline void block_sums(const T (&v)[NV], T* red, T* tot) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    const T s = wave_sum(v[c]);
    if (lane == 0) red[wid * NV + c] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    T s = 0;
    for (int w = 0; w < nw; ++w) s += red[w * NV + threadIdx.x];
    tot[threadIdx.x] = s;
  }
  __syncthreads();
}

// Householder QR of the panel p (m x nb, ld lda; nb <= b). R overwrites the upper triangle of p's first min(m, nb)
// rows, zeros below. V (explicit) goes to v1 (ld ldv1) and to buf (ld m) at column offsets 0 and 2k.
template <typename T, int SB, int R>
__global__ __launch_bounds__(QR_THREADS) void panel_qr_kernel(int m, int nb, T* __restrict__ p, int lda,
                                                              T* __restrict__ v1, int ldv1, T* __restrict__ buf,
                                                              T* __restrict__ tau) {
  __shared__ T red[16 * SB];
  __shared__ T tot[SB];
  __shared__ T sh_alpha;
  const int t = threadIdx.x;
  const int k = min(m, nb);
  T x[R][SB];
  for (int c0 = 0; c0 < nb; c0 += SB) {
    const int sb = min(SB, nb - c0);
    // load the sub-panel (all nb - c0 columns up to SB; columns >= k still receive the reflectors)
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int c = 0; c < SB; ++c) {
        const int i = t + r * QR_THREADS;
        x[r][c] = (i < m && c < sb) ? p[idx(i, c0 + c, lda)] : T(0);
      }
    // apply reflectors 0 .. c0 - 1 (stored in v1) left-looking
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
    // factor the sub-panel column by column
#pragma unroll
    for (int c = 0; c < SB; ++c) {
      const int col = c0 + c;
      if (col < k) {
        // alpha = x(col, c); xnorm^2 = sum over rows > col
        T part1[1];
        part1[0] = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int i = t + r * QR_THREADS;
          if (i == col) sh_alpha = x[r][c];
          if (i > col && i < m) part1[0] += x[r][c] * x[r][c];
        }
        block_sums<T, 1>(part1, red, tot);
        const T alpha = sh_alpha, xn2 = tot[0];
        T beta, tq, scal;
        if (xn2 == T(0)) {
          tq = 0; beta = alpha; scal = 0;
        } else {
          beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
          tq = (beta - alpha) / beta;
          scal = T(1) / (alpha - beta);
        }
        // v in column c: 1 at col, scaled below, R entries above kept
        T v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int i = t + r * QR_THREADS;
          v[r] = (i == col) ? T(1) : ((i > col && i < m) ? x[r][c] * scal : T(0));
          if (i == col) x[r][c] = beta;
          else if (i > col) x[r][c] = 0;  // R below the diagonal
        }
        if (t == 0) tau[col] = tq;
        // store V column explicitly
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int i = t + r * QR_THREADS;
          if (i < m) {
            v1[idx(i, col, ldv1)] = v[r];
            buf[idx(i, col, m)] = v[r];
            buf[idx(i, 2 * k + col, m)] = v[r];
          }
        }
        // apply H = I - tau v v^T to the remaining sub-panel columns
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
            const T f = tq * tot[cc];
#pragma unroll
            for (int r = 0; r < R; ++r) x[r][cc] -= v[r] * f;
          }
        }
      }
    }
    // write back the sub-panel (R part and zeros; columns beyond k keep their reflected values)
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int c = 0; c < SB; ++c) {
        const int i = t + r * QR_THREADS;
        if (i < m && c < sb) p[idx(i, c0 + c, lda)] = x[r][c];
      }
    __syncthreads();  // v1 columns of this sub-panel visible to the next one
  }
}

// T (k x k, upper, ld ldt) of H_1 ... H_k = I - V T V^T from G = V^T V (k x k, ld k) and tau.
template <typename T>
__global__ void larft_kernel(int k, const T* __restrict__ g, const T* __restrict__ tau, T* __restrict__ tt,
                             int ldt) {
  __shared__ T Ts[64][65];
  __shared__ T col[64];
  const int t = threadIdx.x;  // 64 threads
  for (int i = 0; i < k; ++i) {
    // col[0:i] = -tau_i G(0:i, i); then T(0:i, i) = T(0:i, 0:i) col
    if (t < i) col[t] = -tau[i] * g[idx(t, i, k)];
    __syncthreads();
    if (t < i) {
      T s = 0;
      for (int l = t; l < i; ++l) s += Ts[t][l] * col[l];
      Ts[t][i] = s;
    }
    if (t == 0) Ts[i][i] = tau[i];
    __syncthreads();
  }
  for (int j = 0; j < k; ++j)
    if (t < k) tt[idx(t, j, ldt)] = (t <= j) ? Ts[t][j] : T(0);
}

// Lower band of a (diagonals 0 .. b) into band storage ab, diagonals b + 1 .. 2b - 1 zero.
template <typename T>
__global__ void extract_band_kernel(int n, int b, const T* __restrict__ a, int lda, T* __restrict__ ab) {
  const int j = blockIdx.x, ld = band_ld(b);
  for (int dd = threadIdx.x; dd <= ld; dd += blockDim.x) {
    const int i = j + dd;
    if (i < n) ab[(int64_t)i + (int64_t)ld * j] = (dd <= b) ? a[idx(i, j, lda)] : T(0);
  }
}

template <typename T>
void launch_panel_qr(hipStream_t s, int m, int nb, T* p, int lda, T* v1, int ldv1, T* buf, T* tau) {
  constexpr int SB = sizeof(T) == 8 ? 16 : 32;
  if (m <= QR_THREADS)
    hipLaunchKernelGGL((panel_qr_kernel<T, SB, 1>), dim3(1), dim3(QR_THREADS), 0, s, m, nb, p, lda, v1, ldv1, buf,
                       tau);
  else if (m <= 2 * QR_THREADS)
    hipLaunchKernelGGL((panel_qr_kernel<T, SB, 2>), dim3(1), dim3(QR_THREADS), 0, s, m, nb, p, lda, v1, ldv1, buf,
                       tau);
  else if (m <= 4 * QR_THREADS)
    hipLaunchKernelGGL((panel_qr_kernel<T, SB, 4>), dim3(1), dim3(QR_THREADS), 0, s, m, nb, p, lda, v1, ldv1, buf,
                       tau);
  else
    throw std::runtime_error("reduce_to_band: n > 4096 + b is not supported");
  XTB_LAUNCH_CHECK();
}

}  // namespace

template <typename T>
size_t band_workspace(int n, int b) {
  // tau (b) | buf [V W V] (3 n b) | G (b b) | Y (b b) | VT (n b)
  return (size_t)b + 3 * (size_t)n * b + 2 * (size_t)b * b + (size_t)n * b + 64;
}

template <typename T>
void reduce_to_band(hipStream_t s, int n, int b, T* a, int lda, T* v1, T* t1, T* ab, T* work) {
  if (b > 64 || b < 1) throw std::runtime_error("reduce_to_band: band width must be in [1, 64]");
  T* tau = work;
  T* buf = tau + b;
  T* g = buf + 3 * (size_t)n * b;
  T* y = g + (size_t)b * b;
  T* vt = y + (size_t)b * b;
  for (int j = 0; j < n - b - 1; j += b) {
    const int r = j + b, m = n - r, k = std::min(b, m);
    T* panel = a + idx(r, j, lda);
    T* v = v1 + idx(r, j, n);
    launch_panel_qr<T>(s, m, b, panel, lda, v, n, buf, tau);
    // T factor
    gemm<T>(s, 'T', 'N', k, k, m, T(1), buf, m, buf, m, T(0), g, k);
    hipLaunchKernelGGL(larft_kernel<T>, dim3(1), dim3(64), 0, s, k, g, tau, t1 + idx(0, j, b), b);
    XTB_LAUNCH_CHECK();
    T* tt = t1 + idx(0, j, b);
    T* a2 = a + idx(r, r, lda);
    T* x = buf + (size_t)m * k;  // slot 1 of [V | W | V]
    gemm<T>(s, 'N', 'N', m, k, k, T(1), buf, m, tt, b, T(0), vt, m);            // VT = V T
    gemm<T>(s, 'N', 'N', m, k, m, T(1), a2, lda, vt, m, T(0), x, m);            // X = A2 VT
    gemm<T>(s, 'T', 'N', k, k, m, T(1), vt, m, x, m, T(0), y, k);               // Y = VT^T X
    gemm<T>(s, 'N', 'N', m, k, k, T(-0.5), buf, m, y, k, T(1), x, m);           // W = X - V Y / 2
    gemm<T>(s, 'N', 'T', m, m, 2 * k, T(-1), buf, m, buf + (size_t)m * k, m, T(1), a2, lda);  // A2 -= [V W][W V]^T
  }
  hipLaunchKernelGGL(extract_band_kernel<T>, dim3(n), dim3(128), 0, s, n, b, a, lda, ab);
  XTB_LAUNCH_CHECK();
}

template <typename T>
void back_transform_band(hipStream_t s, int n, int b, const T* v1, const T* t1, T* z, int ldz, int m, T* work) {
  // z(r:n, :) = (I - V T V^T) z(r:n, :) per panel, last panel first. work: 2 b m.
  T* w1 = work;
  T* w2 = work + (size_t)b * m;
  int last = -1;
  for (int j = 0; j < n - b - 1; j += b) last = j;
  for (int j = last; j >= 0; j -= b) {
    const int r = j + b, rows = n - r, k = std::min(b, rows);
    const T* v = v1 + idx(r, j, n);
    gemm<T>(s, 'T', 'N', k, m, rows, T(1), v, n, z + r, ldz, T(0), w1, k);       // W1 = V^T Z
    gemm<T>(s, 'N', 'N', k, m, k, T(1), t1 + idx(0, j, b), b, w1, k, T(0), w2, k);  // W2 = T W1
    gemm<T>(s, 'N', 'N', rows, m, k, T(-1), v, n, w2, k, T(1), z + r, ldz);       // Z -= V W2
  }
}

template size_t band_workspace<double>(int, int);
template size_t band_workspace<float>(int, int);
template void reduce_to_band<double>(hipStream_t, int, int, double*, int, double*, double*, double*, double*);
template void reduce_to_band<float>(hipStream_t, int, int, float*, int, float*, float*, float*, float*);
template void back_transform_band<double>(hipStream_t, int, int, const double*, const double*, double*, int, int,
                                          double*);
template void back_transform_band<float>(hipStream_t, int, int, const float*, const float*, float*, int, int,
                                         float*);

}  // namespace xtbeig
