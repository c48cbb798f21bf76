// Stage 1 of the two-stage reduction: dense symmetric -> lower band of width b (the reference's reduceToBand with the
// combined sy2sb update, Stages/BandForm.swift:28-142); the back-transformation through its panels is in backtransform.hip.
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
// After the last panel the lower band is copied into the chase's band storage (element (i, j), 0 <= i - j <= 2b - 1,
// at i + (2b - 1) j; see chase.hip).
#include "common.hpp"
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
                                                              T* __restrict__ tau) {
  __shared__ T red[4 * SB];
  __shared__ T tot[SB];
  __shared__ T sh_alpha;
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
        T v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int i = t + r * QR_THREADS;
          v[r] = (i == col) ? T(1) : ((i > col && i < m) ? x[r][c] * scal : T(0));
          if (i == col) x[r][c] = beta;
          else if (i > col) x[r][c] = 0;
        }
        if (t == 0) tau[col] = tq;
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
              const T f = tq * tot[cc];
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

// The per-panel reductions over the m rows run on cdiv(m, PR) workgroups, PR rows each; k x k partial products per
// row block go to a partial buffer (ld 32, 1024 elements per block) and are summed by the consumer in block order.
constexpr int PR = 64;

// Rows i0 ..< i0 + PR of the m x k (ld m) matrix a into sh (PR x 33), zeros past m and k.
template <typename T>
__device__ inline void stage_rows(int m, int k, int i0, const T* __restrict__ a, T (*sh)[33]) {
  for (int e = threadIdx.x; e < PR * 32; e += blockDim.x) {
    const int i = e % PR, c = e / PR;
    sh[i][c] = (i0 + i < m && c < k) ? a[(int64_t)(i0 + i) + (int64_t)c * m] : T(0);
  }
}

// out[r + 32 c] = sum_i A(i, r) B(i, c) over the staged rows (k x k; sym: r <= c only).
template <typename T>
__device__ inline void rows_atb(int k, const T (*A)[33], const T (*B)[33], T* __restrict__ out, bool sym) {
#pragma unroll
  for (int u = 0; u < 4; ++u) {
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
        out[r + 32 * c] = s0 + s1;
      }
    }
  }
}

// k x k (ld 32) sum of the nblk partials into sh (32 x 33).
template <typename T>
__device__ inline void sum_partials(int k, int nblk, const T* __restrict__ part, T (*sh)[33], bool sym) {
  for (int o = threadIdx.x; o < k * k; o += blockDim.x) {
    const int r = o % k, c = o / k;
    if (sym && r > c) continue;
    T sum = 0;
    for (int q = 0; q < nblk; ++q) sum += part[(int64_t)q * 1024 + r + 32 * c];
    sh[r][c] = sum;
  }
}

// Gp[blk] = V(rows)^T V(rows), upper triangle.
template <typename T>
__global__ __launch_bounds__(256) void panel_gram_kernel(int m, int k, const T* __restrict__ v, T* __restrict__ gp) {
  __shared__ T Vs[PR][33];
  stage_rows<T>(m, k, blockIdx.x * PR, v, Vs);
  __syncthreads();
  rows_atb<T>(k, Vs, Vs, gp + (int64_t)blockIdx.x * 1024, true);
}

// G = sum of the Gram partials; the forward columnwise T (I - V T V^T = H_1 ... H_k; T(0:c, c) = -tau_c T(0:c, 0:c)
// G(0:c, c), T(c, c) = tau_c), written by workgroup 0 to tt (ld ldt); VT(rows) = V(rows) T.
template <typename T>
__global__ __launch_bounds__(256) void panel_tvt_kernel(int m, int k, const T* __restrict__ v,
                                                        const T* __restrict__ tau, const T* __restrict__ gp, int nblk,
                                                        T* __restrict__ tt, int ldt, T* __restrict__ vt) {
  __shared__ T G[32][33];
  __shared__ T Ts[32][33];
  __shared__ T Vs[PR][33];
  const int t = threadIdx.x, i0 = blockIdx.x * PR;
  sum_partials<T>(k, nblk, gp, G, true);
  stage_rows<T>(m, k, i0, v, Vs);
  for (int e = t; e < 32 * 33; e += blockDim.x) (&Ts[0][0])[e] = T(0);
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
                                                       T* __restrict__ yp) {
  __shared__ T Xs[PR][33];
  __shared__ T Ws[PR][33];
  const int t = threadIdx.x, i0 = blockIdx.x * PR;
  const int64_t mk = (int64_t)m * k;
  for (int e = t; e < PR * 32; e += blockDim.x) {
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
  rows_atb<T>(k, Ws, Xs, yp + (int64_t)blockIdx.x * 1024, false);
}

// Y = sum of the Yp partials; W(rows) = X(rows) - V(rows) Y / 2, in place of x.
template <typename T>
__global__ __launch_bounds__(256) void panel_w_kernel(int m, int k, const T* __restrict__ v, const T* __restrict__ yp,
                                                      int nblk, T* __restrict__ x) {
  __shared__ T Y[32][33];
  __shared__ T Vs[PR][33];
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
__global__ void extract_band_kernel(int n, int b, const T* __restrict__ a, int lda, T* __restrict__ ab) {
  const int j = blockIdx.x, ld = band_ld(b);
  for (int dd = threadIdx.x; dd <= ld; dd += blockDim.x) {
    const int i = j + dd;
    if (i < n) ab[(int64_t)i + (int64_t)ld * j] = (dd <= b) ? a[idx(i, j, lda)] : T(0);
  }
}

template <typename T, int SB, int R>
void launch_qr(hipStream_t s, int m, int nb, T* p, int lda, T* v1, int ldv1, T* buf, T* tau) {
  hipLaunchKernelGGL((panel_qr_kernel<T, SB, R>), dim3(1), dim3(QR_THREADS), 0, s, m, nb, p, lda, v1, ldv1, buf, tau);
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
  // tau (b) | buf [V W V] (3 n b) | VT (n b) | split-K partials of X (ceil(n / SPLIT_ROWS) n b) | k x k partials
  // of the row blocks (ceil(n / PR) 1024)
  return (size_t)b + 4 * (size_t)n * b + (size_t)((n + SPLIT_ROWS - 1) / SPLIT_ROWS) * n * b +
         (size_t)((n + PR - 1) / PR) * 1024 + 64;
}

template <typename T>
void reduce_to_band(hipStream_t s, int n, int b, T* a, int lda, T* v1, T* t1, T* ab, T* work) {
  if (b > 32 || b < 1) throw std::runtime_error("reduce_to_band: band width must be in [1, 32]");
  T* tau = work;
  T* buf = tau + b;
  T* vt = buf + 3 * (size_t)n * b;
  T* xpart = vt + (size_t)n * b;
  T* kpart = xpart + (size_t)((n + SPLIT_ROWS - 1) / SPLIT_ROWS) * n * b;
  // v1 is zero outside the panels' V, so that consecutive panels read as one block [V_a V_b ...] from row r_a
  set_zero<T>(s, n, n, v1, n);
  for (int j = 0; j < n - b - 1; j += b) {
    const int r = j + b, m = n - r, k = std::min(b, m);
    T* panel = a + idx(r, j, lda);
    T* v = v1 + idx(r, j, n);
    launch_panel_qr<T>(s, m, b, panel, lda, v, n, buf, tau);
    T* tt = t1 + idx(0, j, b);
    T* a2 = a + idx(r, r, lda);
    T* x = buf + (size_t)m * k;  // slot 1 of [V | W | V]
    const int nblk = cdiv(m, PR);
    hipLaunchKernelGGL(panel_gram_kernel<T>, dim3(nblk), dim3(256), 0, s, m, k, buf, kpart);      // V^T V parts
    XTB_LAUNCH_CHECK();
    hipLaunchKernelGGL(panel_tvt_kernel<T>, dim3(nblk), dim3(256), 0, s, m, k, buf, tau, kpart, nblk, tt, b,
                       vt);  // T, VT = V T
    XTB_LAUNCH_CHECK();
    const int nsplit = gemm_splitk<T>(s, m, k, m, a2, lda, vt, m, xpart, SPLIT_ROWS);  // A2 VT in parts
    hipLaunchKernelGGL(panel_xy_kernel<T>, dim3(nblk), dim3(256), 0, s, m, k, vt, xpart, nsplit, x, kpart);  // X, Y
    XTB_LAUNCH_CHECK();
    hipLaunchKernelGGL(panel_w_kernel<T>, dim3(nblk), dim3(256), 0, s, m, k, buf, kpart, nblk, x);  // W
    XTB_LAUNCH_CHECK();
    gemm<T>(s, 'N', 'T', m, m, 2 * k, T(-1), buf, m, buf + (size_t)m * k, m, T(1), a2, lda);  // A2 -= [V W][W V]^T
  }
  hipLaunchKernelGGL(extract_band_kernel<T>, dim3(n), dim3(128), 0, s, n, b, a, lda, ab);
  XTB_LAUNCH_CHECK();
}

template size_t band_workspace<double>(int, int);
template size_t band_workspace<float>(int, int);
template void reduce_to_band<double>(hipStream_t, int, int, double*, int, double*, double*, double*, double*);
template void reduce_to_band<float>(hipStream_t, int, int, float*, int, float*, float*, float*, float*);
}  // namespace xtbeig
