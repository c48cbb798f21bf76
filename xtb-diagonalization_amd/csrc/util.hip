// Elementwise matrix kernels: copies with precision conversion, symmetrization, identity, and the 32 x 32-tiled
// upper -> full copy behind xtb_dsymmetrize_pad (Exports.swift:117-138 of the reference).
#include "common.hpp"
#include "kernels.hpp"

namespace xtbeig {

namespace {

template <typename S, typename D>
__global__ void copy_kernel(int m, int n, const S* __restrict__ src, int lds, D* __restrict__ dst, int ldd, Zb zb) {
  src = zb.at(src);
  dst = zb.at(dst);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  for (int j = blockIdx.y; j < n; j += gridDim.y)
    if (i < m) dst[idx(i, j, ldd)] = D(src[idx(i, j, lds)]);
}

// dst(i, j) = src(i, j) for i <= j (the upper triangle); dst's strictly lower triangle is not touched.
template <typename T>
__global__ void copy_upper_kernel(int n, const T* __restrict__ src, int lds, T* __restrict__ dst, int ldd, Zb zb) {
  src = zb.at(src);
  dst = zb.at(dst);
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i <= j && i < n) dst[idx(i, j, ldd)] = src[idx(i, j, lds)];
}

// 32 x 32 tiles through LDS: tile (bi, bj) with bi <= bj reads the referenced triangle and writes both.
template <typename T>
__global__ void symmetrize_kernel(int n, T* a, int lda, bool upper, Zb zb) {
  __shared__ T tile[32][33];
  const int bi = blockIdx.x, bj = blockIdx.y;
  if (bi > bj) return;
  a = zb.at(a);
  const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
  // Source tile: upper -> rows bi, cols bj; lower -> rows bj, cols bi (transposed role).
  for (int r = ty; r < 32; r += 8) {
    const int i = (upper ? bi : bj) * 32 + tx, j = (upper ? bj : bi) * 32 + r;
    if (i < n && j < n) tile[r][tx] = a[idx(i, j, lda)];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    // write the mirror: element (i, j) of the source goes to (j, i)
    const int i = (upper ? bj : bi) * 32 + tx, j = (upper ? bi : bj) * 32 + r;
    // dst (i, j) = src (j, i) = tile[i - ...][j - ...]
    if (i < n && j < n) {
      const bool ref = upper ? (j <= i) : (j >= i);  // (j, i) lies in the referenced triangle
      if (ref) a[idx(i, j, lda)] = tile[tx][r];
    }
  }
}

// max |a(i, j)| into *amax by atomicMax on the bit pattern of |a(i, j)| (for non-negative doubles the patterns
// order as the values; a NaN lies above every other).
__global__ __launch_bounds__(256) void absmax_kernel(int m, int n, const double* __restrict__ a, int lda,
                                                     unsigned long long* amax, Zb zb) {
  a = zb.at(a);
  amax = zb.at(amax);
  __shared__ unsigned long long red[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long v = 0;
  for (int j = blockIdx.y; j < n; j += gridDim.y)
    if (i < m) v = max(v, (unsigned long long)__double_as_longlong(fabs(a[idx(i, j, lda)])));
  for (int off = 32; off > 0; off >>= 1) v = max(v, (unsigned long long)__shfl_xor((long long)v, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    v = max(max(red[0], red[1]), max(red[2], red[3]));
    if (v) atomicMax(amax, v);
  }
}

// 2^-e for the FP32 copy (1 while max |src| lies in [2^-60, 2^60], or is zero or not finite)
__device__ inline double fp32_scale(unsigned long long bits) {
  const double m = __longlong_as_double((long long)bits);
  if (!(m > 0.0) || !isfinite(m) || (m >= 0x1p-60 && m <= 0x1p60)) return 1.0;
  int e;
  (void)frexp(m, &e);  // m = f 2^e, f in [0.5, 1)
  return ldexp(1.0, min(1000, max(-1000, -e)));  // a normal double either way (subnormal m: e down to -1073)
}

__global__ void copy_ranged_kernel(int m, int n, const double* __restrict__ src, int lds, float* __restrict__ dst,
                                   int ldd, const unsigned long long* amax, int* status, Zb zb) {
  src = zb.at(src);
  dst = zb.at(dst);
  amax = zb.at(amax);
  status = zb.at(status);
  const unsigned long long bits = *amax;
  const bool finite = isfinite(__longlong_as_double((long long)bits));
  const double sc = fp32_scale(bits);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (!finite && i == 0 && blockIdx.y == 0) atomicExch(status, 1);
  for (int j = blockIdx.y; j < n; j += gridDim.y)
    if (i < m) dst[idx(i, j, ldd)] = finite ? float(sc * src[idx(i, j, lds)]) : 0.0f;
}

__global__ void unscale_values_kernel(int n, double* w, const unsigned long long* amax, Zb zb) {
  w = zb.at(w);
  amax = zb.at(amax);
  const double sc = fp32_scale(*amax);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && sc != 1.0) w[i] /= sc;  // a power of two: exact
}

template <typename T>
__global__ void fill_kernel(int m, int n, T* a, int lda, T diag, Zb zb) {
  a = zb.at(a);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  for (int j = blockIdx.y; j < n; j += gridDim.y)
    if (i < m) a[idx(i, j, lda)] = (i == j) ? diag : T(0);
}

__global__ void symmetrize_pad_kernel(int n, const double* __restrict__ src, int lds, double* __restrict__ dst,
                                      int ldd) {
  __shared__ double tile[32][33];
  const int bi = blockIdx.x, bj = blockIdx.y;
  if (bi > bj) return;
  const int tx = threadIdx.x, ty = threadIdx.y;
  for (int r = ty; r < 32; r += 8) {
    const int i = bi * 32 + tx, j = bj * 32 + r;
    if (i < n && j < n && i <= j) {
      const double v = src[idx(i, j, lds)];
      tile[r][tx] = v;
      dst[idx(i, j, ldd)] = v;
    }
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int i = bj * 32 + tx, j = bi * 32 + r;  // (i, j) in the lower triangle, mirror of (j, i)
    if (i < n && j < n && j <= i) dst[idx(i, j, ldd)] = tile[tx][r];
  }
}

// dst(j, i) = src(i, j) through 32 x 32 LDS tiles; src m x n, dst n x m.
template <typename S, typename D>
__global__ void transpose_kernel(int m, int n, const S* __restrict__ src, int lds, D* __restrict__ dst, int ldd,
                                 Zb zb) {
  __shared__ D tile[32][33];
  src = zb.at(src);
  dst = zb.at(dst);
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
  for (int r = ty; r < 32; r += 8) {
    const int i = i0 + tx, j = j0 + r;
    if (i < m && j < n) tile[r][tx] = D(src[idx(i, j, lds)]);
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int j = j0 + tx, i = i0 + r;  // dst row j, column i
    if (i < m && j < n) dst[idx(j, i, ldd)] = tile[tx][r];
  }
}

}  // namespace

template <typename S, typename D>
void transpose_matrix(hipStream_t s, int m, int n, const S* src, int lds, D* dst, int ldd) {
  if (m <= 0 || n <= 0) return;
  if (cdiv(n, 32) > 65535) throw std::runtime_error("transpose: matrix too wide");
  hipLaunchKernelGGL((transpose_kernel<S, D>), dim3(cdiv(m, 32), cdiv(n, 32), zgrid()), dim3(32, 8), 0, s, m, n, src,
                     lds, dst, ldd, zb());
  XTB_LAUNCH_CHECK();
}

template <typename S, typename D>
void copy_matrix(hipStream_t s, int m, int n, const S* src, int lds, D* dst, int ldd) {
  if (m <= 0 || n <= 0) return;
  hipLaunchKernelGGL((copy_kernel<S, D>), dim3(cdiv(m, 256), n < 65535 ? n : 65535, zgrid()), dim3(256), 0, s, m, n,
                     src, lds, dst, ldd, zb());
  XTB_LAUNCH_CHECK();
}

void copy_to_fp32_ranged(hipStream_t s, int n, const double* src, int lds, float* dst, int ldd,
                         unsigned long long* amax, int* status) {
  if (n <= 0) return;
  set_zero<double>(s, 1, 1, reinterpret_cast<double*>(amax), 1);
  hipLaunchKernelGGL(absmax_kernel, dim3(cdiv(n, 256), std::min(n, 64), zgrid()), dim3(256), 0, s, n, n, src, lds,
                     amax, zb());
  XTB_LAUNCH_CHECK();
  hipLaunchKernelGGL(copy_ranged_kernel, dim3(cdiv(n, 256), n < 65535 ? n : 65535, zgrid()), dim3(256), 0, s, n, n,
                     src, lds, dst, ldd, amax, status, zb());
  XTB_LAUNCH_CHECK();
}

void unscale_values(hipStream_t s, int n, double* w, const unsigned long long* amax) {
  if (n <= 0) return;
  hipLaunchKernelGGL(unscale_values_kernel, dim3(cdiv(n, 256), 1, zgrid()), dim3(256), 0, s, n, w, amax, zb());
  XTB_LAUNCH_CHECK();
}

template <typename T>
void copy_upper(hipStream_t s, int n, const T* src, int lds, T* dst, int ldd) {
  if (n <= 0) return;
  hipLaunchKernelGGL(copy_upper_kernel<T>, dim3(cdiv(n, 256), n, zgrid()), dim3(256), 0, s, n, src, lds, dst, ldd,
                     zb());
  XTB_LAUNCH_CHECK();
}

template <typename T>
void symmetrize(hipStream_t s, char uplo, int n, T* a, int lda) {
  if (n <= 1) return;
  const int nt = cdiv(n, 32);
  hipLaunchKernelGGL(symmetrize_kernel<T>, dim3(nt, nt, zgrid()), dim3(32, 8), 0, s, n, a, lda,
                     uplo == 'U' || uplo == 'u', zb());
  XTB_LAUNCH_CHECK();
}

template <typename T>
void set_zero(hipStream_t s, int m, int n, T* a, int lda) {
  if (m <= 0 || n <= 0) return;
  hipLaunchKernelGGL(fill_kernel<T>, dim3(cdiv(m, 256), n < 65535 ? n : 65535, zgrid()), dim3(256), 0, s, m, n, a,
                     lda, T(0), zb());
  XTB_LAUNCH_CHECK();
}

template <typename T>
void set_identity(hipStream_t s, int n, T* a, int lda) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_kernel<T>, dim3(cdiv(n, 256), n < 65535 ? n : 65535, zgrid()), dim3(256), 0, s, n, n, a,
                     lda, T(1), zb());
  XTB_LAUNCH_CHECK();
}

void symmetrize_pad(hipStream_t s, int n, const double* src, int lds, double* dst, int ldd) {
  if (n <= 0) return;
  const int nt = cdiv(n, 32);
  hipLaunchKernelGGL(symmetrize_pad_kernel, dim3(nt, nt), dim3(32, 8), 0, s, n, src, lds, dst, ldd);
  XTB_LAUNCH_CHECK();
}

template void transpose_matrix<double, double>(hipStream_t, int, int, const double*, int, double*, int);
template void transpose_matrix<double, float>(hipStream_t, int, int, const double*, int, float*, int);
template void transpose_matrix<float, float>(hipStream_t, int, int, const float*, int, float*, int);
template void copy_matrix<double, double>(hipStream_t, int, int, const double*, int, double*, int);
template void copy_matrix<double, float>(hipStream_t, int, int, const double*, int, float*, int);
template void copy_matrix<float, double>(hipStream_t, int, int, const float*, int, double*, int);
template void copy_matrix<float, float>(hipStream_t, int, int, const float*, int, float*, int);
template void symmetrize<double>(hipStream_t, char, int, double*, int);
template void copy_upper<double>(hipStream_t, int, const double*, int, double*, int);
template void symmetrize<float>(hipStream_t, char, int, float*, int);
template void set_zero<double>(hipStream_t, int, int, double*, int);
template void set_zero<float>(hipStream_t, int, int, float*, int);
template void set_identity<double>(hipStream_t, int, double*, int);
template void set_identity<float>(hipStream_t, int, float*, int);

}  // namespace xtbeig
