// Shared device/host helpers for the MI355X (gfx950) eigensolver library.
// Matrices are column-major (Fortran ABI), element (i, j) of a matrix with leading dimension ld at i + j * ld.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

namespace xtbeig {

struct HipError : std::runtime_error {
  hipError_t code;
  HipError(hipError_t c, const char* what) : std::runtime_error(what), code(c) {}
};

#define XTB_HIP(expr)                                                                            \
  do {                                                                                           \
    hipError_t _e = (expr);                                                                      \
    if (_e != hipSuccess) {                                                                      \
      char _buf[512];                                                                            \
      snprintf(_buf, sizeof(_buf), "%s:%d: %s -> %s", __FILE__, __LINE__, #expr,                \
               hipGetErrorString(_e));                                                           \
      throw ::xtbeig::HipError(_e, _buf);                                                        \
    }                                                                                            \
  } while (0)

#define XTB_LAUNCH_CHECK() XTB_HIP(hipGetLastError())

// Dynamic LDS beyond 64 KiB (gfx950 has 160 KiB per workgroup): raise the kernel's limit before launching.
template <typename K>
inline void allow_lds(K kernel, size_t bytes) {
  if (bytes > 64 * 1024)
    XTB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)bytes));
}

__host__ __device__ inline int64_t idx(int64_t i, int64_t j, int64_t ld) { return i + j * ld; }
__host__ __device__ inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// ---- wavefront (64-lane) reductions -------------------------------------------------------------
template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
template <typename T>
__device__ inline T wave_max(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Block-wide sum for blockDim.x a multiple of 64 (<= 1024). red: shared scratch of >= 16 elements.
// Every thread returns the total. Contains two barriers.
template <typename T>
__device__ inline T block_sum(T v, T* red) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  T s = 0;
  for (int w = 0; w < nw; ++w) s += red[w];
  return s;
}

template <typename T>
__device__ inline T sign_of(T a, T b) {  // |a| with the sign of b (Fortran SIGN)
  return b >= T(0) ? fabs(a) : -fabs(a);
}

}  // namespace xtbeig
