// Shared device/host helpers for the MI355X (gfx950) eigensolver library.
// Matrices are column-major (Fortran ABI), element (i, j) of a matrix with leading dimension ld at i + j * ld.
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <type_traits>

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

// Dynamic LDS beyond 64 KiB (gfx950 has 160 KiB per workgroup): raise the kernel's limit before launching. The
// attribute is process-wide, so it is always raised to the whole 160 KiB, never to one call's size: concurrent calls
// from several threads (different n, different LDS sizes) then all write the same value, and none can lower the
// limit under another's launch.
constexpr size_t LDS_MAX_BYTES = 160 * 1024;
template <typename K>
inline void allow_lds(K kernel, size_t bytes) {
  if (bytes <= 64 * 1024) return;
  // the limit is on dynamic LDS: what the kernel's static LDS leaves of the 160 KiB
  hipFuncAttributes attr;
  XTB_HIP(hipFuncGetAttributes(&attr, reinterpret_cast<const void*>(kernel)));
  const size_t room = LDS_MAX_BYTES - std::min(LDS_MAX_BYTES, (size_t)attr.sharedSizeBytes);
  if (bytes > room) throw std::runtime_error("allow_lds: more LDS than a workgroup has");
  XTB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)room));
}

__host__ __device__ inline int64_t idx(int64_t i, int64_t j, int64_t ld) { return i + j * ld; }
__host__ __device__ inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// ---- problem batches ------------------------------------------------------------------------------------------------
// A batched solve enqueues the pipeline once for `count` problems of one size whose buffers lie in equal arena slices
// `stride` bytes apart. Every kernel of the pipeline carries the problem in the slow part of its grid's z index
// (z = problem * zin + the kernel's own z, zin its own z extent) and moves each pointer into that problem's slice with
// Zb::at before use. A single solve is the batch of one: stride 0, at() returns its argument, the grid is unchanged.
// Tables shared by every problem of the batch (the divide-and-conquer tree, the bulge back-transformation plan) are
// not moved.
struct Zb {
  size_t stride;  // bytes between consecutive problems' slices (0: a single problem)
  int zin;        // the kernel's own z extent
  // Uniform branches keep the division off a single solve's path (these kernels are short and latency-bound).
  __device__ inline int prob() const { return stride == 0 ? 0 : zin == 1 ? (int)blockIdx.z : (int)blockIdx.z / zin; }
  __device__ inline int z() const { return (stride == 0 || zin == 1) ? (zin == 1 ? 0 : (int)blockIdx.z)
                                                                      : (int)blockIdx.z % zin; }
  // byte arithmetic on the pointer itself (no round trip through an integer), so that the compiler still knows a
  // kernel argument points to global memory and emits global rather than flat loads and stores
  template <typename P>
  __device__ inline P* at(P* p) const {
    using Byte = typename std::conditional<std::is_const<P>::value, const char, char>::type;
    if (stride == 0) return p;
    return p ? reinterpret_cast<P*>(reinterpret_cast<Byte*>(p) + (size_t)prob() * stride) : p;
  }
};
// The batch that the calling host thread is enqueuing (count 1, stride 0 outside a BatchScope).
struct Batch {
  int count = 1;
  size_t stride = 0;
};
Batch& current_batch();
inline Zb zb(int zin = 1) { return Zb{current_batch().stride, zin}; }
// grid z extent of a launch whose own z extent is zin
inline unsigned zgrid(int zin = 1) {
  const long long z = (long long)zin * current_batch().count;
  if (z > 65535) throw std::runtime_error("batched launch: grid z extent " + std::to_string(z) + " past 65535");
  return (unsigned)z;
}
// Sets the calling thread's batch for its lifetime (nested scopes restore the outer one).
struct BatchScope {
  Batch saved;
  BatchScope(int count, size_t stride) : saved(current_batch()) {
    current_batch().count = count;
    current_batch().stride = stride;
  }
  ~BatchScope() { current_batch() = saved; }
  BatchScope(const BatchScope&) = delete;
  BatchScope& operator=(const BatchScope&) = delete;
};

// ---- wavefront (64-lane) reductions -------------------------------------------------------------
// DPP moves within rows of 16 lanes (quad permutes, half-row and row mirrors), for 32-bit lanes; 64-bit values move
// as two halves.
template <int CTRL>
__device__ inline float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ inline double dpp_mov(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
constexpr int DPP_XOR1 = 0xB1;         // quad_perm [1, 0, 3, 2]
constexpr int DPP_XOR2 = 0x4E;         // quad_perm [2, 3, 0, 1]
constexpr int DPP_HALF_MIRROR = 0x141;  // lane i <-> 7 - i within 8
constexpr int DPP_MIRROR = 0x140;       // lane i <-> 15 - i within 16
constexpr int DPP_ROR8 = 0x128;         // lane i <-> i ^ 8 within 16 (row rotate by 8)

// The value of lane l ^ D (D = 16 or 32) through the gfx950 row and half swaps: v_permlane16_swap exchanges rows 1
// and 3 of its first operand with rows 0 and 2 of its second, v_permlane32_swap the upper half of the first with the
// lower half of the second; with both operands the same value, a lane finds its partner's value in the result that
// took the other row (half).
template <int D>
__device__ inline unsigned xor_lanes_u32(unsigned v) {
  static_assert(D == 16 || D == 32, "xor_lanes: D must be 16 or 32");
  const int lane = threadIdx.x & 63;
  if constexpr (D == 16) {
    const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (lane & 16) ? r[0] : r[1];
  } else {
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (lane & 32) ? r[0] : r[1];
  }
}
template <int D>
__device__ inline float xor_lanes(float v) {
  return __uint_as_float(xor_lanes_u32<D>(__float_as_uint(v)));
}
template <int D>
__device__ inline double xor_lanes(double v) {
  const unsigned lo = xor_lanes_u32<D>((unsigned)__double2loint(v));
  const unsigned hi = xor_lanes_u32<D>((unsigned)__double2hiint(v));
  return __hiloint2double((int)hi, (int)lo);
}

// Sum over aligned groups of G lanes (G a power of two <= 64); every lane of a group receives its group's total.
// The mirror steps pair whole sub-groups once those are uniform.
template <int G, typename T>
__device__ inline T group_sum_lanes(T v) {
  if (G >= 2) v += dpp_mov<DPP_XOR1>(v);
  if (G >= 4) v += dpp_mov<DPP_XOR2>(v);
  if (G >= 8) v += dpp_mov<DPP_HALF_MIRROR>(v);
  if (G >= 16) v += dpp_mov<DPP_MIRROR>(v);
  if (G >= 32) v += __shfl_xor(v, 16, 64);
  if (G >= 64) v += __shfl_xor(v, 32, 64);
  return v;
}

template <typename T>
__device__ inline T wave_sum(T v) {
  return group_sum_lanes<64>(v);
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
