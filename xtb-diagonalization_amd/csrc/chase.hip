// Stage 2 of the two-stage reduction: band (width b) -> tridiagonal by Schwarz bulge chasing, one Householder
// reflector per bulge (the reference's chaseBulgesLower / chaseBulgeLower, Stages/BulgeChasing.swift:123-321), and
// the back-transformation through its reflectors (Stages/BulgeBackTransformation.swift:11-30).
//
// Chase: one workgroup; each wavefront owns one sweep at a time (sweep s on wave s mod W) and chases its bulges down
// the band. Bulge k of sweep s waits until sweep s - 1 has finished bulges 0 .. k + 3 (the reference's pipelining
// rule, BulgeChasing.swift:143-153), published through LDS progress counters. Every pair of bulges whose windows
// overlap therefore runs in sequential order, and the result equals the one-sweep-at-a-time chase. Per bulge, with
// H = I - w w^T on rows R = rs ..< re, the lower triangle of the window splits into three blocks:
//   columns lo ..< rs, rows R:   A[R, j] -= w (w^T A[R, j])        lane = column
//   rows re ..< hi, columns R:   A[i, R] -= (A[i, R] w) w^T        lane = row
//   rows and columns R:          A_RR -= w q^T + q w^T, p = A_RR w, q = p - (w^T p) w / 2
// Band storage: element (i, j), 0 <= i - j <= 2b - 1, at i + (2b - 1) j. The reference's compact band uses leading
// dimension 2b for diagonals 0 ..< 2b + 1 (BulgeChasing.swift:8-12); the chase itself never reaches further than
// 2b - 1 below the diagonal (left block: rs + b - 1 - (rs - b); block below: re + b - 1 - rs), so one diagonal less
// is stored and no two stored elements share an address. The band lives in LDS when its 2b n elements fit (FP32 up
// to b = 16, FP64 up to b = 8 at n = 1204), else in global memory (L2-resident at these sizes).
//
// Back-transformation: Z = H_1 H_2 ... H_N Z. Bulges of one level p = 4 s + k act on disjoint rows (the reference's
// ordering argument, BulgeChasing.swift:133-140), so levels run from last to first and the reflectors of a level in
// parallel, one wavefront each, on a block of Z's columns kept in LDS.
#include "common.hpp"
#include "kernels.hpp"

namespace xtbeig {

namespace {

constexpr int CHASE_WAVES = 16;

__device__ inline int ld_acquire(const int* p) { return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP); }
__device__ inline void st_release(int* p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP); }

template <typename T>
struct Band {
  T* ab;
  int ld;  // 2b
  __device__ inline T& operator()(int i, int j) const { return ab[(int64_t)i + (int64_t)ld * j]; }
};

// One bulge (s, k) by one wavefront. Returns false when the sweep has no such bulge.
template <typename T>
__device__ inline bool chase_bulge(int n, int b, const Band<T>& A, T* __restrict__ v2, int ldv2, int s, int k,
                                   T* wsh, T* qsh) {
  const int lane = threadIdx.x & 63;
  const int rs = s + 1 + k * b;
  if (rs >= n) return false;
  const int re = min(rs + b, n), len = re - rs;
  const int c = (k == 0) ? s : rs - b;
  // larfg on x = A[R, c]
  const T xi = (lane < len) ? A(rs + lane, c) : T(0);
  const T alpha = __shfl(xi, 0, 64);
  const T xn2 = wave_sum((lane >= 1 && lane < len) ? xi * xi : T(0));
  T tau = 0, beta = alpha, scal = 0;
  if (xn2 != T(0)) {
    beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
    tau = (beta - alpha) / beta;
    scal = T(1) / (alpha - beta);
  }
  const T sq = sqrt(tau);
  const T wi = (lane == 0) ? sq : ((lane < len) ? sq * (xi * scal) : T(0));
  if (lane < len) {
    v2[(int64_t)(rs + lane) + (int64_t)ldv2 * s] = wi;
    wsh[lane] = wi;
  }
  __builtin_amdgcn_wave_barrier();
  if (tau != T(0)) {
    const int lo = max(rs - b, 0), hi = min(re + b, n);
    // left block: columns lo ..< rs, lane = column
    if (lane < rs - lo) {
      const int j = lo + lane;
      T dot = 0;
      for (int i = 0; i < len; ++i) dot += wsh[i] * A(rs + i, j);
      for (int i = 0; i < len; ++i) A(rs + i, j) -= wsh[i] * dot;
    }
    // block below: rows re ..< hi, lane = row
    if (lane < hi - re) {
      const int i = re + lane;
      T y = 0;
      for (int jj = 0; jj < len; ++jj) y += A(i, rs + jj) * wsh[jj];
      for (int jj = 0; jj < len; ++jj) A(i, rs + jj) -= y * wsh[jj];
    }
    // diagonal block R x R (lower triangle stored): p = A_RR w
    T p = 0;
    if (lane < len) {
      const int i = rs + lane;
      for (int jj = 0; jj < len; ++jj) {
        const T a = (jj <= lane) ? A(i, rs + jj) : A(rs + jj, i);
        p += a * wsh[jj];
      }
    }
    const T half = wave_sum(lane < len ? wi * p : T(0)) * T(0.5);
    const T q = p - half * wi;
    if (lane < len) qsh[lane] = q;
    __builtin_amdgcn_wave_barrier();
    if (lane < len) {
      const int i = rs + lane;
      for (int jj = 0; jj <= lane; ++jj) A(i, rs + jj) -= wi * qsh[jj] + q * wsh[jj];
    }
  }
  __threadfence_block();
  // exact zeros where the reflector annihilated
  if (lane < len) A(rs + lane, c) = (lane == 0) ? beta : T(0);
  __threadfence_block();
  return true;
}

template <typename T, bool LDS>
__global__ __launch_bounds__(64 * CHASE_WAVES) void chase_kernel(int n, int b, T* __restrict__ ab_global,
                                                                 T* __restrict__ v2, int ldv2, T* __restrict__ d,
                                                                 T* __restrict__ e, int* __restrict__ status) {
  // dynamic LDS: per-wave w and q vectors (b each), then the band when it is LDS-resident
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
          const int need = (s - 1) * KB + (k + 4);
          long spins = 0;
          while (ld_acquire(&prog[pw]) < need) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > (1L << 24)) { failed = true; break; }
          }
          if (failed) break;
        }
        const bool more = chase_bulge<T>(n, b, A, v2, ldv2, s, k, wbuf + wid * b, qbuf + wid * b);
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

// Z = H_1 ... H_N Z on a block of CB columns (LDS, n x CB), one level of disjoint reflectors at a time.
template <typename T, int CB>
__global__ __launch_bounds__(1024) void bulge_back_kernel(int n, int b, const T* __restrict__ v2, int ldv2,
                                                          T* __restrict__ z, int ldz, int m) {
  extern __shared__ char smem_raw[];
  T* zs = reinterpret_cast<T*>(smem_raw);  // n x CB, ld n
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6, nw = blockDim.x >> 6;
  const int c0 = blockIdx.x * CB;
  const int cols = min(CB, m - c0);
  for (int q = t; q < n * CB; q += blockDim.x) {
    const int i = q % n, c = q / n;
    zs[q] = (c < cols) ? z[idx(i, c0 + c, ldz)] : T(0);
  }
  __syncthreads();
  const int sweeps = n - 2;
  const int kmax0 = (n - 2) / b + 1;
  const int pmax = 4 * (sweeps - 1) + kmax0;
  // lanes: row r = lane % 32 (two passes for b > 32), column c = lane / 32 (+2 per pass)
  for (int p = pmax; p >= 0; --p) {
    // reflectors (s, k) with 4 s + k = p: s from max(0, ceil((p - kmax0) / 4)) to min(sweeps - 1, p / 4)
    const int smin = max(0, (p - kmax0 + 3) / 4), smax = min(sweeps - 1, p / 4);
    for (int s = smin + wid; s <= smax; s += nw) {
      const int k = p - 4 * s;
      const int rs = s + 1 + k * b;
      if (rs >= n) continue;
      const int len = min(rs + b, n) - rs;
      const T* w = v2 + (int64_t)rs + (int64_t)ldv2 * s;
      // w values per lane (rows lane and lane + 64 unused: b <= 64)
      const T wl = (lane < len) ? w[lane] : T(0);
      for (int c = 0; c < CB; ++c) {
        T* zc = zs + (int64_t)c * n + rs;
        const T zv = (lane < len) ? zc[lane] : T(0);
        const T dot = wave_sum(wl * zv);
        if (lane < len) zc[lane] = zv - wl * dot;
      }
    }
    __syncthreads();
  }
  for (int q = t; q < n * cols; q += blockDim.x) {
    const int i = q % n, c = q / n;
    z[idx(i, c0 + c, ldz)] = zs[q];
  }
}

}  // namespace

template <typename T>
void chase_bulges(hipStream_t s, int n, int b, T* ab, T* v2, T* d, T* e, int* status) {
  if (b > 64) throw std::runtime_error("chase_bulges: band width must be <= 64");
  const size_t vec_bytes = 2 * sizeof(T) * CHASE_WAVES * (size_t)b;
  const size_t band_bytes = sizeof(T) * (size_t)band_elems(n, b);
  if (vec_bytes + band_bytes + sizeof(int) * CHASE_WAVES <= 160 * 1024) {
    allow_lds(chase_kernel<T, true>, vec_bytes + band_bytes);
    hipLaunchKernelGGL((chase_kernel<T, true>), dim3(1), dim3(64 * CHASE_WAVES), vec_bytes + band_bytes, s, n, b, ab,
                       v2, n, d, e, status);
  } else {
    hipLaunchKernelGGL((chase_kernel<T, false>), dim3(1), dim3(64 * CHASE_WAVES), vec_bytes, s, n, b, ab, v2, n, d,
                       e, status);
  }
  XTB_LAUNCH_CHECK();
}

template <typename T>
void back_transform_bulges(hipStream_t s, int n, int b, const T* v2, T* z, int ldz, int m) {
  if (n <= 2 || b <= 1 || m <= 0) return;
  constexpr int CB = sizeof(T) == 8 ? 8 : 16;
  const size_t bytes = sizeof(T) * (size_t)n * CB;
  if (bytes > 160 * 1024) throw std::runtime_error("back_transform_bulges: n too large for the LDS column block");
  allow_lds(bulge_back_kernel<T, CB>, bytes);
  hipLaunchKernelGGL((bulge_back_kernel<T, CB>), dim3(cdiv(m, CB)), dim3(1024), bytes, s, n, b, v2, n, z, ldz, m);
  XTB_LAUNCH_CHECK();
}

template void chase_bulges<double>(hipStream_t, int, int, double*, double*, double*, double*, int*);
template void chase_bulges<float>(hipStream_t, int, int, float*, float*, float*, float*, int*);
template void back_transform_bulges<double>(hipStream_t, int, int, const double*, double*, int, int);
template void back_transform_bulges<float>(hipStream_t, int, int, const float*, float*, int, int);

}  // namespace xtbeig
