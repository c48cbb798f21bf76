// Both back-transformations of the two-stage reduction as streams of compact-WY blocks Z <- (I - V T V^T) Z applied by
// one persistent kernel: each workgroup keeps a block of CB columns of Z in LDS and walks the whole sequence of blocks
// (the columns are independent, so the workgroups never synchronize).
//
// Bulge reflectors (the reference's backTransformBulgesBlocked, Stages/BulgeBackTransformation.swift:47-85): the
// reflectors of g consecutive sweeps s0 ..< s0 + g at one bulge index k form one block of height b + g - 1 (column c,
// sweep s0 + c, starts c rows lower), applied for k ascending and, within k, s0 descending. That order is a
// regrouping of the chase order in which only pairs acting on disjoint rows change places (a pair (s, k), (s', k')
// with s < s', k < k' starts at least b + 1 rows apart; the same sweep's reflectors are disjoint), so any g is valid.
// The blocks (explicit V with unit diagonal, forward T) are built once per solve, one workgroup per block.
//
// Band panels (the reference's backTransformBand, Stages/BandFormBackTransformation.swift:10-21): the stored V and T
// of every panel, last panel first.
#include "common.hpp"
#include "kernels.hpp"

#include <vector>

namespace xtbeig {

namespace {

// Build bulge block q: V (height x count, ld ldv, explicit) and T (count x count, ld g, forward columnwise).
template <typename T>
__global__ __launch_bounds__(256) void bulge_blocks_kernel(int n, int b, int g, const WyBlock* __restrict__ blocks,
                                                           const T* __restrict__ v2, int ldv2, T* __restrict__ vout,
                                                           T* __restrict__ tout) {
  __shared__ T G[64][65];
  __shared__ T tau[64];
  __shared__ T Ts[64][65];
  const WyBlock bk = blocks[blockIdx.x];
  const int t = threadIdx.x;
  const int h = bk.height, cnt = bk.count, ldv = bk.ldv;
  T* v = vout + bk.v_off;
  T* tt = tout + bk.t_off;
  const int s0 = bk.s0, k = bk.k;
  // V columns: sweep s = s0 + c, rows rs ..< rs + len at V rows c ..< c + len
  for (int e = t; e < ldv * cnt; e += blockDim.x) {
    const int i = e % ldv, c = e / ldv;
    const int s = s0 + c, rs = s + 1 + k * b;
    const int len = (rs < n) ? min(b, n - rs) : 0;
    T val = 0;
    if (i < h && rs < n) {
      const T w0 = v2[(int64_t)rs + (int64_t)ldv2 * s];
      const int ii = i - c;
      if (w0 != T(0) && ii >= 0 && ii < len) val = (ii == 0) ? T(1) : v2[(int64_t)(rs + ii) + (int64_t)ldv2 * s] / w0;
    }
    v[i + (int64_t)c * ldv] = val;
  }
  for (int c = t; c < cnt; c += blockDim.x) {
    const int s = s0 + c, rs = s + 1 + k * b;
    const T w0 = (rs < n) ? v2[(int64_t)rs + (int64_t)ldv2 * s] : T(0);
    tau[c] = w0 * w0;
  }
  __syncthreads();
  // Gram G = V^T V (upper part needed)
  for (int e = t; e < cnt * cnt; e += blockDim.x) {
    const int r = e % cnt, c = e / cnt;
    if (r < c) {
      T sum = 0;
      for (int i = 0; i < h; ++i) sum += v[i + (int64_t)r * ldv] * v[i + (int64_t)c * ldv];
      G[r][c] = sum;
    }
  }
  __syncthreads();
  // T(0:c, c) = -tau_c T(0:c, 0:c) G(0:c, c), T(c, c) = tau_c
  for (int c = 0; c < cnt; ++c) {
    if (t < c) {
      T sum = 0;
      for (int l = t; l < c; ++l) sum += Ts[t][l] * (-tau[c] * G[l][c]);
      Ts[t][c] = sum;
    }
    if (t == 0) Ts[c][c] = tau[c];
    __syncthreads();
  }
  for (int e = t; e < bk.ldt * cnt; e += blockDim.x) {
    const int r = e % bk.ldt, c = e / bk.ldt;
    tt[r + (int64_t)c * bk.ldt] = (r < cnt && r <= c) ? Ts[r][c] : T(0);
  }
}

// Z(:, c0 ..< c0 + CB) <- B_last ... B_first Z with every block in LDS order of the list.
template <typename T, int CB>
__global__ __launch_bounds__(256) void wy_apply_kernel(int n, int m, const WyBlock* __restrict__ blocks, int nblocks,
                                                       const T* __restrict__ vbase, const T* __restrict__ tbase,
                                                       T* __restrict__ z, int ldz) {
  extern __shared__ char smem_raw[];
  T* zs = reinterpret_cast<T*>(smem_raw);  // n x CB, ld n
  T* w1 = zs + (size_t)n * CB;             // 64 x CB
  T* w2 = w1 + 64 * CB;                    // 64 x CB
  const int t = threadIdx.x;
  const int c0 = blockIdx.x * CB;
  const int cols = min(CB, m - c0);
  for (int e = t; e < n * CB; e += blockDim.x) {
    const int i = e % n, c = e / n;
    zs[e] = (c < cols) ? z[idx(i, c0 + c, ldz)] : T(0);
  }
  __syncthreads();
  for (int q = 0; q < nblocks; ++q) {
    const WyBlock bk = blocks[q];
    const int r0 = bk.row0, h = bk.height, cnt = bk.count, ldv = bk.ldv, ldt = bk.ldt;
    const T* v = vbase + bk.v_off;
    const T* tt = tbase + bk.t_off;
    // W1 = V^T Z(rows)
    for (int o = t; o < cnt * CB; o += blockDim.x) {
      const int c = o / CB, col = o % CB;
      const T* vc = v + (int64_t)c * ldv;
      const T* zc = zs + (size_t)col * n + r0;
      T sum = 0;
      for (int i = 0; i < h; ++i) sum += vc[i] * zc[i];
      w1[c * CB + col] = sum;
    }
    __syncthreads();
    // W2 = T W1 (T upper triangular)
    for (int o = t; o < cnt * CB; o += blockDim.x) {
      const int c = o / CB, col = o % CB;
      T sum = 0;
      for (int l = c; l < cnt; ++l) sum += tt[c + (int64_t)l * ldt] * w1[l * CB + col];
      w2[c * CB + col] = sum;
    }
    __syncthreads();
    // Z(rows) -= V W2
    for (int o = t; o < h * CB; o += blockDim.x) {
      const int i = o % h, col = o / h;
      T sum = 0;
      for (int c = 0; c < cnt; ++c) sum += v[i + (int64_t)c * ldv] * w2[c * CB + col];
      zs[(size_t)col * n + r0 + i] -= sum;
    }
    __syncthreads();
  }
  for (int e = t; e < n * cols; e += blockDim.x) {
    const int i = e % n, c = e / n;
    z[idx(i, c0 + c, ldz)] = zs[e];
  }
}

template <typename T>
void launch_apply(hipStream_t s, int n, int m, const WyBlock* blocks, int nblocks, const T* vbase, const T* tbase,
                  T* z, int ldz) {
  if (m <= 0 || nblocks <= 0) return;
  constexpr int CB = sizeof(T) == 8 ? 8 : 16;
  const size_t bytes = sizeof(T) * ((size_t)n * CB + 2 * 64 * CB);
  if (bytes > 160 * 1024) throw std::runtime_error("wy back-transform: n too large for the LDS column block");
  allow_lds(wy_apply_kernel<T, CB>, bytes);
  hipLaunchKernelGGL((wy_apply_kernel<T, CB>), dim3(cdiv(m, CB)), dim3(256), bytes, s, n, m, blocks, nblocks, vbase,
                     tbase, z, ldz);
  XTB_LAUNCH_CHECK();
}

}  // namespace

int bulge_group(int b) { return std::min(64, std::max(2 * b, 8)); }

WyPlan::WyPlan(int n_, int b_) : n(n_), b(b_), g(bulge_group(b_)) {
  // bulge blocks in application order: k ascending, s0 descending
  std::vector<WyBlock> bl;
  const int sweeps = n - 2;
  const int ldv = b + g;  // >= height
  if (b > 1 && sweeps > 0) {
    for (int k = 0; 1 + k * b < n; ++k) {
      for (int s0 = ((sweeps - 1) / g) * g; s0 >= 0; s0 -= g) {
        const int r0 = s0 + 1 + k * b;
        if (r0 >= n) continue;
        WyBlock w{};
        w.row0 = r0;
        w.height = std::min(b + g - 1, n - r0);
        w.count = std::min(std::min(g, sweeps - s0), w.height);
        w.s0 = s0;
        w.k = k;
        w.ldv = ldv;
        w.ldt = g;
        w.v_off = (int64_t)bl.size() * ldv * g;
        w.t_off = (int64_t)bl.size() * g * g;
        bl.push_back(w);
      }
    }
  }
  nbulge = (int)bl.size();
  // band panels, last first; V and T point into v1 (ld n) and t1 (ld b)
  std::vector<WyBlock> pl;
  int last = -1;
  for (int j = 0; j < n - b - 1; j += b) last = j;
  for (int j = last; j >= 0; j -= b) {
    const int r = j + b;
    WyBlock w{};
    w.row0 = r;
    w.height = n - r;
    w.count = std::min(b, n - r);
    w.ldv = n;
    w.ldt = b;
    w.v_off = (int64_t)r + (int64_t)j * n;
    w.t_off = (int64_t)j * b;
    pl.push_back(w);
  }
  npanel = (int)pl.size();
  if (nbulge) {
    XTB_HIP(hipMalloc(&bulge_dev, sizeof(WyBlock) * nbulge));
    XTB_HIP(hipMemcpy(bulge_dev, bl.data(), sizeof(WyBlock) * nbulge, hipMemcpyHostToDevice));
  }
  if (npanel) {
    XTB_HIP(hipMalloc(&panel_dev, sizeof(WyBlock) * npanel));
    XTB_HIP(hipMemcpy(panel_dev, pl.data(), sizeof(WyBlock) * npanel, hipMemcpyHostToDevice));
  }
}

WyPlan::~WyPlan() {
  if (bulge_dev) (void)hipFree(bulge_dev);
  if (panel_dev) (void)hipFree(panel_dev);
}

size_t WyPlan::v_elems() const { return (size_t)nbulge * (b + g) * g + 1; }
size_t WyPlan::t_elems() const { return (size_t)nbulge * g * g + 1; }

template <typename T>
void back_transform_bulges_wy(hipStream_t s, const WyPlan& p, const T* v2, T* z, int ldz, int m, T* vwork,
                              T* twork) {
  if (p.nbulge == 0 || m <= 0) return;
  if (p.g > 64) throw std::runtime_error("wy back-transform: group larger than 64");
  hipLaunchKernelGGL(bulge_blocks_kernel<T>, dim3(p.nbulge), dim3(256), 0, s, p.n, p.b, p.g, p.bulge_dev, v2, p.n,
                     vwork, twork);
  XTB_LAUNCH_CHECK();
  launch_apply<T>(s, p.n, m, p.bulge_dev, p.nbulge, vwork, twork, z, ldz);
}

template <typename T>
void back_transform_band_wy(hipStream_t s, const WyPlan& p, const T* v1, const T* t1, T* z, int ldz, int m) {
  launch_apply<T>(s, p.n, m, p.panel_dev, p.npanel, v1, t1, z, ldz);
}

template void back_transform_bulges_wy<double>(hipStream_t, const WyPlan&, const double*, double*, int, int, double*,
                                               double*);
template void back_transform_bulges_wy<float>(hipStream_t, const WyPlan&, const float*, float*, int, int, float*,
                                              float*);
template void back_transform_band_wy<double>(hipStream_t, const WyPlan&, const double*, const double*, double*, int,
                                             int);
template void back_transform_band_wy<float>(hipStream_t, const WyPlan&, const float*, const float*, float*, int,
                                            int);

}  // namespace xtbeig
