// Grouped solves: many independent generalized eigenproblems of uneven n, each kernel of the solve launched once over
// a whole group of problems (configuration 3's 256 problems of n in [124, 1204]).
//
// Grouping. Problems are padded to a common size N (n rounded up to a multiple of `granule`) and grouped by N. A
// group's problems get equal slices of the handle's arena, `stride` bytes apart, and the solve is enqueued once under
// a BatchScope: every launch of the pipeline carries the group in its grid's z extent and every kernel moves its
// pointers into its problem's slice (Zb). The chase of 40 problems is then one launch of 40 workgroups on 40 CUs, a
// dense-to-band panel step or a D&C level one launch for all 40, each product one batched GEMM; the launch count of
// a group is that of one solve.
//
// Padding. A problem of size n < N is solved as H = [H_p 0; 0 D], S = [S_p 0; 0 I]. The Cholesky factor and X = U^-1
// are block diagonal, the congruence leaves the blocks uncoupled, every Householder vector of the padded columns is
// zero (tau = 0: the panels and bulges test for a zero column), the reflectors of the problem's own columns are zero
// in the padded rows, and each D&C merge's rank-one term is zero on the padded poles, which therefore deflate. The
// padded block's eigenpairs are (D_kk, e_k), and the problem's own eigenvectors are exactly zero in the padded rows.
// D is set after the congruence (pad_spectrum): distinct values inside the spectral range of the problem's H', so
// neither the D&C's scale (its deflation tolerance) nor the refinement's eigenvalue gaps change. After the solve the
// columns that vanish in the padded rows are the problem's own, kept in their (ascending) order; a count that is not
// n is reported in status word 3 (value 2).
//
// Concurrency. The stages of a small group leave most of the GPU idle (the chase of 5 problems runs on 5 CUs, a D&C
// level's kernels are short), so groups run at once on up to `slots` sub-handles with their own streams and arenas:
// chunks go most expensive first (cost count x N^3) to the least-loaded slot, and the host enqueues them round robin
// so that every stream starts early.
//
// Limits. A slice holds the padded H, S, w, C and a full solve's workspace; a chunk is sized so that its slices fit
// the slot's share of the memory budget (a fraction of the device's free memory) and so that no launch's grid z
// extent (the problem count times the kernel's own z extent, at most twice the widest D&C level) exceeds 65535.
#include "common.hpp"
#include "kernels.hpp"
#include "solver.hpp"

#include <algorithm>
#include <functional>
#include <map>

namespace xtbeig {

Batch& current_batch() {
  thread_local Batch b;
  return b;
}

namespace {

struct ProbIO {
  const double* h;
  const double* s;
  double* w;
  double* c;
  int ldh, lds, ldc, n;
};

// Slice layout of a group of padded size N: status words (the handle's 64-int status block, the refinement check's
// ratios at byte 128), the column positions of the compaction, H, S, w, C, then the solve's workspace.
struct SliceLayout {
  size_t pos, h, s, w, c, arena, bytes;
  SliceLayout(int N, int b, bool mixed) {
    const size_t nn = (size_t)N * N;
    auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
    pos = 512;
    h = up(pos + sizeof(int) * (size_t)N);
    s = up(h + sizeof(double) * nn);
    w = up(s + sizeof(double) * nn);
    c = up(w + sizeof(double) * (size_t)N);
    arena = up(c + sizeof(double) * nn);
    bytes = (arena + sygvd_bytes(N, b, mixed, N) + 4095) & ~size_t(4095);
  }
};

// H, S of problem blockIdx.z into its slice, padded to N (H zero, S the identity outside the problem's block); the
// slice's status block cleared.
__global__ void group_pack_kernel(int N, const ProbIO* __restrict__ io, char* __restrict__ base, size_t stride,
                                  size_t off_h, size_t off_s) {
  const int p = blockIdx.z, j = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
  const ProbIO q = io[p];
  char* sl = base + (size_t)p * stride;
  if (j == 0 && blockIdx.x == 0 && threadIdx.x < 128) reinterpret_cast<int*>(sl)[threadIdx.x] = 0;
  if (i >= N) return;
  double* H = reinterpret_cast<double*>(sl + off_h);
  double* S = reinterpret_cast<double*>(sl + off_s);
  const bool in = i < q.n && j < q.n;
  H[idx(i, j, N)] = in ? q.h[idx(i, j, q.ldh)] : 0.0;
  S[idx(i, j, N)] = in ? q.s[idx(i, j, q.lds)] : (i == j ? 1.0 : 0.0);
}

template <bool MAX>
__device__ inline double block_reduce(double v, double* red) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(v, off, 64);
    v = MAX ? fmax(v, o) : fmin(v, o);
  }
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double r = red[0];
  for (int q = 1; q < (int)(blockDim.x >> 6); ++q) r = MAX ? fmax(r, red[q]) : fmin(r, red[q]);
  return r;
}

__global__ __launch_bounds__(256) void pad_spectrum_kernel(int n, const int* __restrict__ nvec, double* hp, int ldhp,
                                                           double* h, int ldh, Zb zb) {
  __shared__ double red[4];
  const int np = nvec[zb.prob()], t = threadIdx.x;
  if (np >= n) return;
  hp = zb.at(hp);
  h = zb.at(h);
  double lo = INFINITY, hi = -INFINITY;
  for (int i = t; i < np; i += blockDim.x) {
    const double d = hp[idx(i, i, ldhp)];
    lo = fmin(lo, d);
    hi = fmax(hi, d);
  }
  lo = block_reduce<false>(lo, red);
  hi = block_reduce<true>(hi, red);
  double spread = hi - lo;
  if (!(spread > 0.0)) {
    // a constant diagonal c: lambda_max >= c + |a_ij| for every off-diagonal a_ij (the 2 x 2 principal submatrix's
    // larger eigenvalue, by interlacing), so (c, c + max |a_ij|) is inside the range; H' = c I: (c, 2c)
    double m = 0.0;
    for (int64_t e = t; e < (int64_t)np * np; e += blockDim.x) {
      const int i = (int)(e % np), j = (int)(e / np);
      if (i != j) m = fmax(m, fabs(hp[idx(i, j, ldhp)]));
    }
    spread = block_reduce<true>(m, red);
    if (!(spread > 0.0)) spread = fabs(lo) > 0.0 ? fabs(lo) : 1.0;
  }
  for (int k = np + t; k < n; k += blockDim.x) {
    const double v = lo + spread * (double)(k - np + 1) / (double)(n - np + 1);
    hp[idx(k, k, ldhp)] = v;
    if (h) h[idx(k, k, ldh)] = v;
  }
}

// The compaction of problem blockIdx.z: column j of C (N x N) is one of the problem's own when it is exactly zero in
// the padded rows; pos[j] = its place among them (in column order), or -1. Status word 3 |= 2 when the count of own
// columns is not the problem's n.
__global__ __launch_bounds__(256) void group_select_kernel(int N, const ProbIO* __restrict__ io,
                                                           char* __restrict__ base, size_t stride, size_t off_c,
                                                           size_t off_pos) {
  __shared__ int part[256];
  const int p = blockIdx.z, t = threadIdx.x;
  char* sl = base + (size_t)p * stride;
  const double* C = reinterpret_cast<const double*>(sl + off_c);
  int* pos = reinterpret_cast<int*>(sl + off_pos);
  const int np = io[p].n;
  const int per = cdiv(N, 256), j0 = t * per, j1 = min(N, j0 + per);
  int own = 0;
  for (int j = j0; j < j1; ++j) {
    bool padded = false;
    for (int i = np; i < N; ++i) padded |= !(C[idx(i, j, N)] == 0.0);
    pos[j] = padded ? -1 : 0;
    own += padded ? 0 : 1;
  }
  part[t] = own;
  __syncthreads();
  if (t == 0) {
    int run = 0;
    for (int q = 0; q < 256; ++q) {
      const int v = part[q];
      part[q] = run;
      run += v;
    }
    if (run != np) atomicOr(reinterpret_cast<int*>(sl) + 3, 2);
  }
  __syncthreads();
  int next = part[t];
  for (int j = j0; j < j1; ++j)
    if (pos[j] == 0) pos[j] = next++;
}

// w and C of problem blockIdx.z out of its slice: column j of the padded C to the problem's column pos[j].
__global__ void group_unpack_kernel(int N, const ProbIO* __restrict__ io, const char* __restrict__ base,
                                    size_t stride, size_t off_w, size_t off_c, size_t off_pos) {
  const int p = blockIdx.z, j = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
  const ProbIO q = io[p];
  const char* sl = base + (size_t)p * stride;
  const int dst = reinterpret_cast<const int*>(sl + off_pos)[j];
  if (dst < 0 || dst >= q.n) return;
  const double* C = reinterpret_cast<const double*>(sl + off_c);
  if (i < q.n) q.c[idx(i, dst, q.ldc)] = C[idx(i, j, N)];
  if (i == 0) q.w[dst] = reinterpret_cast<const double*>(sl + off_w)[j];
}

// the status words of the chunk's problems into the table (count x NSTATUS)
__global__ void group_status_kernel(int count, const char* __restrict__ base, size_t stride, int* __restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= count * NSTATUS) return;
  out[q] = reinterpret_cast<const int*>(base + (size_t)(q / NSTATUS) * stride)[q % NSTATUS];
}

}  // namespace

void pad_spectrum(hipStream_t s, int n, const int* nvec, double* hp, int ldhp, double* h, int ldh) {
  hipLaunchKernelGGL(pad_spectrum_kernel, dim3(1, 1, zgrid()), dim3(256), 0, s, n, nvec, hp, ldhp, h, ldh, zb());
  XTB_LAUNCH_CHECK();
}

void sygvd_grouped(Handle& h, int count, const int* n, const double* const* hm, const int* ldh,
                   const double* const* sm, const int* lds, double* const* w, double* const* cm, const int* ldc,
                   bool mixed, int granule, int slots, int* status, float* latency_ms) {
  if (count <= 0) return;
  if (granule <= 0) granule = 64;
  // padded size per problem: n rounded up to the granule, unless that passes the size limit
  std::map<int, std::vector<int>, std::greater<int>> groups;
  for (int q = 0; q < count; ++q) {
    int N = (n[q] + granule - 1) / granule * granule;
    const int bN = std::max(1, std::min(h.band_for(N, mixed), N - 1));
    if (N - bN > MAX_N) N = n[q];
    groups[N].push_back(q);
  }
  slots = std::max(1, std::min(slots <= 0 ? 4 : slots, 64));
  size_t free_b = 0, total_b = 0;
  XTB_HIP(hipSetDevice(h.device));
  XTB_HIP(hipMemGetInfo(&free_b, &total_b));
  size_t held = h.arena_bytes;
  for (auto& sub : h.subs) held += sub->arena_bytes;
  const size_t budget = std::max<size_t>(std::min<size_t>((size_t)(0.6 * (double)(free_b + held)), size_t(64) << 30),
                                         size_t(1) << 30);
  struct Chunk {
    int N, b, first, count;
    SliceLayout lay;
    bool padded;
    double cost;
  };
  std::vector<Chunk> chunks;
  std::vector<ProbIO> tab;
  std::vector<int> nv, order;
  // a group is cut into equal chunks when its slices exceed a slot's share of the memory budget or its launches the
  // grid's z limit
  const size_t slot_budget = std::max<size_t>(budget / slots, size_t(256) << 20);
  for (auto& g : groups) {
    const int N = g.first;
    const std::vector<int>& list = g.second;
    int b = h.band_for(N, mixed);
    b = std::max(1, std::min(b, std::max(1, N - 1)));
    if (b > MAX_BAND || N - b > MAX_N)
      throw std::runtime_error("grouped solve: n = " + std::to_string(N) + " exceeds the largest supported size");
    const SliceLayout lay(N, b, mixed);
    const DcPlan& plan = h.plan(N);
    int zmax = 64;  // the largest own z extent of a launch: the D&C merge products (two per merge of a level)
    for (const auto& lv : plan.lvls) zmax = std::max(zmax, 2 * lv.count);
    const int cap = (int)std::max<size_t>(1, std::min<size_t>(slot_budget / lay.bytes, (size_t)(65535 / zmax)));
    const int pieces = std::max(1, cdiv((int)list.size(), cap));
    const int per = cdiv((int)list.size(), pieces);
    for (size_t at = 0; at < list.size(); at += per) {
      const int cnt = (int)std::min<size_t>(per, list.size() - at);
      Chunk ch{N, b, (int)tab.size(), cnt, lay, false, (double)cnt * N * N * N};
      for (int k = 0; k < cnt; ++k) {
        const int q = list[at + k];
        tab.push_back(ProbIO{hm[q], sm[q], w[q], cm[q], ldh[q], lds[q], ldc[q], n[q]});
        nv.push_back(n[q]);
        order.push_back(q);
        ch.padded |= n[q] < N;
      }
      chunks.push_back(ch);
    }
  }
  // chunks onto slots, most expensive first onto the least-loaded slot; one slot: the handle itself
  slots = std::min<int>(slots, (int)chunks.size());
  std::vector<Handle*> hs;
  if (slots == 1) {
    hs.push_back(&h);
  } else {
    ensure_subs(h, slots);
    for (int sl = 0; sl < slots; ++sl) hs.push_back(h.subs[sl].get());
  }
  std::vector<int> by_cost(chunks.size());
  for (size_t ci = 0; ci < chunks.size(); ++ci) by_cost[ci] = (int)ci;
  std::stable_sort(by_cost.begin(), by_cost.end(), [&](int a, int b) { return chunks[a].cost > chunks[b].cost; });
  std::vector<std::vector<int>> lists(slots);
  std::vector<double> load(slots, 0.0);
  for (int ci : by_cost) {
    const int sl = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    lists[sl].push_back(ci);
    load[sl] += chunks[ci].cost;
  }
  // every slot's arena and plans before anything is enqueued
  for (int sl = 0; sl < slots; ++sl) {
    size_t need = 0;
    for (int ci : lists[sl]) {
      const Chunk& ch = chunks[ci];
      need = std::max(need, (size_t)ch.count * ch.lay.bytes);
      (void)hs[sl]->plan(ch.N);
      (void)hs[sl]->backplan(ch.N, ch.b);
    }
    hs[sl]->reserve(need);
  }
  // device tables: the problems' buffers and sizes, then the status words (count x NSTATUS), in chunk order
  const size_t tab_bytes = (sizeof(ProbIO) * tab.size() + 255) & ~size_t(255);
  const size_t nv_bytes = (sizeof(int) * nv.size() + 255) & ~size_t(255);
  const size_t st_bytes = sizeof(int) * NSTATUS * (size_t)count;
  h.reserve_aux(tab_bytes + nv_bytes + st_bytes);
  char* aux = static_cast<char*>(h.aux);
  ProbIO* tab_dev = reinterpret_cast<ProbIO*>(aux);
  int* nv_dev = reinterpret_cast<int*>(aux + tab_bytes);
  int* st_dev = reinterpret_cast<int*>(aux + tab_bytes + nv_bytes);
  char* stage = static_cast<char*>(h.host_staging(tab_bytes + nv_bytes));
  std::copy(tab.begin(), tab.end(), reinterpret_cast<ProbIO*>(stage));
  std::copy(nv.begin(), nv.end(), reinterpret_cast<int*>(stage + tab_bytes));
  hipStream_t st = h.stream;
  XTB_HIP(hipMemcpyAsync(aux, stage, tab_bytes + nv_bytes, hipMemcpyHostToDevice, st));
  std::vector<hipEvent_t> ev(chunks.size() + 1 + slots, nullptr);
  for (auto& e : ev) XTB_HIP(hipEventCreate(&e));
  hipEvent_t* done = ev.data() + chunks.size() + 1;
  XTB_HIP(hipEventRecord(ev[0], st));
  for (int sl = 0; sl < slots; ++sl)
    if (hs[sl] != &h) XTB_HIP(hipStreamWaitEvent(hs[sl]->stream, ev[0], 0));
  auto enqueue = [&](Handle& hh, int ci) {
    const Chunk& ch = chunks[ci];
    const int N = ch.N;
    const SliceLayout& L = ch.lay;
    const ProbIO* io = tab_dev + ch.first;
    hipStream_t s = hh.stream;
    char* base = static_cast<char*>(hh.arena);
    hipLaunchKernelGGL(group_pack_kernel, dim3(cdiv(N, 256), N, ch.count), dim3(256), 0, s, N, io, base, L.bytes, L.h,
                       L.s);
    XTB_LAUNCH_CHECK();
    {
      BatchScope scope(ch.count, L.bytes);
      int* const keep = hh.dinfo;
      hh.dinfo = reinterpret_cast<int*>(base);
      double* H = reinterpret_cast<double*>(base + L.h);
      const PadSpec pad{nv_dev + ch.first, H, N};
      try {
        sygvd_in(hh, base + L.arena, N, H, N, reinterpret_cast<double*>(base + L.s), N,
                 reinterpret_cast<double*>(base + L.w), reinterpret_cast<double*>(base + L.c), N, mixed, ch.b,
                 nullptr, ch.padded ? &pad : nullptr);
      } catch (...) {
        hh.dinfo = keep;
        throw;
      }
      hh.dinfo = keep;
    }
    hipLaunchKernelGGL(group_select_kernel, dim3(1, 1, ch.count), dim3(256), 0, s, N, io, base, L.bytes, L.c, L.pos);
    XTB_LAUNCH_CHECK();
    hipLaunchKernelGGL(group_unpack_kernel, dim3(cdiv(N, 256), N, ch.count), dim3(256), 0, s, N, io, base, L.bytes,
                       L.w, L.c, L.pos);
    XTB_LAUNCH_CHECK();
    hipLaunchKernelGGL(group_status_kernel, dim3(cdiv(ch.count * NSTATUS, 256)), dim3(256), 0, s, ch.count, base,
                       L.bytes, st_dev + (size_t)NSTATUS * ch.first);
    XTB_LAUNCH_CHECK();
    XTB_HIP(hipEventRecord(ev[ci + 1], s));
  };
  try {
    // round robin over the slots, so that every slot starts while the host is still enqueueing
    size_t longest = 0;
    for (auto& l : lists) longest = std::max(longest, l.size());
    for (size_t pos = 0; pos < longest; ++pos)
      for (int sl = 0; sl < slots; ++sl)
        if (pos < lists[sl].size()) enqueue(*hs[sl], lists[sl][pos]);
    for (int sl = 0; sl < slots; ++sl)
      if (hs[sl] != &h) {
        XTB_HIP(hipEventRecord(done[sl], hs[sl]->stream));
        XTB_HIP(hipStreamWaitEvent(st, done[sl], 0));
      }
  } catch (...) {
    for (int sl = 0; sl < slots; ++sl) (void)hipStreamSynchronize(hs[sl]->stream);
    for (auto& e : ev) (void)hipEventDestroy(e);
    throw;
  }
  XTB_HIP(hipStreamSynchronize(st));
  std::vector<int> sts((size_t)NSTATUS * count);
  XTB_HIP(hipMemcpy(sts.data(), st_dev, st_bytes, hipMemcpyDeviceToHost));
  // merged into the handle's words: the first nonzero value of each
  int tot[NSTATUS];
  XTB_HIP(hipMemcpy(tot, h.dinfo, sizeof(tot), hipMemcpyDeviceToHost));
  bool changed = false;
  for (int k = 0; k < count; ++k)
    for (int q = 0; q < NSTATUS; ++q)
      if (sts[(size_t)NSTATUS * k + q] && !tot[q]) {
        tot[q] = sts[(size_t)NSTATUS * k + q];
        changed = true;
      }
  if (changed) XTB_HIP(hipMemcpy(h.dinfo, tot, sizeof(tot), hipMemcpyHostToDevice));
  for (size_t ci = 0; ci < chunks.size(); ++ci) {
    const Chunk& ch = chunks[ci];
    float ms = 0;
    if (latency_ms) XTB_HIP(hipEventElapsedTime(&ms, ev[0], ev[ci + 1]));
    for (int k = 0; k < ch.count; ++k) {
      const int q = order[ch.first + k];
      if (latency_ms) latency_ms[q] = ms;
      if (status)
        for (int s = 0; s < NSTATUS; ++s) status[(size_t)NSTATUS * q + s] = sts[(size_t)NSTATUS * (ch.first + k) + s];
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
}

}  // namespace xtbeig
