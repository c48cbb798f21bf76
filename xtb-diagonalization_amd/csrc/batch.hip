// Many solves: a solve captured once as a HIP graph and replayed (the SCC cycle's repeated eigensolves of one n),
// batches of uneven-n problems spread over concurrent streams, and the per-bucket band-width autotuner. (Batches
// solved inside the kernels, a group at a time, are in grouped.hip.)
//
// Graphs. A solve enqueues several hundred kernels; replaying them from a graph removes the host's launch work and the
// gaps between launches. The graph for (n, precision, band, refinement steps) is captured after one uncaptured warm-up
// solve (the D&C plan, the back-transformation plan and the arena exist then, so the capture records kernels only: no
// allocation, no memset, no host synchronization). It reads H and S from its own staging buffers and writes w and C
// there; a call copies in, launches, copies out. A graph remembers the arena generation it was captured against: when a
// later, larger solve reallocates the arena, the graph is captured again.
//
// Batches. Each generalized solve at these n is latency-bound (the chase runs on one CU, the D&C levels and panel
// factorizations on few), so a batch runs up to `slots` solves at once on sub-handles with their own streams and
// arenas. Problems are assigned longest first (cost ~ n^3) to the least-loaded slot; the slots wait for the caller's
// stream before starting and the caller's stream waits for all slots at the end. HIP events around every solve give
// the per-solve latency. For eager solves the arenas and plans of every problem are set up before anything is
// enqueued; with use_graph each solve goes through its slot's graph cache.
//
// Autotuner. For each n bucket of the tuned table within [nmin, nmax] the candidate band widths are timed on a
// synthetic symmetric pair of the bucket's geometric-middle size (two timed solves after a warm-up, the minimum
// kept); the fastest is written to the handle's table, which every later solve on the handle consults.
#include "common.hpp"
#include "kernels.hpp"
#include "solver.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace xtbeig {

Handle::GraphEntry::~GraphEntry() {
  if (exec) (void)hipGraphExecDestroy(exec);
  if (h) (void)hipFree(h);
  if (s) (void)hipFree(s);
  if (w) (void)hipFree(w);
  if (c) (void)hipFree(c);
}

namespace {

// Synthetic symmetric test matrices for the autotuner: H with entries in [-1, 1] (a hash of the position), S = I +
// 0.1 (hash) with a dominant diagonal (positive definite).
__device__ inline double hash01(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return (x >> 8) * (1.0 / 16777216.0);
}

__global__ void synthetic_pair_kernel(int n, double* __restrict__ h, double* __restrict__ s) {
  const int j = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int lo = min(i, j), hi = max(i, j);
  const uint32_t key = (uint32_t)lo * 2654435761U + (uint32_t)hi * 40503U;
  h[i + (int64_t)j * n] = 2.0 * hash01(key) - 1.0;
  s[i + (int64_t)j * n] = (i == j) ? 1.0 + 0.1 * n : 0.1 * hash01(key ^ 0x9e3779b9U);
}

}  // namespace

void sygvd_graph(Handle& h, int n, const double* hm, int ldh, const double* sm, int lds, double* w, double* cm,
                 int ldc, bool mixed) {
  if (n <= 0) return;
  const int b = h.band_for(n, mixed);
  const auto key = std::make_tuple(n, mixed ? 1 : 0, b, mixed ? h.refine_iters : 0);
  auto it = h.graphs.find(key);
  Handle::GraphEntry* g = (it == h.graphs.end()) ? nullptr : it->second.get();
  const size_t nn = (size_t)n * n;
  if (!g) {
    auto e = std::make_unique<Handle::GraphEntry>();
    XTB_HIP(hipMalloc(&e->h, nn * sizeof(double)));
    XTB_HIP(hipMalloc(&e->s, nn * sizeof(double)));
    XTB_HIP(hipMalloc(&e->w, n * sizeof(double)));
    XTB_HIP(hipMalloc(&e->c, nn * sizeof(double)));
    g = e.get();
    h.graphs.emplace(key, std::move(e));
  }
  hipStream_t st = h.stream;
  XTB_HIP(hipMemcpy2DAsync(g->h, n * sizeof(double), hm, ldh * sizeof(double), n * sizeof(double), n,
                           hipMemcpyDeviceToDevice, st));
  XTB_HIP(hipMemcpy2DAsync(g->s, n * sizeof(double), sm, lds * sizeof(double), n * sizeof(double), n,
                           hipMemcpyDeviceToDevice, st));
  if (!g->exec || g->generation != h.arena_generation) {
    if (g->exec) {
      XTB_HIP(hipGraphExecDestroy(g->exec));
      g->exec = nullptr;
    }
    // warm-up (plans, arena, kernel attributes), then the capture of the same solve on the same buffers
    sygvd_device(h, n, g->h, n, g->s, n, g->w, g->c, n, mixed, b);
    hipGraph_t graph = nullptr;
    XTB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    try {
      sygvd_device(h, n, g->h, n, g->s, n, g->w, g->c, n, mixed, b);
    } catch (...) {
      (void)hipStreamEndCapture(st, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      throw;
    }
    XTB_HIP(hipStreamEndCapture(st, &graph));
    const hipError_t rc = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    XTB_HIP(rc);
    g->generation = h.arena_generation;
  }
  XTB_HIP(hipGraphLaunch(g->exec, st));
  XTB_HIP(hipMemcpyAsync(w, g->w, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  XTB_HIP(hipMemcpy2DAsync(cm, ldc * sizeof(double), g->c, n * sizeof(double), n * sizeof(double), n,
                           hipMemcpyDeviceToDevice, st));
}

void ensure_subs(Handle& h, int slots) {
  while ((int)h.subs.size() < slots) h.subs.push_back(std::make_unique<Handle>(h.device));
  for (auto& sub : h.subs) {
    sub->band_override = h.band_override;
    sub->refine_iters = h.refine_iters;
    sub->test_secular_maxit = h.test_secular_maxit;
    std::copy(&h.tuned_band[0][0], &h.tuned_band[0][0] + 2 * Handle::NBUCKET, &sub->tuned_band[0][0]);
  }
}

void sygvd_batched(Handle& h, int count, const int* n, const double* const* hm, const int* ldh,
                   const double* const* sm, const int* lds, double* const* w, double* const* cm, const int* ldc,
                   bool mixed, int slots, bool use_graph, float* latency_ms) {
  if (count <= 0) return;
  slots = std::max(1, std::min(slots, count));
  ensure_subs(h, slots);
  // longest first onto the least-loaded slot
  std::vector<int> order(count);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return n[a] > n[b]; });
  std::vector<double> load(slots, 0.0);
  std::vector<std::vector<int>> lists(slots);
  for (int q : order) {
    const int sl = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    lists[sl].push_back(q);
    load[sl] += std::pow((double)n[q], 3.0);
  }
  // Everything an eager solve would allocate, up front: each slot's arena for its largest problem and the D&C and
  // back-transformation plans of every size it will see, so that the enqueue loop below never waits on the host.
  if (!use_graph) {
    for (int sl = 0; sl < slots; ++sl) {
      Handle& sub = *h.subs[sl];
      size_t need = 0;
      for (int q : lists[sl]) {
        if (n[q] <= 1) continue;
        const int bq = std::max(1, std::min(sub.band_for(n[q], mixed), n[q] - 1));
        need = std::max(need, sygvd_bytes(n[q], bq, mixed, n[q]));
        (void)sub.plan(n[q]);
        (void)sub.backplan(n[q], bq);
      }
      if (need) sub.reserve(need);
    }
  }
  hipEvent_t ready;
  XTB_HIP(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
  XTB_HIP(hipEventRecord(ready, h.stream));
  std::vector<hipEvent_t> ev(2 * (size_t)count, nullptr);
  for (auto& e : ev) XTB_HIP(hipEventCreate(&e));
  std::vector<hipEvent_t> done(slots, nullptr);
  for (int sl = 0; sl < slots; ++sl) XTB_HIP(hipStreamWaitEvent(h.subs[sl]->stream, ready, 0));
  // enqueue round robin over the slots, so that every slot starts while the host is still enqueueing
  size_t longest = 0;
  for (auto& l : lists) longest = std::max(longest, l.size());
  for (size_t pos = 0; pos < longest; ++pos)
    for (int sl = 0; sl < slots; ++sl) {
      if (pos >= lists[sl].size()) continue;
      Handle& sub = *h.subs[sl];
      const int q = lists[sl][pos];
      XTB_HIP(hipEventRecord(ev[2 * q], sub.stream));
      if (use_graph) sygvd_graph(sub, n[q], hm[q], ldh[q], sm[q], lds[q], w[q], cm[q], ldc[q], mixed);
      else sygvd_device(sub, n[q], hm[q], ldh[q], sm[q], lds[q], w[q], cm[q], ldc[q], mixed, 0);
      XTB_HIP(hipEventRecord(ev[2 * q + 1], sub.stream));
    }
  for (int sl = 0; sl < slots; ++sl) {
    XTB_HIP(hipEventCreateWithFlags(&done[sl], hipEventDisableTiming));
    XTB_HIP(hipEventRecord(done[sl], h.subs[sl]->stream));
    XTB_HIP(hipStreamWaitEvent(h.stream, done[sl], 0));
  }
  XTB_HIP(hipStreamSynchronize(h.stream));
  // the slots' status words: any failure is reported on the caller's handle
  for (int sl = 0; sl < slots; ++sl) {
    int st[NSTATUS];
    XTB_HIP(hipMemcpy(st, h.subs[sl]->dinfo, NSTATUS * sizeof(int), hipMemcpyDeviceToHost));
    XTB_HIP(hipMemset(h.subs[sl]->dinfo, 0, NSTATUS * sizeof(int)));
    int tot[NSTATUS];
    XTB_HIP(hipMemcpy(tot, h.dinfo, NSTATUS * sizeof(int), hipMemcpyDeviceToHost));
    bool changed = false;
    for (int k = 0; k < NSTATUS; ++k)
      if (st[k] && !tot[k]) {
        tot[k] = st[k];
        changed = true;
      }
    if (changed) XTB_HIP(hipMemcpy(h.dinfo, tot, NSTATUS * sizeof(int), hipMemcpyHostToDevice));
  }
  if (latency_ms)
    for (int q = 0; q < count; ++q) XTB_HIP(hipEventElapsedTime(&latency_ms[q], ev[2 * q], ev[2 * q + 1]));
  for (auto& e : ev) (void)hipEventDestroy(e);
  for (auto& e : done) (void)hipEventDestroy(e);
  (void)hipEventDestroy(ready);
}

void autotune_bands(Handle& h, bool mixed, int nmin, int nmax, int* buckets_out, int* bands_out, float* ms_out) {
  const int cand[] = {8, 16, 32};
  const int ncand = 3;
  const int keep_override = h.band_override;
  int idx_out = 0;
  for (int q = 0; q < Handle::NBUCKET; ++q) {
    const int lo = std::max(bucket_edge(q), 2), hi = bucket_edge(q + 1) - 1;
    if (hi < nmin || lo > nmax) continue;
    const int a = std::max(lo, nmin), z = std::min(hi, nmax);
    const int n = std::max(2, (int)std::lround(std::sqrt((double)a * (double)z)));
    const size_t nn = (size_t)n * n;
    double *hm = nullptr, *sm = nullptr, *w = nullptr, *cm = nullptr;
    XTB_HIP(hipMalloc(&hm, nn * sizeof(double)));
    XTB_HIP(hipMalloc(&sm, nn * sizeof(double)));
    XTB_HIP(hipMalloc(&w, n * sizeof(double)));
    XTB_HIP(hipMalloc(&cm, nn * sizeof(double)));
    hipLaunchKernelGGL(synthetic_pair_kernel, dim3(cdiv(n, 256), n), dim3(256), 0, h.stream, n, hm, sm);
    XTB_LAUNCH_CHECK();
    int best_b = 0;
    float best_ms = 1e30f;
    hipEvent_t e0, e1;
    XTB_HIP(hipEventCreate(&e0));
    XTB_HIP(hipEventCreate(&e1));
    try {
      for (int ci = 0; ci < ncand; ++ci) {
        const int b = std::min(cand[ci], std::max(2, n - 1));
        h.band_override = b;
        sygvd_device(h, n, hm, n, sm, n, w, cm, n, mixed, b);  // warm-up: plans, arena
        float t = 1e30f;
        for (int r = 0; r < 2; ++r) {
          XTB_HIP(hipEventRecord(e0, h.stream));
          sygvd_device(h, n, hm, n, sm, n, w, cm, n, mixed, b);
          XTB_HIP(hipEventRecord(e1, h.stream));
          XTB_HIP(hipEventSynchronize(e1));
          float ms = 0;
          XTB_HIP(hipEventElapsedTime(&ms, e0, e1));
          t = std::min(t, ms);
        }
        int st[NSTATUS];
        XTB_HIP(hipMemcpy(st, h.dinfo, NSTATUS * sizeof(int), hipMemcpyDeviceToHost));
        if ((st[0] | st[1] | st[2] | st[3]) == 0 && t < best_ms) {
          best_ms = t;
          best_b = b;
        }
        XTB_HIP(hipMemset(h.dinfo, 0, NSTATUS * sizeof(int)));
      }
    } catch (...) {
      h.band_override = keep_override;
      (void)hipFree(hm); (void)hipFree(sm); (void)hipFree(w); (void)hipFree(cm);
      (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
      throw;
    }
    h.band_override = keep_override;
    (void)hipFree(hm); (void)hipFree(sm); (void)hipFree(w); (void)hipFree(cm);
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
    h.tuned_band[mixed ? 1 : 0][q] = best_b;
    if (buckets_out) buckets_out[idx_out] = bucket_edge(q);
    if (bands_out) bands_out[idx_out] = best_b;
    if (ms_out) ms_out[idx_out] = best_ms;
    ++idx_out;
  }
}

}  // namespace xtbeig
