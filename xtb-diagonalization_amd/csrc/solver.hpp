// Handle (one per caller thread or per device stream) and the solver drivers.
#pragma once
#include "common.hpp"
#include "kernels.hpp"

#include <map>
#include <tuple>
#include <vector>
#include <memory>
#include <string>

namespace xtbeig {

// status words per handle (xtbeig_synchronize)
constexpr int NSTATUS = 4;
// the widest band the dense-to-band panels, the chase and the back-transforms take, and the largest n of a solve
// (the panel factorization holds a column of up to 2048 rows below the band in one workgroup's registers)
constexpr int MAX_BAND = 64;
constexpr int MAX_N = 2048;
// the largest n of the allocation-free Cholesky entry point (its scratch array is sized for it)
constexpr int NOALLOC_MAX_N = 2048;

struct Handle {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  // the solve's second stream (forming the back-transformation's orthogonal factor while the chase and the divide
  // and conquer run) and the events that fork it off and join it back
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_band = nullptr, ev_chase = nullptr, ev_join = nullptr;
  void* arena = nullptr;
  size_t arena_bytes = 0;
  // second device buffer, for work that wraps a solve (the solve itself carves the arena)
  void* aux = nullptr;
  size_t aux_bytes = 0;
  // device status words: [0] potrf info, [1] chase stall, [2] stedc info, [3] refinement check failed; doubles at
  // dinfo + 32: the last refinement check's two ratios (residual, orthogonality; <= 1 passes)
  int* dinfo = nullptr;
  double* check_ratios() const { return reinterpret_cast<double*>(dinfo + 32); }
  // device address of the allocation-free entry point's scratch array on this handle's device (noalloc.hip)
  void* scratch_dev = nullptr;
  void* pinned = nullptr;
  size_t pinned_bytes = 0;
  // device buffers of the host-array entry points, kept between calls (grown on demand): slot k is the k-th buffer
  // a call holds at once
  std::vector<std::pair<void*, size_t>> io_slots;
  int io_top = 0;
  int refine_iters = 2;
  // test setting: a cap on the divide and conquer's secular-equation iterations (-1: the solver's own limit)
  int test_secular_maxit = -1;
  int band_override = 0;
  // SCC precision schedule: settings (fp32_rmsq 0 = FP64 only) and the counters of the current sequence
  double scc_fp32_rmsq = 0.0;
  int scc_fp32_maxiter = 0, scc_fp32_refine = 0;
  int scc_fp32_cycles = 0, scc_fp64_restart = 0, scc_cycles = 0;
  bool scc_fp64_latched = false;
  // autotuned band widths per n bucket (0: not tuned, the built-in table applies), [0] FP64, [1] FP32 reduction
  static constexpr int NBUCKET = 9;
  int tuned_band[2][NBUCKET] = {};
  // arena reallocations so far: captured graphs remember the generation they point into
  int arena_generation = 0;
  // the geometry of an SCC sequence: S, U and X = U^-1 (n x n each, ld n), factored once by set_geometry and read by
  // every cycle's solve until the next set_geometry
  double* geo = nullptr;
  size_t geo_bytes = 0;
  int geo_n = 0;  // 0: no geometry set
  double *geo_s = nullptr, *geo_u = nullptr, *geo_x = nullptr;
  void set_geometry(int n, const double* sm, int lds);
  std::map<int, std::unique_ptr<DcPlan>> plans;
  std::map<std::pair<int, int>, std::unique_ptr<BackPlan>> backplans;

  explicit Handle(int dev);
  ~Handle();
  Handle(const Handle&) = delete;
  Handle& operator=(const Handle&) = delete;
  void reserve(size_t bytes);
  void reserve_aux(size_t bytes);
  void* host_staging(size_t bytes);
  void* io_slot(int k, size_t bytes);
  const DcPlan& plan(int n);
  const BackPlan& backplan(int n, int b);
  // band width for an n x n solve: the caller's override, else the tuned table, else the built-in one
  int band_for(int n, bool fp32) const;
  // graph cache (batch.hip) and the concurrent sub-handles of the batched solve
  struct GraphEntry {
    hipGraphExec_t exec = nullptr;
    double *h = nullptr, *s = nullptr, *w = nullptr, *c = nullptr;  // staging: H, S in; w, C out (ld n)
    int generation = -1;                                           // arena generation the graph points into
    GraphEntry() = default;
    GraphEntry(const GraphEntry&) = delete;
    GraphEntry& operator=(const GraphEntry&) = delete;
    ~GraphEntry();
  };
  // keyed by (n, precision, band, refinement steps): the step count is a host loop, baked into the capture
  std::map<std::tuple<int, int, int, int>, std::unique_ptr<GraphEntry>> graphs;
  std::vector<std::unique_ptr<Handle>> subs;
};

// noalloc.hip: the scratch array's device address on `device` (current), looked up once per device; the
// allocation-free Cholesky + inverse on it.
void* noalloc_scratch(int device);
size_t noalloc_scratch_bytes();
void potrf_inverse_noalloc(Handle& h, int n, double* s, int lds, double* u, int ldu, double* x, int ldx);

// n bucket edges of the tuned table: bucket q holds bucket_edge(q) <= n < bucket_edge(q + 1)
int bucket_of(int n);
int bucket_edge(int q);

int default_band(int n, bool fp32);

template <typename T>
void syevd_device(Handle& h, int n, const T* a, int lda, double* w, T* z, int ldz, int b, int m = -1);
// u_out (optional): where S's Cholesky factor U (n x n, ld n, zero below the diagonal) stays in the handle's arena
// until the next call on the handle.
void sygvd_device(Handle& h, int n, const double* hm, int ldh, const double* sm, int lds, double* w, double* cm,
                  int ldc, bool mixed, int b, const double** u_out = nullptr);
// arena bytes of a generalized solve (m eigenvector columns)
size_t sygvd_bytes(int n, int b, bool mixed, int m);
// The generalized solve of sygvd_device (band b resolved and checked) on the workspace at base (sygvd_bytes(n, b,
// mixed, n) bytes), with the handle's streams and status words; pad: see PadSpec.
void sygvd_in(Handle& h, void* base, int n, const double* hm, int ldh, const double* sm, int lds, double* w,
              double* cm, int ldc, bool mixed, int b, const double** u_out, const PadSpec* pad);
// Grouped solves (grouped.hip): count problems padded to multiples of `granule` and solved group by group, every
// kernel of a group's pipeline launched once over the whole group. status (count x NSTATUS ints, may be null): each
// problem's status words; latency_ms (may be null): each problem's time to the end of its group.
// slots: groups (chunks of a group) run concurrently on up to this many sub-handle streams.
void sygvd_grouped(Handle& h, int count, const int* n, const double* const* hm, const int* ldh,
                   const double* const* sm, const int* lds, double* const* w, double* const* cm, const int* ldc,
                   bool mixed, int granule, int slots, int* status, float* latency_ms);
void stedc_device(Handle& h, int n, double* d, double* e, double* z, int ldz);
// The cycle's generalized solve on the handle's geometry (set_geometry): all n eigenvalues, the first m eigenvectors.
void sygvd_geometry(Handle& h, int n, int m, const double* hm, int ldh, double* w, double* cm, int ldc, bool mixed,
                    int b);
// Z's upper triangle = X^T H X (X upper triangular, n x n); Z's strictly lower triangle untouched.
void congruence_device(Handle& h, int n, const double* hm, int ldh, const double* x, int ldx, double* z, int ldz);
// C(:, 0:m) = X C(:, 0:m).
void back_solve_device(Handle& h, int n, int m, const double* x, int ldx, double* cm, int ldc);

// batch.hip: the generalized solve through a captured graph (one per (n, precision, band), inputs and outputs
// staged through the graph's own buffers), the batched solve over concurrent sub-handles, and the band autotuner.
void sygvd_graph(Handle& h, int n, const double* hm, int ldh, const double* sm, int lds, double* w, double* cm,
                 int ldc, bool mixed);
void sygvd_batched(Handle& h, int count, const int* n, const double* const* hm, const int* ldh,
                   const double* const* sm, const int* lds, double* const* w, double* const* cm, const int* ldc,
                   bool mixed, int slots, bool use_graph, float* latency_ms);
// h.subs grown to at least `slots` sub-handles, each given h's band, refinement and test settings.
void ensure_subs(Handle& h, int slots);
void autotune_bands(Handle& h, bool mixed, int nmin, int nmax, int* buckets_out, int* bands_out, float* ms_out);

}  // namespace xtbeig
