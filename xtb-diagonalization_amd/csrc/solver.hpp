// Handle (one per caller thread or per device stream) and the solver drivers.
#pragma once
#include "common.hpp"
#include "kernels.hpp"

#include <map>
#include <memory>
#include <string>

namespace xtbeig {

struct Handle {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  void* arena = nullptr;
  size_t arena_bytes = 0;
  // second device buffer, for work that wraps a solve (the solve itself carves the arena)
  void* aux = nullptr;
  size_t aux_bytes = 0;
  // device status words: [0] potrf info, [1] chase stall, [2] stedc info
  int* dinfo = nullptr;
  void* pinned = nullptr;
  size_t pinned_bytes = 0;
  int refine_iters = 2;
  int band_override = 0;
  // SCC precision schedule: settings (fp32_rmsq 0 = FP64 only) and the counters of the current sequence
  double scc_fp32_rmsq = 0.0;
  int scc_fp32_maxiter = 0, scc_fp32_refine = 0;
  int scc_fp32_cycles = 0, scc_fp64_restart = 0, scc_cycles = 0;
  bool scc_fp64_latched = false;
  std::map<int, std::unique_ptr<DcPlan>> plans;
  std::map<std::pair<int, int>, std::unique_ptr<BackPlan>> backplans;

  explicit Handle(int dev);
  ~Handle();
  Handle(const Handle&) = delete;
  Handle& operator=(const Handle&) = delete;
  void reserve(size_t bytes);
  void reserve_aux(size_t bytes);
  void* host_staging(size_t bytes);
  const DcPlan& plan(int n);
  const BackPlan& backplan(int n, int b);
};

int default_band(int n, bool fp32);

template <typename T>
void syevd_device(Handle& h, int n, const T* a, int lda, double* w, T* z, int ldz, int b);
void sygvd_device(Handle& h, int n, const double* hm, int ldh, const double* sm, int lds, double* w, double* cm,
                  int ldc, bool mixed, int b);
void stedc_device(Handle& h, int n, double* d, double* e, double* z, int ldz);

}  // namespace xtbeig
