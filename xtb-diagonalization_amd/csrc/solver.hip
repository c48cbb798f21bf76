// The eigensolver drivers: per-handle device workspace, the standard solve (two-stage reduction, divide and conquer,
// both back-transformations) and the generalized solve in FP64 or in FP32 with FP64 refinement.
//
// Generalized solve H C = S C Lambda (itype 1), as xTB's fact_solve path (the reference: scripts/kernels/
// eigbench.f90:101-168, scripts/kernels/congruence.c:1-52):
//   S = U^T U and X = U^{-1}                          FP64 always (chol.hip)
//   H' = X^T H X                                      two GEMMs, FP64 always (rounded to FP32 for the FP32 solve)
//   H' Z = Z Lambda                                    syevd_core: dense -> band (MFMA), bulge chase, FP64 D&C,
//                                                      bulge and band back-transforms, FP64 or FP32
//   C = X Z                                           GEMM, FP64 or FP32
//   FP32 only: FP64 Ogita-Aishima refinement of (Lambda, C) against (H, S), the eigenpair sort and the a-posteriori
//   check (status word 3).
#include "solver.hpp"

#include <algorithm>
#include <cstring>

namespace xtbeig {

// Band width per n: 8 in both precisions. The chase of a band of 8 is the cheaper per sweep whether its band fits one
// CU's LDS (FP64 up to n = 1255, FP32 up to about 2500) or not (then its last columns stay on chip): at n = 1788 in
// FP64, 14.7 ms at band 8 against 31.3 ms at band 16 with both bands in device memory; and the twice as many
// dense-to-band panels cost less than that saves.
int default_band(int n, bool fp32) {
  (void)fp32;
  if (n <= 3) return 1;
  return std::min(8, std::max(2, n - 1));
}

Handle::Handle(int dev) : device(dev) {
  XTB_HIP(hipSetDevice(device));
  // the handle's stream is an ordinary (blocking) stream: work on the legacy default stream (a caller's memset or
  // copy of an input, a framework's default-stream kernels) is ordered with it without an explicit event; the second
  // stream is internal, forked from and joined to this one by events
  XTB_HIP(hipStreamCreateWithFlags(&stream, hipStreamDefault));
  own_stream = true;
  XTB_HIP(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking));
  XTB_HIP(hipEventCreateWithFlags(&ev_band, hipEventDisableTiming));
  XTB_HIP(hipEventCreateWithFlags(&ev_chase, hipEventDisableTiming));
  XTB_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
  XTB_HIP(hipMalloc(&dinfo, 64 * sizeof(int)));
  XTB_HIP(hipMemset(dinfo, 0, 64 * sizeof(int)));
  scratch_dev = noalloc_scratch(device);
}

Handle::~Handle() {
  if (stream) (void)hipStreamSynchronize(stream);
  if (stream2) (void)hipStreamSynchronize(stream2);
  if (ev_band) (void)hipEventDestroy(ev_band);
  if (ev_chase) (void)hipEventDestroy(ev_chase);
  if (ev_join) (void)hipEventDestroy(ev_join);
  if (stream2) (void)hipStreamDestroy(stream2);
  if (arena) (void)hipFree(arena);
  if (aux) (void)hipFree(aux);
  if (geo) (void)hipFree(geo);
  if (dinfo) (void)hipFree(dinfo);
  if (pinned) (void)hipHostFree(pinned);
  for (auto& sl : io_slots)
    if (sl.first) (void)hipFree(sl.first);
  if (own_stream && stream) (void)hipStreamDestroy(stream);
}

void Handle::reserve(size_t bytes) {
  if (bytes <= arena_bytes) return;
  XTB_HIP(hipSetDevice(device));
  if (arena) {
    XTB_HIP(hipStreamSynchronize(stream));
    XTB_HIP(hipFree(arena));
    arena = nullptr;
    arena_bytes = 0;
  }
  XTB_HIP(hipMalloc(&arena, bytes));
  arena_bytes = bytes;
  ++arena_generation;
}

int bucket_edge(int q) {
  static const int edges[Handle::NBUCKET + 1] = {0, 192, 320, 512, 768, 1024, 1536, 2048, 4096, 1 << 30};
  return edges[q];
}

int bucket_of(int n) {
  int q = 0;
  while (q + 1 < Handle::NBUCKET && n >= bucket_edge(q + 1)) ++q;
  return q;
}

int Handle::band_for(int n, bool fp32) const {
  if (band_override > 0) return band_override;
  const int t = tuned_band[fp32 ? 1 : 0][bucket_of(n)];
  if (t > 0) return std::min(t, std::max(2, n - 1));
  return default_band(n, fp32);
}

void Handle::reserve_aux(size_t bytes) {
  if (bytes <= aux_bytes) return;
  XTB_HIP(hipSetDevice(device));
  if (aux) {
    XTB_HIP(hipStreamSynchronize(stream));
    XTB_HIP(hipFree(aux));
    aux = nullptr;
    aux_bytes = 0;
  }
  XTB_HIP(hipMalloc(&aux, bytes));
  aux_bytes = bytes;
}

void* Handle::host_staging(size_t bytes) {
  if (bytes > pinned_bytes) {
    if (pinned) XTB_HIP(hipHostFree(pinned));
    pinned = nullptr;
    XTB_HIP(hipHostMalloc(&pinned, bytes, hipHostMallocDefault));
    pinned_bytes = bytes;
  }
  return pinned;
}

void* Handle::io_slot(int k, size_t bytes) {
  if ((int)io_slots.size() <= k) io_slots.resize(k + 1, {nullptr, 0});
  auto& sl = io_slots[k];
  if (bytes > sl.second) {
    // the previous contents are no longer in use: every host-array call synchronizes before it returns
    if (sl.first) XTB_HIP(hipFree(sl.first));
    sl.first = nullptr;
    XTB_HIP(hipMalloc(&sl.first, bytes));
    sl.second = bytes;
  }
  return sl.first;
}

const DcPlan& Handle::plan(int n) {
  auto it = plans.find(n);
  if (it != plans.end()) return *it->second;
  auto p = std::make_unique<DcPlan>(n);
  const DcPlan& ref = *p;
  plans.emplace(n, std::move(p));
  return ref;
}

const BackPlan& Handle::backplan(int n, int b) {
  const auto key = std::make_pair(n, b);
  auto it = backplans.find(key);
  if (it != backplans.end()) return *it->second;
  auto p = std::make_unique<BackPlan>(n, b);
  const BackPlan& ref = *p;
  backplans.emplace(key, std::move(p));
  return ref;
}

// ---- workspace carving --------------------------------------------------------------------------------------------

struct Carver {
  char* base;
  size_t off = 0;
  explicit Carver(void* b) : base(static_cast<char*>(b)) {}
  template <typename T>
  T* take(size_t count) {
    off = (off + 255) & ~size_t(255);
    T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += count * sizeof(T);
    return p;
  }
};

template <typename T>
struct StdBuffers {
  T *v1, *t1, *ab, *bwork, *v2, *dT, *eT, *btwork, *zt, *qt;
  double *d64, *e64, *z64, *dcwork;
  int* dciwork;
};

template <typename T>
StdBuffers<T> carve_std(Carver& c, int n, int b) {
  StdBuffers<T> s;
  const size_t nn = (size_t)n * n;
  s.v1 = c.take<T>(nn);
  s.t1 = c.take<T>((size_t)b * n);
  s.ab = c.take<T>(band_elems(n, b));
  s.bwork = c.take<T>(band_workspace<T>(n, b));
  s.v2 = c.take<T>(nn);
  s.dT = c.take<T>(n);
  s.eT = c.take<T>(n);
  s.btwork = c.take<T>(back_band_workspace<T>(n, b, n));
  s.zt = c.take<T>((size_t)zt_ld(n) * zt_rows(n));
  s.qt = c.take<T>((size_t)zt_ld(n) * zt_rows(n));
  s.d64 = c.take<double>(n);
  s.e64 = c.take<double>(n);
  s.z64 = c.take<double>(nn);
  s.dcwork = c.take<double>(dc_workspace_doubles(n));
  s.dciwork = c.take<int>(dc_workspace_ints(n));
  return s;
}

// Eigenpairs of the symmetric a (n x n, full storage, destroyed): w ascending (FP64), z (n x n).
// z = Q1 Q2 Zdc: Q1 from the dense-to-band panels, Q2 from the chase, Zdc the tridiagonal's eigenvectors. M = Q1 Q2
// is formed on the handle's second stream while the chase and the divide and conquer run on the first: Q1 as the band
// back-transformation of the identity (as soon as the band is there), then M = Q1 Q2 by the bulge reflectors applied
// from the right (once the chase is done); the first stream then needs one product, z = M Zdc.
// pre (optional, n x n, ld ldpre): z = pre Q1 Q2 Zdc instead, with P = pre M formed on the second stream into pm
// (n x n, ld n) right after M, so that the first stream still needs one product after the divide and conquer.
// m (1 <= m <= n, default n): only the first m eigenvectors are formed (all n eigenvalues are): the last product
// costs 2 n^2 m instead of 2 n^3, as xTB's occupied-only back-transformation.
template <typename T>
void syevd_core(Handle& h, const StdBuffers<T>& s, int n, int b, T* a, int lda, double* w, T* z, int ldz,
                const T* pre = nullptr, int ldpre = 0, T* pm = nullptr, int m = -1) {
  hipStream_t st = h.stream, st2 = h.stream2;
  if (m < 0 || m > n) m = n;
  if (n == 1) {
    copy_matrix<T, double>(st, 1, 1, a, lda, w, 1);
    set_identity(st, 1, z, ldz);
    return;
  }
  const int ldzt = zt_ld(n), rows = zt_rows(n);
  const BackPlan& bp = h.backplan(n, b);
  reduce_to_band<T>(st, n, b, a, lda, s.v1, s.t1, s.ab, s.bwork);
  XTB_HIP(hipEventRecord(h.ev_band, st));
  // first stream: the chase, then the tridiagonal's eigenpairs. The chase is enqueued before the second stream's
  // work, so that it is queued behind the panels however long the host takes to enqueue the launches below.
  chase_bulges<T>(st, n, b, s.ab, s.v2, s.dT, s.eT, h.dinfo + 1);
  XTB_HIP(hipEventRecord(h.ev_chase, st));
  XTB_HIP(hipStreamWaitEvent(st2, h.ev_band, 0));
  // second stream: zt = (Q1 I)^T, qt = Q1 (both ldzt x rows, the rows past n zero)
  set_zero<T>(st2, ldzt, rows, s.zt, ldzt);
  set_identity<T>(st2, n, s.zt, ldzt);
  back_transform_band<T>(st2, bp, s.v1, s.t1, s.zt, ldzt, n, s.btwork);
  set_zero<T>(st2, ldzt, rows - n, s.qt + (size_t)n * ldzt, ldzt);
  transpose_matrix<T, T>(st2, n, n, s.zt, ldzt, s.qt, ldzt);
  // second stream: qt <- qt Q2 = Q1 Q2 (qt read as a column-major matrix with leading dimension ldzt), enqueued
  // before the divide and conquer so that it starts as soon as the chase ends
  XTB_HIP(hipStreamWaitEvent(st2, h.ev_chase, 0));
  back_transform_bulges<T>(st2, bp, s.v2, s.qt, ldzt, n, true);
  if (pre) gemm<T>(st2, 'N', 'N', n, n, n, T(1), pre, ldpre, s.qt, ldzt, T(0), pm, n);
  XTB_HIP(hipEventRecord(h.ev_join, st2));
  copy_matrix<T, double>(st, n, 1, s.dT, n, s.d64, n);
  copy_matrix<T, double>(st, n - 1, 1, s.eT, n, s.e64, n);
  stedc(st, h.plan(n), s.d64, s.e64, s.z64, n, s.dcwork, s.dciwork, h.dinfo + 2, h.test_secular_maxit);
  copy_matrix<double, double>(st, n, 1, s.d64, n, w, n);
  XTB_HIP(hipStreamWaitEvent(st, h.ev_join, 0));
  // z = M Zdc (Zdc rounded to T in zt, which the second stream no longer needs)
  const T* zdc;
  if (std::is_same<T, double>::value) {
    zdc = reinterpret_cast<const T*>(s.z64);
  } else {
    copy_matrix<double, T>(st, n, m, s.z64, n, s.zt, n);
    zdc = s.zt;
  }
  if (pre) gemm<T>(st, 'N', 'N', n, m, n, T(1), pm, n, zdc, n, T(0), z, ldz);
  else gemm<T>(st, 'N', 'N', n, m, n, T(1), s.qt, ldzt, zdc, n, T(0), z, ldz);
}

template <typename T>
size_t syevd_bytes(int n, int b) {
  Carver c(nullptr);
  c.take<T>((size_t)n * n);  // working copy
  carve_std<T>(c, n, b);
  return c.off + 4096;
}

// The dense-to-band panel factorization holds a panel column in registers, at most 2048 rows below the band (8 rows
// per thread of one workgroup), and the divide and conquer stages each merge in LDS: larger solves are refused before
// any work is enqueued.
static void check_size(int n, int b) {
  if (b > MAX_BAND)
    throw std::runtime_error("band width " + std::to_string(b) + " exceeds the widest supported band, " +
                             std::to_string(MAX_BAND));
  if (n - b > MAX_N)
    throw std::runtime_error("n = " + std::to_string(n) + " exceeds the largest supported size, 2048 + band width (" +
                             std::to_string(b) + ")");
}

template <typename T>
void syevd_device(Handle& h, int n, const T* a, int lda, double* w, T* z, int ldz, int b, int m) {
  if (n <= 0) return;
  if (b <= 0) b = h.band_for(n, sizeof(T) == 4);
  b = std::max(1, std::min(b, std::max(1, n - 1)));
  check_size(n, b);
  h.reserve(syevd_bytes<T>(n, b));
  Carver c(h.arena);
  T* work = c.take<T>((size_t)n * n);
  StdBuffers<T> s = carve_std<T>(c, n, b);
  copy_matrix<T, T>(h.stream, n, n, a, lda, work, n);
  syevd_core<T>(h, s, n, b, work, n, w, z, ldz, nullptr, 0, nullptr, m);
}

// The arena of the solve's steps after the Cholesky factorization (the congruence, the standard solve, the
// back-solve and, FP32 + refinement, the refinement), for m eigenvectors.
static void carve_core_bytes(Carver& c, int n, int b, bool mixed, int m) {
  const size_t nn = (size_t)n * n;
  if (!mixed) {
    c.take<double>(nn);  // H X, then X Q1 Q2
    c.take<double>(nn);  // H'
    carve_std<double>(c, n, b);
  } else {
    c.take<float>(nn);  // X32
    c.take<float>(nn);  // X Q1 Q2
    c.take<float>(nn);  // H' rounded
    c.take<float>(nn);  // C32
    carve_std<float>(c, n, b);
    c.take<double>(refine_workspace(n));  // also H X and H' in FP64 before the refinement
    c.take<int>(refine_iwork(n));
    if (m < n) c.take<double>(nn);  // all n refined eigenvectors, of which m are returned
    c.take<double>(16);              // max |H'| for the FP32 copy's range
  }
}

size_t sygvd_bytes(int n, int b, bool mixed, int m) {
  Carver c(nullptr);
  const size_t nn = (size_t)n * n;
  c.take<double>(nn);  // S copy
  c.take<double>(nn);  // U
  c.take<double>(nn);  // X
  c.take<double>(potrf_workspace(n));
  carve_core_bytes(c, n, b, mixed, m);
  return c.off + 4096;
}

// Everything after S = U^T U: H' = X^T H X (FP64), the standard solve of H' (FP64, or FP32 and rounded), C = X Z for
// the first m columns, and for the mixed path the FP64 refinement against (H, S) and the check. x: X = U^-1 (n x n,
// ld ldx, zero below the diagonal); sm is read by the refinement only.
// pad (grouped solves): problems of their own size pad->n[p] < n padded to n; H' gets the padded block's diagonal
// (pad_spectrum) before the reduction, and H (writable there) the same diagonal for the refinement.
static void sygvd_core(Handle& h, Carver& c, int n, const double* hm, int ldh, const double* sm, int lds,
                       const double* x, int ldx, double* w, double* cm, int ldc, bool mixed, int b, int m,
                       const PadSpec* pad = nullptr) {
  hipStream_t st = h.stream;
  const size_t nn = (size_t)n * n;
  if (!mixed) {
    double* hx = c.take<double>(nn);
    double* hp = c.take<double>(nn);
    StdBuffers<double> s = carve_std<double>(c, n, b);
    gemm<double>(st, 'N', 'N', n, n, n, 1.0, hm, ldh, x, ldx, 0.0, hx, n);
    gemm_upper_tn<double>(st, n, n, 1.0, x, ldx, hx, n, 0.0, hp, n);  // X^T H X is symmetric: upper tiles, mirrored
    symmetrize<double>(st, 'U', n, hp, n);
    if (pad) pad_spectrum(st, n, pad->n, hp, n, nullptr, 0);
    // C = X Z = (X Q1 Q2) Zdc, X Q1 Q2 formed beside the divide and conquer (into hx, free by then)
    syevd_core<double>(h, s, n, b, hp, n, w, cm, ldc, x, ldx, hx, m);
  } else {
    float* x32 = c.take<float>(nn);
    float* pm = c.take<float>(nn);
    float* hp = c.take<float>(nn);
    float* c32 = c.take<float>(nn);
    StdBuffers<float> s = carve_std<float>(c, n, b);
    double* rwork = c.take<double>(refine_workspace(n));
    int* riwork = c.take<int>(refine_iwork(n));
    double* cfull = m < n ? c.take<double>(nn) : cm;
    const int ldfull = m < n ? n : ldc;
    unsigned long long* amax = reinterpret_cast<unsigned long long*>(c.take<double>(16));
    // H' = X^T H X in FP64, then rounded: formed in FP32, its rounding error grows with cond(S) (about
    // eps32 cond(S) ||H||), which at cond(S) = 1e4 leaves the FP32 eigenvectors too far off for the refinement to
    // converge; in FP64 the FP32 solve starts from eps32 ||H'|| whatever S. The two products are a small part of
    // the solve. H X and H' sit in the refinement's workspace, which is free until the refinement starts.
    double* hx64 = rwork;
    double* hp64 = rwork + nn;
    copy_matrix<double, float>(st, n, n, x, ldx, x32, n);
    gemm<double>(st, 'N', 'N', n, n, n, 1.0, hm, ldh, x, ldx, 0.0, hx64, n);
    gemm_upper_tn<double>(st, n, n, 1.0, x, ldx, hx64, n, 0.0, hp64, n);
    symmetrize<double>(st, 'U', n, hp64, n);
    if (pad) pad_spectrum(st, n, pad->n, hp64, n, pad->h, pad->ldh);
    // rounded to FP32 inside its range (a power-of-two scale for an H' near or past it, undone on the eigenvalues;
    // a non-finite H' is not solved in FP32 and fails the check)
    copy_to_fp32_ranged(st, n, hp64, n, hp, n, amax, h.dinfo + 3);
    // the refinement needs the whole basis: all n vectors are formed and refined, the first m returned
    syevd_core<float>(h, s, n, b, hp, n, w, c32, n, x32, n, pm);
    unscale_values(st, n, w, amax);
    copy_matrix<float, double>(st, n, n, c32, n, cfull, ldfull);
    refine_oa(st, n, hm, ldh, sm, lds, cfull, ldfull, w, h.refine_iters, rwork, riwork, h.dinfo + 3,
              h.check_ratios());
    if (m < n) copy_matrix<double, double>(st, n, m, cfull, ldfull, cm, ldc);
  }
}

void sygvd_device(Handle& h, int n, const double* hm, int ldh, const double* sm, int lds, double* w, double* cm,
                  int ldc, bool mixed, int b, const double** u_out) {
  if (n <= 0) return;
  if (b <= 0) b = h.band_for(n, mixed);
  b = std::max(1, std::min(b, std::max(1, n - 1)));
  check_size(n, b);
  h.reserve(sygvd_bytes(n, b, mixed, n));
  sygvd_in(h, h.arena, n, hm, ldh, sm, lds, w, cm, ldc, mixed, b, u_out, nullptr);
}

void sygvd_in(Handle& h, void* base, int n, const double* hm, int ldh, const double* sm, int lds, double* w,
              double* cm, int ldc, bool mixed, int b, const double** u_out, const PadSpec* pad) {
  hipStream_t st = h.stream;
  Carver c(base);
  const size_t nn = (size_t)n * n;
  double* scopy = c.take<double>(nn);
  double* u = c.take<double>(nn);
  double* x = c.take<double>(nn);
  double* tmp = c.take<double>(potrf_workspace(n));
  copy_matrix<double, double>(st, n, n, sm, lds, scopy, n);
  potrf_inverse(st, n, scopy, n, u, n, x, n, tmp, h.dinfo + 0, h.stream2, h.ev_band, h.ev_join);
  if (u_out) *u_out = u;  // no later stage carves over it
  sygvd_core(h, c, n, hm, ldh, sm, lds, x, n, w, cm, ldc, mixed, b, n, pad);
}

// ---- the geometry (S, U, X = U^-1) kept on the handle across the cycles of one geometry -----------------------------

void Handle::set_geometry(int n, const double* sm, int lds) {
  if (n <= 0) throw std::runtime_error("set_geometry: n must be positive");
  const size_t nn = (size_t)n * n;
  const size_t bytes = 3 * nn * sizeof(double);
  if (bytes > geo_bytes) {
    XTB_HIP(hipSetDevice(device));
    if (geo) {
      XTB_HIP(hipStreamSynchronize(stream));
      XTB_HIP(hipFree(geo));
      geo = nullptr;
      geo_bytes = 0;
    }
    XTB_HIP(hipMalloc(&geo, bytes));
    geo_bytes = bytes;
  }
  geo_n = 0;
  geo_s = geo;
  geo_u = geo + nn;
  geo_x = geo + 2 * nn;
  reserve(nn * sizeof(double) + potrf_workspace(n) * sizeof(double) + 4096);
  Carver c(arena);
  double* scopy = c.take<double>(nn);
  double* tmp = c.take<double>(potrf_workspace(n));
  copy_matrix<double, double>(stream, n, n, sm, lds, geo_s, n);
  copy_matrix<double, double>(stream, n, n, sm, lds, scopy, n);
  potrf_inverse(stream, n, scopy, n, geo_u, n, geo_x, n, tmp, dinfo + 0, stream2, ev_band, ev_join);
  geo_n = n;
}

void sygvd_geometry(Handle& h, int n, int m, const double* hm, int ldh, double* w, double* cm, int ldc, bool mixed,
                    int b) {
  if (n <= 0) return;
  if (h.geo_n != n) throw std::runtime_error("sygvd_geometry: no geometry of this n on the handle (set it first)");
  if (m < 1 || m > n) throw std::runtime_error("sygvd_geometry: need 1 <= m <= n");
  if (b <= 0) b = h.band_for(n, mixed);
  b = std::max(1, std::min(b, std::max(1, n - 1)));
  check_size(n, b);
  Carver c0(nullptr);
  carve_core_bytes(c0, n, b, mixed, m);
  h.reserve(c0.off + 4096);
  Carver c(h.arena);
  sygvd_core(h, c, n, hm, ldh, h.geo_s, n, h.geo_x, n, w, cm, ldc, mixed, b, m);
}

void congruence_device(Handle& h, int n, const double* hm, int ldh, const double* x, int ldx, double* z, int ldz) {
  if (n <= 0) return;
  const size_t nn = (size_t)n * n;
  h.reserve(2 * nn * sizeof(double) + 4096);
  Carver c(h.arena);
  double* hx = c.take<double>(nn);
  double* zz = c.take<double>(nn);
  gemm<double>(h.stream, 'N', 'N', n, n, n, 1.0, hm, ldh, x, ldx, 0.0, hx, n);
  gemm_upper_tn<double>(h.stream, n, n, 1.0, x, ldx, hx, n, 0.0, zz, n);
  copy_upper<double>(h.stream, n, zz, n, z, ldz);
}

void back_solve_device(Handle& h, int n, int m, const double* x, int ldx, double* cm, int ldc) {
  if (n <= 0 || m <= 0) return;
  if (m > n) throw std::runtime_error("back_solve: need m <= n");
  h.reserve((size_t)n * m * sizeof(double) + 4096);
  double* t = static_cast<double*>(h.arena);
  gemm<double>(h.stream, 'N', 'N', n, m, n, 1.0, x, ldx, cm, ldc, 0.0, t, n);
  copy_matrix<double, double>(h.stream, n, m, t, n, cm, ldc);
}

void stedc_device(Handle& h, int n, double* d, double* e, double* z, int ldz) {
  if (n <= 0) return;
  Carver c0(nullptr);
  c0.take<double>(dc_workspace_doubles(n));
  c0.take<int>(dc_workspace_ints(n));
  h.reserve(c0.off + 4096);
  Carver c(h.arena);
  double* work = c.take<double>(dc_workspace_doubles(n));
  int* iwork = c.take<int>(dc_workspace_ints(n));
  stedc(h.stream, h.plan(n), d, e, z, ldz, work, iwork, h.dinfo + 2, h.test_secular_maxit);
}

template void syevd_device<double>(Handle&, int, const double*, int, double*, double*, int, int, int);
template void syevd_device<float>(Handle&, int, const float*, int, double*, float*, int, int, int);

}  // namespace xtbeig
