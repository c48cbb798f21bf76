// The C ABI (include/xtbeig.h). Host-memory entry points run on a per-thread handle (no process-global mutable state:
// two threads of one process never share a handle or workspace); device-memory entry points on the caller's handle.
// Every entry point catches internal errors and reports them (return code, LAPACK info, xtbeig_last_error).
#include "../../include/xtbeig.h"
#include "solver.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>
#include <algorithm>

using xtbeig::Handle;

struct xtbeig_handle {
  Handle h;
  explicit xtbeig_handle(int dev) : h(dev) {}
};

namespace {

thread_local std::string tl_error;
thread_local std::unique_ptr<xtbeig_handle> tl_handle;

int set_error(const char* what) {
  tl_error = what;
  return 1;
}

Handle& thread_handle() {
  if (!tl_handle) {
    int dev = 0;
    if (const char* e = std::getenv("XTBEIG_DEVICE")) dev = std::atoi(e);
    tl_handle.reset(new xtbeig_handle(dev));
  }
  return tl_handle->h;
}

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    return set_error(e.what());
  } catch (...) {
    return set_error("unknown error");
  }
}

// Device status words -> host, cleared.
void read_status(Handle& h, int* st) {
  XTB_HIP(hipStreamSynchronize(h.stream));
  XTB_HIP(hipMemcpy(st, h.dinfo, xtbeig::NSTATUS * sizeof(int), hipMemcpyDeviceToHost));
  XTB_HIP(hipMemsetAsync(h.dinfo, 0, xtbeig::NSTATUS * sizeof(int), h.stream));
  XTB_HIP(hipStreamSynchronize(h.stream));
}

// A device buffer of a host-array call: the next of the calling thread's handle slots, kept between calls (a
// hipMalloc / hipFree pair per call cost more than the transfers). Buffers are released in reverse order.
struct DevBuf {
  void* p = nullptr;
  explicit DevBuf(size_t bytes) {
    Handle& h = thread_handle();
    p = h.io_slot(h.io_top++, bytes > 0 ? bytes : 1);
  }
  ~DevBuf() { --thread_handle().io_top; }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  template <typename T> T* as() { return static_cast<T*>(p); }
};

void h2d(Handle& h, void* dst, const void* src, size_t bytes) {
  XTB_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h.stream));
}
void d2h(Handle& h, void* dst, const void* src, size_t bytes) {
  XTB_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h.stream));
}
// Host copy of n columns (elements i0(j) ..< i1(j) of column j) between column-major arrays, split over up to 8
// host threads when it is large: one thread moves about 10 GB/s, the pinned-to-device transfer 50.
template <typename T, typename R>
void host_copy_cols(T* dst, size_t ldd, const T* src, size_t lds, int n, R range) {
  auto part = [&](int j0, int j1) {
    for (int j = j0; j < j1; ++j) {
      int i0, i1;
      range(j, i0, i1);
      if (i1 > i0) std::memcpy(dst + (size_t)j * ldd + i0, src + (size_t)j * lds + i0, sizeof(T) * (size_t)(i1 - i0));
    }
  };
  int i0, i1;
  range(n - 1, i0, i1);
  const size_t bytes = sizeof(T) * (size_t)n * (size_t)(i1 > i0 ? i1 - i0 : 1);
  const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
  const int nt = (int)std::min<size_t>({(size_t)8, (size_t)hw, bytes / (1u << 20), (size_t)n});
  if (nt <= 1) {
    part(0, n);
    return;
  }
  std::vector<std::thread> pool;
  for (int q = 1; q < nt; ++q) pool.emplace_back(part, (int)((int64_t)n * q / nt), (int)((int64_t)n * (q + 1) / nt));
  part(0, (int)((int64_t)n / nt));
  for (auto& th : pool) th.join();
}

// Column-major sub-matrix copies with leading dimensions. Small contiguous ones: one copy from the caller's memory.
// The rest go through the handle's pinned staging buffer, packed or unpacked on the host (by several threads when
// large) and moved in one transfer: a pitched copy from pageable memory moves column by column, and a large copy
// from pageable memory runs at the single-threaded staging rate.
constexpr size_t STAGED_BYTES = 1u << 20;
template <typename T>
void h2d_mat(Handle& h, T* dst, int ldd, const T* src, int lds, int m, int n) {
  if (m <= 0 || n <= 0) return;
  const size_t bytes = sizeof(T) * (size_t)m * n;
  if (lds == m && ldd == m && bytes < STAGED_BYTES) {
    h2d(h, dst, src, bytes);
    return;
  }
  XTB_HIP(hipStreamSynchronize(h.stream));  // the staging buffer may still feed an earlier copy
  T* stage = static_cast<T*>(h.host_staging(bytes));
  host_copy_cols(stage, (size_t)m, src, (size_t)lds, n, [m](int, int& i0, int& i1) { i0 = 0; i1 = m; });
  XTB_HIP(hipMemcpy2DAsync(dst, sizeof(T) * ldd, stage, sizeof(T) * m, sizeof(T) * m, n, hipMemcpyHostToDevice,
                           h.stream));
  XTB_HIP(hipStreamSynchronize(h.stream));
}
template <typename T>
void d2h_mat(Handle& h, T* dst, int ldd, const T* src, int lds, int m, int n) {
  if (m <= 0 || n <= 0) return;
  const size_t bytes = sizeof(T) * (size_t)m * n;
  if (lds == m && ldd == m && bytes < STAGED_BYTES) {
    d2h(h, dst, src, bytes);
    return;
  }
  XTB_HIP(hipStreamSynchronize(h.stream));
  T* stage = static_cast<T*>(h.host_staging(bytes));
  XTB_HIP(hipMemcpy2DAsync(stage, sizeof(T) * m, src, sizeof(T) * lds, sizeof(T) * m, n, hipMemcpyDeviceToHost,
                           h.stream));
  XTB_HIP(hipStreamSynchronize(h.stream));
  host_copy_cols(dst, (size_t)ldd, stage, (size_t)m, n, [m](int, int& i0, int& i1) { i0 = 0; i1 = m; });
}

bool is_char(const char* c, char u) { return c && (*c == u || *c == (char)(u + 32)); }

// The referenced triangle only (upper: i <= j; lower: i >= j) of the device n x n matrix (ld n) into host memory,
// the other triangle of the host matrix left as it is. Synchronizes the handle's stream.
template <typename T>
void d2h_triangle(Handle& h, bool upper, T* dst, int ldd, const T* src, int n) {
  XTB_HIP(hipStreamSynchronize(h.stream));
  T* stage = static_cast<T*>(h.host_staging(sizeof(T) * (size_t)n * n));
  XTB_HIP(hipMemcpyAsync(stage, src, sizeof(T) * (size_t)n * n, hipMemcpyDeviceToHost, h.stream));
  XTB_HIP(hipStreamSynchronize(h.stream));
  host_copy_cols(dst, (size_t)ldd, stage, (size_t)n, n, [upper, n](int j, int& i0, int& i1) {
    i0 = upper ? 0 : j;
    i1 = upper ? j + 1 : n;
  });
}

}  // namespace

template <typename T>
int reduce_to_band_impl(xtbeig_handle_t handle, int n, int b, T* a, int lda, T* v1, T* t1, T* ab) {
  return guarded([&] {
    Handle& h = handle->h;
    h.reserve(xtbeig::band_workspace<T>(n, b) * sizeof(T) + 4096);
    xtbeig::reduce_to_band<T>(h.stream, n, b, a, lda, v1, t1, ab, static_cast<T*>(h.arena));
  });
}

// The back-transformations work on Z^T: the column-major z is transposed into the arena and back.
template <typename T>
int back_band_impl(xtbeig_handle_t handle, int n, int b, const T* v1, const T* t1, T* z, int ldz, int m) {
  return guarded([&] {
    Handle& h = handle->h;
    if (n <= 0 || m <= 0) return;
    const xtbeig::BackPlan& p = h.backplan(n, b);
    const int ldzt = xtbeig::zt_ld(m);
    const size_t zt_elems = (size_t)ldzt * xtbeig::zt_rows(n);
    h.reserve((zt_elems + xtbeig::back_band_workspace<T>(n, b, m)) * sizeof(T) + 8192);
    T* zt = static_cast<T*>(h.arena);
    xtbeig::transpose_matrix<T, T>(h.stream, n, m, z, ldz, zt, ldzt);
    xtbeig::back_transform_band<T>(h.stream, p, v1, t1, zt, ldzt, m, zt + zt_elems);
    xtbeig::transpose_matrix<T, T>(h.stream, m, n, zt, ldzt, z, ldz);
  });
}

template <typename T>
int back_bulges_impl(xtbeig_handle_t handle, int n, int b, const T* v2, T* z, int ldz, int m) {
  return guarded([&] {
    Handle& h = handle->h;
    if (n <= 0 || m <= 0) return;
    const xtbeig::BackPlan& p = h.backplan(n, b);
    const int ldzt = xtbeig::zt_ld(m);
    h.reserve((size_t)ldzt * xtbeig::zt_rows(n) * sizeof(T) + 8192);
    T* zt = static_cast<T*>(h.arena);
    xtbeig::set_zero<T>(h.stream, ldzt, xtbeig::zt_rows(n) - n, zt + (size_t)n * ldzt, ldzt);
    xtbeig::transpose_matrix<T, T>(h.stream, n, m, z, ldz, zt, ldzt);
    xtbeig::back_transform_bulges<T>(h.stream, p, v2, zt, ldzt, m);
    xtbeig::transpose_matrix<T, T>(h.stream, m, n, zt, ldzt, z, ldz);
  });
}

template <typename T>
void syevd_host(const char* jobz, const char* uplo, const int* n_, T* a, const int* lda_, T* w, T* work,
                       const int* lwork, int* iwork, const int* liwork, int* info) {
  const int n = *n_, lda = *lda_;
  const bool wantz = is_char(jobz, 'V');
  const long long lmin = n <= 1 ? 1 : (wantz ? 1 + 6LL * n + 2LL * n * n : 2LL * n + 1);
  const long long limin = (n <= 1 || !wantz) ? 1 : 3 + 5LL * n;
  *info = 0;
  if (!wantz && !is_char(jobz, 'N')) *info = -1;
  else if (!is_char(uplo, 'U') && !is_char(uplo, 'L')) *info = -2;
  else if (n < 0) *info = -3;
  else if (lda < (n > 1 ? n : 1)) *info = -5;
  if (*info != 0) return;
  if (*lwork == -1 || *liwork == -1) {
    work[0] = (T)lmin;
    iwork[0] = (int)limin;
    return;
  }
  if (*lwork < lmin) { *info = -8; return; }
  if (*liwork < limin) { *info = -10; return; }
  if (n == 0) return;
  int st[xtbeig::NSTATUS] = {0, 0, 0, 0};
  const int rc = guarded([&] {
    Handle& h = thread_handle();
    DevBuf da(sizeof(T) * (size_t)n * n), dz(sizeof(T) * (size_t)n * n), dw(sizeof(double) * n),
        dwt(sizeof(T) * n);
    h2d_mat<T>(h, da.as<T>(), n, a, lda, n, n);
    xtbeig::symmetrize<T>(h.stream, is_char(uplo, 'U') ? 'U' : 'L', n, da.as<T>(), n);
    xtbeig::syevd_device<T>(h, n, da.as<T>(), n, dw.as<double>(), dz.as<T>(), n, h.band_override);
    xtbeig::copy_matrix<double, T>(h.stream, n, 1, dw.as<double>(), n, dwt.as<T>(), n);
    d2h(h, w, dwt.as<T>(), sizeof(T) * n);
    if (wantz) d2h_mat<T>(h, a, lda, dz.as<T>(), n, n, n);
    read_status(h, st);
  });
  if (rc) *info = n + 1;
  else if (st[1] || st[2]) *info = st[2] ? st[2] : n + 1;
}

// Generalized solve from host memory. T is the caller's type; the GPU works on FP64 copies (the FP32 interface runs
// the FP32 + FP64-refinement path and rounds the results).
template <typename T>
static void sygvd_host(bool mixed, const int* itype, const char* jobz, const char* uplo, const int* n_, T* a,
                       const int* lda_, T* b, const int* ldb_, T* w, T* work, const int* lwork, int* iwork,
                       const int* liwork, int* info) {
  const int n = *n_, lda = *lda_, ldb = *ldb_;
  const bool wantz = is_char(jobz, 'V');
  const long long lmin = n <= 1 ? 1 : (wantz ? 1 + 6LL * n + 2LL * n * n : 2LL * n + 1);
  const long long limin = (n <= 1 || !wantz) ? 1 : 3 + 5LL * n;
  *info = 0;
  if (*itype != 1) *info = -1;  // itype 2 / 3 (A B x, B A x) are not part of xTB's path
  else if (!wantz && !is_char(jobz, 'N')) *info = -2;
  else if (!is_char(uplo, 'U') && !is_char(uplo, 'L')) *info = -3;
  else if (n < 0) *info = -4;
  else if (lda < (n > 1 ? n : 1)) *info = -6;
  else if (ldb < (n > 1 ? n : 1)) *info = -8;
  if (*info != 0) return;
  if (*lwork == -1 || *liwork == -1) {
    work[0] = (T)lmin;
    iwork[0] = (int)limin;
    return;
  }
  if (*lwork < lmin) { *info = -11; return; }
  if (*liwork < limin) { *info = -13; return; }
  if (n == 0) return;
  int st[xtbeig::NSTATUS] = {0, 0, 0, 0};
  const int rc = guarded([&] {
    Handle& h = thread_handle();
    const size_t nn = (size_t)n * n;
    DevBuf dh(sizeof(double) * nn), ds(sizeof(double) * nn), dc(sizeof(double) * nn), dw(sizeof(double) * n),
        dx(sizeof(double) * nn), dt(sizeof(T) * nn);
    const char ul = is_char(uplo, 'U') ? 'U' : 'L';
    auto upload = [&](double* dst, const T* src, int ld) {
      if (std::is_same<T, double>::value) {
        h2d_mat<double>(h, dst, n, reinterpret_cast<const double*>(src), ld, n, n);
      } else {
        h2d_mat<T>(h, dt.as<T>(), n, src, ld, n, n);
        xtbeig::copy_matrix<T, double>(h.stream, n, n, dt.as<T>(), n, dst, n);
      }
    };
    auto download = [&](T* dst, int ld, const double* src, int rows, int cols, bool tri, bool upper) {
      const T* from;
      if (std::is_same<T, double>::value) {
        from = reinterpret_cast<const T*>(src);
      } else {
        xtbeig::copy_matrix<double, T>(h.stream, rows, cols, src, n, dt.as<T>(), n);
        from = dt.as<T>();
      }
      if (tri) d2h_triangle<T>(h, upper, dst, ld, from, n);
      else d2h_mat<T>(h, dst, ld, from, n, rows, cols);
    };
    upload(dh.as<double>(), a, lda);
    upload(ds.as<double>(), b, ldb);
    xtbeig::symmetrize<double>(h.stream, ul, n, dh.as<double>(), n);
    xtbeig::symmetrize<double>(h.stream, ul, n, ds.as<double>(), n);
    const double* uf = nullptr;  // the solve's own Cholesky factor of S, in the handle's arena
    xtbeig::sygvd_device(h, n, dh.as<double>(), n, ds.as<double>(), n, dw.as<double>(), dc.as<double>(), n, mixed,
                         h.band_override, &uf);
    if (mixed) {
      // an FP32 solve whose refined pairs fail the check, or whose chase or D&C failed, is redone in FP64
      read_status(h, st);
      if (st[0] == 0 && (st[1] || st[2] || st[3])) {
        xtbeig::sygvd_device(h, n, dh.as<double>(), n, ds.as<double>(), n, dw.as<double>(), dc.as<double>(), n,
                             false, 0, &uf);
        for (int k = 0; k < xtbeig::NSTATUS; ++k) st[k] = 0;
      }
    }
    download(w, n, dw.as<double>(), n, 1, false, false);
    if (wantz) download(a, lda, dc.as<double>(), n, n, false, false);
    int st2[xtbeig::NSTATUS] = {0, 0, 0, 0};
    read_status(h, st2);
    for (int k = 0; k < xtbeig::NSTATUS; ++k) st[k] = st[k] ? st[k] : st2[k];
    if (st[0] == 0) {
      // b on exit: the Cholesky factor in the referenced triangle (U, or L = U^T), the other triangle untouched,
      // as LAPACK's dsygvd leaves it
      if (ul == 'L') xtbeig::transpose_matrix<double, double>(h.stream, n, n, uf, n, dx.as<double>(), n);
      download(b, ldb, ul == 'U' ? uf : dx.as<double>(), n, n, true, ul == 'U');
    }
  });
  if (rc) *info = 2 * n + 1;
  else if (st[0]) *info = n + st[0];
  else if (st[1] || st[2]) *info = st[2] ? st[2] : n;
}

// One-stage reduction from host memory (upper triangle referenced; full: both triangles stored and returned).
// Returns guarded()'s code: nonzero when a HIP error, an allocation or a launch failure stopped the reduction (the
// message in xtbeig_last_error); the void xtb_* entry points can report it only there.
template <typename T>
static int sytrd_host(int n, T* a, int lda, T* d, T* e, T* tau, bool full) {
  if (n <= 0) return 0;
  return guarded([&] {
    Handle& h = thread_handle();
    const size_t nn = (size_t)n * n;
    h.reserve(xtbeig::sytrd_workspace<T>(n) * sizeof(T) + 4096);
    DevBuf da(sizeof(T) * nn), dd(sizeof(T) * n), de(sizeof(T) * n), dt(sizeof(T) * n);
    h2d_mat<T>(h, da.as<T>(), n, a, lda, n, n);
    xtbeig::sytrd_upper<T>(h.stream, n, da.as<T>(), n, dd.as<T>(), de.as<T>(), dt.as<T>(),
                           static_cast<T*>(h.arena), full);
    if (full) d2h_mat<T>(h, a, lda, da.as<T>(), n, n, n);
    else d2h_triangle<T>(h, true, a, lda, da.as<T>(), n);
    d2h(h, d, dd.as<T>(), sizeof(T) * n);
    if (n > 1) {
      d2h(h, e, de.as<T>(), sizeof(T) * (n - 1));
      d2h(h, tau, dt.as<T>(), sizeof(T) * (n - 1));
    }
    XTB_HIP(hipStreamSynchronize(h.stream));
  });
}

// z = Q z for m columns from host memory, Q from sytrd_host (a, tau).
template <typename T>
static int ormtr_host(int n, int m, const T* a, int lda, const T* tau, T* z, int ldz) {
  if (n <= 1 || m <= 0) return 0;
  return guarded([&] {
    Handle& h = thread_handle();
    h.reserve(xtbeig::ormtr_workspace<T>(n, m) * sizeof(T) + 4096);
    DevBuf da(sizeof(T) * (size_t)n * n), dt(sizeof(T) * n), dz(sizeof(T) * (size_t)n * m);
    h2d_mat<T>(h, da.as<T>(), n, a, lda, n, n);
    h2d(h, dt.as<T>(), tau, sizeof(T) * (n - 1));
    h2d_mat<T>(h, dz.as<T>(), n, z, ldz, n, m);
    xtbeig::ormtr_upper<T>(h.stream, n, m, da.as<T>(), n, dt.as<T>(), dz.as<T>(), n, static_cast<T*>(h.arena));
    d2h_mat<T>(h, z, ldz, dz.as<T>(), n, n, m);
    XTB_HIP(hipStreamSynchronize(h.stream));
  });
}

// LAPACK xsytrd: uplo 'U' only (the convention dormtr 'U' and the reference use); lwork >= 1, optimal n * 32.
template <typename T>
static void sytrd_lapack(const char* uplo, const int* n_, T* a, const int* lda_, T* d, T* e, T* tau, T* work,
                         const int* lwork, int* info) {
  const int n = *n_, lda = *lda_;
  *info = 0;
  if (!is_char(uplo, 'U')) *info = -1;
  else if (n < 0) *info = -2;
  else if (lda < (n > 1 ? n : 1)) *info = -4;
  else if (*lwork < 1 && *lwork != -1) *info = -9;
  if (*info != 0) return;
  if (*lwork == -1) {
    work[0] = (T)((n > 1 ? n : 1) * 32);
    return;
  }
  if (sytrd_host<T>(n, a, lda, d, e, tau, false) != 0) *info = n + 1;
}

extern "C" {

const char* xtbeig_last_error(void) { return tl_error.c_str(); }

int xtbeig_create(xtbeig_handle_t* handle, int device) {
  return guarded([&] { *handle = new xtbeig_handle(device); });
}

int xtbeig_destroy(xtbeig_handle_t handle) {
  return guarded([&] { delete handle; });
}

int xtbeig_set_stream(xtbeig_handle_t handle, void* stream) {
  return guarded([&] {
    Handle& h = handle->h;
    XTB_HIP(hipStreamSynchronize(h.stream));
    if (h.own_stream) XTB_HIP(hipStreamDestroy(h.stream));
    h.own_stream = false;
    h.stream = static_cast<hipStream_t>(stream);
  });
}

void* xtbeig_get_stream(xtbeig_handle_t handle) { return handle ? (void*)handle->h.stream : nullptr; }

int xtbeig_set_tuning(xtbeig_handle_t handle, int band, int refine_iterations) {
  if (!handle) return set_error("xtbeig_set_tuning: null handle");
  if (band < 0 || band > xtbeig::MAX_BAND || refine_iterations < 0 || refine_iterations > 8)
    return set_error(("xtbeig_set_tuning: band must be in [0, " + std::to_string(xtbeig::MAX_BAND) +
                      "] and refine_iterations in [0, 8]").c_str());
  handle->h.band_override = band;
  handle->h.refine_iters = refine_iterations;
  return 0;
}

int xtbeig_default_band(int n, int fp32) { return xtbeig::default_band(n, fp32 != 0); }

int xtbeig_set_test_limits(xtbeig_handle_t handle, int secular_maxit) {
  if (secular_maxit < -1) return set_error("xtbeig_set_test_limits: secular_maxit must be >= -1");
  return guarded([&] {
    Handle& h = handle ? handle->h : thread_handle();
    h.test_secular_maxit = secular_maxit;
    for (auto& sub : h.subs) sub->test_secular_maxit = secular_maxit;
  });
}

int xtbeig_synchronize(xtbeig_handle_t handle, int* status) {
  int st[xtbeig::NSTATUS] = {0, 0, 0, 0};
  int rc = guarded([&] { read_status(handle->h, st); });
  if (rc) return rc;
  if (status)
    for (int k = 0; k < xtbeig::NSTATUS; ++k) status[k] = st[k];
  return (st[0] || st[1] || st[2] || st[3]) ? 2 : 0;
}

int xtbeig_refine_check(xtbeig_handle_t handle, double* ratios) {
  if (!handle || !ratios) return set_error("xtbeig_refine_check: null argument");
  return guarded([&] {
    Handle& h = handle->h;
    XTB_HIP(hipStreamSynchronize(h.stream));
    XTB_HIP(hipMemcpy(ratios, h.check_ratios(), 2 * sizeof(double), hipMemcpyDeviceToHost));
  });
}

int xtbeig_dev_sygvd(xtbeig_handle_t handle, int precision, int n, const double* hm, int ldh, const double* sm,
                     int lds, double* w, double* c, int ldc) {
  if (n < 0 || ldh < (n > 1 ? n : 1) || lds < (n > 1 ? n : 1) || ldc < (n > 1 ? n : 1))
    return set_error("xtbeig_dev_sygvd: bad dimensions");
  if (precision != XTBEIG_FP64 && precision != XTBEIG_FP32_REFINE)
    return set_error("xtbeig_dev_sygvd: unknown precision");
  return guarded([&] {
    Handle& h = handle->h;
    xtbeig::sygvd_device(h, n, hm, ldh, sm, lds, w, c, ldc, precision == XTBEIG_FP32_REFINE, h.band_override);
  });
}

int xtbeig_dev_dsyevd(xtbeig_handle_t handle, int n, const double* a, int lda, double* w, double* z, int ldz) {
  return guarded([&] {
    Handle& h = handle->h;
    xtbeig::syevd_device<double>(h, n, a, lda, w, z, ldz, h.band_override);
  });
}

int xtbeig_dev_ssyevd(xtbeig_handle_t handle, int n, const float* a, int lda, double* w, float* z, int ldz) {
  return guarded([&] {
    Handle& h = handle->h;
    xtbeig::syevd_device<float>(h, n, a, lda, w, z, ldz, h.band_override);
  });
}

int xtbeig_dev_dsyevd_partial(xtbeig_handle_t handle, int n, int m, const double* a, int lda, double* w, double* z,
                              int ldz) {
  if (!handle) return set_error("xtbeig_dev_dsyevd_partial: null handle");
  if (n < 0 || m < (n > 0 ? 1 : 0) || m > n || lda < (n > 1 ? n : 1) || ldz < (n > 1 ? n : 1))
    return set_error("xtbeig_dev_dsyevd_partial: bad dimensions (need 1 <= m <= n)");
  return guarded([&] {
    Handle& h = handle->h;
    xtbeig::syevd_device<double>(h, n, a, lda, w, z, ldz, h.band_override, m);
  });
}

int xtbeig_dev_set_geometry(xtbeig_handle_t handle, int n, const double* s, int lds) {
  if (!handle) return set_error("xtbeig_dev_set_geometry: null handle");
  if (n < 1 || lds < n) return set_error("xtbeig_dev_set_geometry: bad dimensions");
  return guarded([&] { handle->h.set_geometry(n, s, lds); });
}

int xtbeig_dev_congruence(xtbeig_handle_t handle, int n, const double* hm, int ldh, double* z, int ldz) {
  if (!handle) return set_error("xtbeig_dev_congruence: null handle");
  if (n < 1 || ldh < n || ldz < n) return set_error("xtbeig_dev_congruence: bad dimensions");
  return guarded([&] {
    Handle& h = handle->h;
    if (h.geo_n != n) throw std::runtime_error("xtbeig_dev_congruence: no geometry of this n on the handle");
    xtbeig::congruence_device(h, n, hm, ldh, h.geo_x, n, z, ldz);
  });
}

int xtbeig_dev_back_solve(xtbeig_handle_t handle, int n, int m, double* c, int ldc) {
  if (!handle) return set_error("xtbeig_dev_back_solve: null handle");
  if (n < 1 || m < 0 || m > n || ldc < n) return set_error("xtbeig_dev_back_solve: bad dimensions");
  return guarded([&] {
    Handle& h = handle->h;
    if (h.geo_n != n) throw std::runtime_error("xtbeig_dev_back_solve: no geometry of this n on the handle");
    xtbeig::back_solve_device(h, n, m, h.geo_x, n, c, ldc);
  });
}

int xtbeig_dev_sygvd_geometry(xtbeig_handle_t handle, int precision, int n, int m, const double* hm, int ldh,
                              double* w, double* c, int ldc) {
  if (!handle) return set_error("xtbeig_dev_sygvd_geometry: null handle");
  if (precision != XTBEIG_FP64 && precision != XTBEIG_FP32_REFINE)
    return set_error("xtbeig_dev_sygvd_geometry: unknown precision");
  if (n < 1 || m < 1 || m > n || ldh < n || ldc < n) return set_error("xtbeig_dev_sygvd_geometry: bad dimensions");
  return guarded([&] {
    Handle& h = handle->h;
    xtbeig::sygvd_geometry(h, n, m, hm, ldh, w, c, ldc, precision == XTBEIG_FP32_REFINE, h.band_override);
  });
}

int xtbeig_dev_dstedc(xtbeig_handle_t handle, int n, double* d, double* e, double* z, int ldz) {
  return guarded([&] { xtbeig::stedc_device(handle->h, n, d, e, z, ldz); });
}

int xtbeig_dev_dgemm(xtbeig_handle_t handle, char ta, char tb, int m, int n, int k, double alpha, const double* a,
                     int lda, const double* b, int ldb, double beta, double* c, int ldc) {
  return guarded([&] { xtbeig::gemm<double>(handle->h.stream, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc); });
}

int xtbeig_dev_sgemm(xtbeig_handle_t handle, char ta, char tb, int m, int n, int k, float alpha, const float* a,
                     int lda, const float* b, int ldb, float beta, float* c, int ldc) {
  return guarded([&] { xtbeig::gemm<float>(handle->h.stream, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc); });
}

int xtbeig_dev_potrf_inverse(xtbeig_handle_t handle, int n, double* s, int lds, double* u, int ldu, double* x,
                             int ldx) {
  return guarded([&] {
    Handle& h = handle->h;
    h.reserve(xtbeig::potrf_workspace(n) * sizeof(double) + 4096);
    xtbeig::potrf_inverse(h.stream, n, s, lds, u, ldu, x, ldx, static_cast<double*>(h.arena), h.dinfo);
  });
}

int xtbeig_dev_potrf_inverse_noalloc(xtbeig_handle_t handle, int n, double* s, int lds, double* u, int ldu,
                                     double* x, int ldx) {
  if (!handle) return set_error("xtbeig_dev_potrf_inverse_noalloc: null handle");
  if (n < 0 || lds < (n > 1 ? n : 1) || ldu < (n > 1 ? n : 1) || ldx < (n > 1 ? n : 1))
    return set_error("xtbeig_dev_potrf_inverse_noalloc: bad dimensions");
  return guarded([&] { xtbeig::potrf_inverse_noalloc(handle->h, n, s, lds, u, ldu, x, ldx); });
}

long long xtbeig_band_elems(int n, int b) { return (long long)xtbeig::band_elems(n, b); }

int xtbeig_dev_dreduce_to_band(xtbeig_handle_t handle, int n, int b, double* a, int lda, double* v1, double* t1,
                               double* ab) {
  return reduce_to_band_impl<double>(handle, n, b, a, lda, v1, t1, ab);
}
int xtbeig_dev_sreduce_to_band(xtbeig_handle_t handle, int n, int b, float* a, int lda, float* v1, float* t1,
                               float* ab) {
  return reduce_to_band_impl<float>(handle, n, b, a, lda, v1, t1, ab);
}

int xtbeig_dev_dchase(xtbeig_handle_t handle, int n, int b, double* ab, double* v2, double* d, double* e) {
  return guarded([&] { xtbeig::chase_bulges<double>(handle->h.stream, n, b, ab, v2, d, e, handle->h.dinfo + 1); });
}
int xtbeig_dev_schase(xtbeig_handle_t handle, int n, int b, float* ab, float* v2, float* d, float* e) {
  return guarded([&] { xtbeig::chase_bulges<float>(handle->h.stream, n, b, ab, v2, d, e, handle->h.dinfo + 1); });
}
int xtbeig_dev_dback_bulges(xtbeig_handle_t handle, int n, int b, const double* v2, double* z, int ldz, int m) {
  return back_bulges_impl<double>(handle, n, b, v2, z, ldz, m);
}
int xtbeig_dev_sback_bulges(xtbeig_handle_t handle, int n, int b, const float* v2, float* z, int ldz, int m) {
  return back_bulges_impl<float>(handle, n, b, v2, z, ldz, m);
}
int xtbeig_dev_dback_band(xtbeig_handle_t handle, int n, int b, const double* v1, const double* t1, double* z,
                          int ldz, int m) {
  return back_band_impl<double>(handle, n, b, v1, t1, z, ldz, m);
}
int xtbeig_dev_sback_band(xtbeig_handle_t handle, int n, int b, const float* v1, const float* t1, float* z, int ldz,
                          int m) {
  return back_band_impl<float>(handle, n, b, v1, t1, z, ldz, m);
}

int xtbeig_dev_dc_merge(xtbeig_handle_t handle, int n, int n1, double rho, double* d, double* q, int ldq,
                        double* d_deflated, double* q2, int* k_ctot, double* dlamda, double* w, int* indxc,
                        double* rho_out) {
  return guarded([&] {
    Handle& h = handle->h;
    h.reserve((xtbeig::dc_workspace_doubles(n) + 64) * sizeof(double) + (xtbeig::dc_workspace_ints(n) + 64) * 4 +
              8192);
    double* work = static_cast<double*>(h.arena);
    int* iwork = reinterpret_cast<int*>(work + xtbeig::dc_workspace_doubles(n) + 64);
    xtbeig::dc_merge_probe(h.stream, n, n1, rho, d, q, ldq, work, iwork, h.dinfo + 2, d_deflated, q2, k_ctot, dlamda,
                           w, indxc, rho_out);
  });
}

// ---- host-memory entry points -----------------------------------------------------------------------------------

// The void entry points of the sytrd.c symbol set refuse leading dimensions below n before touching the caller's
// arrays (the message in xtbeig_last_error; the reference's calls have no error path).
static bool lds_ok(const char* who, int n, int lda) {
  if (n > 0 && lda < n) {
    set_error((std::string(who) + ": leading dimension must be >= n").c_str());
    return false;
  }
  return true;
}

void xtb_dstedc(const int* n_, double* d, double* e, double* z, const int* ldz_, double* work, const int* lwork,
                int* iwork, const int* liwork, int* info) {
  const int n = *n_, ldz = *ldz_;
  const long long lmin = n <= 1 ? 1 : 1 + 4LL * n + (long long)n * n, limin = n <= 1 ? 1 : 3 + 5LL * n;
  if (*lwork == -1 || *liwork == -1) {
    work[0] = (double)lmin;
    iwork[0] = (int)limin;
    *info = 0;
    return;
  }
  if (n < 0) { *info = -2; return; }
  if (ldz < (n > 1 ? n : 1)) { *info = -5; return; }
  if (*lwork < lmin) { *info = -7; return; }
  if (*liwork < limin) { *info = -9; return; }
  *info = 0;
  if (n == 0) return;
  int st[xtbeig::NSTATUS] = {0, 0, 0, 0};
  const int rc = guarded([&] {
    Handle& h = thread_handle();
    DevBuf dd(sizeof(double) * n), de(sizeof(double) * n), dz(sizeof(double) * (size_t)n * n);
    h2d(h, dd.as<double>(), d, sizeof(double) * n);
    if (n > 1) h2d(h, de.as<double>(), e, sizeof(double) * (n - 1));
    xtbeig::stedc_device(h, n, dd.as<double>(), de.as<double>(), dz.as<double>(), n);
    d2h(h, d, dd.as<double>(), sizeof(double) * n);
    d2h_mat<double>(h, z, ldz, dz.as<double>(), n, n, n);
    read_status(h, st);
  });
  if (rc) *info = n + 1;
  else if (st[2]) *info = st[2];
}

void xtb_sstedc(const int* n_, float* d, float* e, float* z, const int* ldz_, float* work, const int* lwork,
                int* iwork, const int* liwork, int* info) {
  const int n = *n_, ldz = *ldz_;
  const long long lmin = n <= 1 ? 1 : 1 + 4LL * n + (long long)n * n, limin = n <= 1 ? 1 : 3 + 5LL * n;
  if (*lwork == -1 || *liwork == -1) {
    work[0] = (float)lmin;
    iwork[0] = (int)limin;
    *info = 0;
    return;
  }
  if (n < 0) { *info = -2; return; }
  if (ldz < (n > 1 ? n : 1)) { *info = -5; return; }
  if (*lwork < lmin) { *info = -7; return; }
  if (*liwork < limin) { *info = -9; return; }
  *info = 0;
  if (n == 0) return;
  int st[xtbeig::NSTATUS] = {0, 0, 0, 0};
  const int rc = guarded([&] {
    Handle& h = thread_handle();
    DevBuf df(sizeof(float) * n * 2), dd(sizeof(double) * n), de(sizeof(double) * n),
        dz(sizeof(double) * (size_t)n * n), dzf(sizeof(float) * (size_t)n * n);
    float* fd = df.as<float>();
    float* fe = fd + n;
    h2d(h, fd, d, sizeof(float) * n);
    if (n > 1) h2d(h, fe, e, sizeof(float) * (n - 1));
    xtbeig::copy_matrix<float, double>(h.stream, n, 1, fd, n, dd.as<double>(), n);
    if (n > 1) xtbeig::copy_matrix<float, double>(h.stream, n - 1, 1, fe, n, de.as<double>(), n);
    xtbeig::stedc_device(h, n, dd.as<double>(), de.as<double>(), dz.as<double>(), n);
    xtbeig::copy_matrix<double, float>(h.stream, n, 1, dd.as<double>(), n, fd, n);
    xtbeig::copy_matrix<double, float>(h.stream, n, n, dz.as<double>(), n, dzf.as<float>(), n);
    d2h(h, d, fd, sizeof(float) * n);
    d2h_mat<float>(h, z, ldz, dzf.as<float>(), n, n, n);
    read_status(h, st);
  });
  if (rc) *info = n + 1;
  else if (st[2]) *info = st[2];
}

void xtb_dsymmetrize_pad(const int* n_, const double* src, const int* lds_, double* dst, const int* ldd_) {
  const int n = *n_, lds = *lds_, ldd = *ldd_;
  if (n <= 0) return;
  if (lds < n || ldd < n) {  // nothing is written (the reference's call has no error path: the message is all)
    set_error("xtb_dsymmetrize_pad: leading dimensions must be >= n");
    return;
  }
  guarded([&] {
    Handle& h = thread_handle();
    DevBuf ds(sizeof(double) * (size_t)n * n), dt(sizeof(double) * (size_t)n * n);
    h2d_mat<double>(h, ds.as<double>(), n, src, lds, n, n);
    xtbeig::symmetrize_pad(h.stream, n, ds.as<double>(), n, dt.as<double>(), n);
    d2h_mat<double>(h, dst, ldd, dt.as<double>(), n, n, n);
    XTB_HIP(hipStreamSynchronize(h.stream));
  });
}


void xtb_dcongruence_upper(const int* n_, const double* hm, const int* ldh_, const double* x, const int* ldx_,
                           double* z, const int* ldz_) {
  const int n = *n_, ldh = *ldh_, ldx = *ldx_, ldz = *ldz_;
  if (n <= 0) return;
  if (!lds_ok("xtb_dcongruence_upper", n, ldh) || !lds_ok("xtb_dcongruence_upper", n, ldx) ||
      !lds_ok("xtb_dcongruence_upper", n, ldz))
    return;
  guarded([&] {
    Handle& h = thread_handle();
    const size_t nn = (size_t)n * n;
    DevBuf dh(sizeof(double) * nn), dx(sizeof(double) * nn), dz(sizeof(double) * nn);
    h2d_mat<double>(h, dh.as<double>(), n, hm, ldh, n, n);
    h2d_mat<double>(h, dx.as<double>(), n, x, ldx, n, n);
    xtbeig::congruence_device(h, n, dh.as<double>(), n, dx.as<double>(), n, dz.as<double>(), n);
    d2h_triangle<double>(h, true, z, ldz, dz.as<double>(), n);
  });
}

void xtb_dtrmm_back(const int* n_, const int* m_, const double* x, const int* ldx_, double* c, const int* ldc_) {
  const int n = *n_, m = *m_, ldx = *ldx_, ldc = *ldc_;
  if (n <= 0 || m <= 0) return;
  if (m > n) {
    set_error("xtb_dtrmm_back: need m <= n");
    return;
  }
  if (!lds_ok("xtb_dtrmm_back", n, ldx) || !lds_ok("xtb_dtrmm_back", n, ldc)) return;
  guarded([&] {
    Handle& h = thread_handle();
    DevBuf dx(sizeof(double) * (size_t)n * n), dc(sizeof(double) * (size_t)n * m);
    h2d_mat<double>(h, dx.as<double>(), n, x, ldx, n, n);
    h2d_mat<double>(h, dc.as<double>(), n, c, ldc, n, m);
    xtbeig::back_solve_device(h, n, m, dx.as<double>(), n, dc.as<double>(), n);
    d2h_mat<double>(h, c, ldc, dc.as<double>(), n, n, m);
    XTB_HIP(hipStreamSynchronize(h.stream));
  });
}

void xtbeig_dsyevd(const char* jobz, const char* uplo, const int* n, double* a, const int* lda, double* w,
                   double* work, const int* lwork, int* iwork, const int* liwork, int* info) {
  syevd_host<double>(jobz, uplo, n, a, lda, w, work, lwork, iwork, liwork, info);
}
void xtbeig_ssyevd(const char* jobz, const char* uplo, const int* n, float* a, const int* lda, float* w, float* work,
                   const int* lwork, int* iwork, const int* liwork, int* info) {
  syevd_host<float>(jobz, uplo, n, a, lda, w, work, lwork, iwork, liwork, info);
}

void xtbeig_dsygvd(const int* itype, const char* jobz, const char* uplo, const int* n, double* a, const int* lda,
                   double* b, const int* ldb, double* w, double* work, const int* lwork, int* iwork,
                   const int* liwork, int* info) {
  sygvd_host<double>(false, itype, jobz, uplo, n, a, lda, b, ldb, w, work, lwork, iwork, liwork, info);
}

void xtbeig_dsygvd_mixed(const int* itype, const char* jobz, const char* uplo, const int* n, double* a,
                         const int* lda, double* b, const int* ldb, double* w, double* work, const int* lwork,
                         int* iwork, const int* liwork, int* info) {
  sygvd_host<double>(true, itype, jobz, uplo, n, a, lda, b, ldb, w, work, lwork, iwork, liwork, info);
}

void xtbeig_ssygvd(const int* itype, const char* jobz, const char* uplo, const int* n, float* a, const int* lda,
                   float* b, const int* ldb, float* w, float* work, const int* lwork, int* iwork, const int* liwork,
                   int* info) {
  sygvd_host<float>(true, itype, jobz, uplo, n, a, lda, b, ldb, w, work, lwork, iwork, liwork, info);
}

void xtb_dsytrd_upper(const int* n, double* a, const int* lda, double* d, double* e, double* tau, double* work) {
  (void)work;
  if (lds_ok("xtb_dsytrd_upper", *n, *lda)) sytrd_host<double>(*n, a, *lda, d, e, tau, false);
}
void xtb_ssytrd_upper(const int* n, float* a, const int* lda, float* d, float* e, float* tau, float* work) {
  (void)work;
  if (lds_ok("xtb_ssytrd_upper", *n, *lda)) sytrd_host<float>(*n, a, *lda, d, e, tau, false);
}
void xtb_dsytrd_full(const int* n, double* a, const int* lda, double* d, double* e, double* tau, double* work) {
  (void)work;
  if (lds_ok("xtb_dsytrd_full", *n, *lda)) sytrd_host<double>(*n, a, *lda, d, e, tau, true);
}
void xtb_dormtr_upper(const int* n, const int* m, const double* a, const int* lda, const double* tau, double* z,
                      const int* ldz, double* work) {
  (void)work;
  if (lds_ok("xtb_dormtr_upper", *n, *lda) && lds_ok("xtb_dormtr_upper", *n, *ldz))
    ormtr_host<double>(*n, *m, a, *lda, tau, z, *ldz);
}
void xtb_sormtr_upper(const int* n, const int* m, const float* a, const int* lda, const float* tau, float* z,
                      const int* ldz, float* work) {
  (void)work;
  if (lds_ok("xtb_sormtr_upper", *n, *lda) && lds_ok("xtb_sormtr_upper", *n, *ldz))
    ormtr_host<float>(*n, *m, a, *lda, tau, z, *ldz);
}
void xtbeig_dsytrd(const char* uplo, const int* n, double* a, const int* lda, double* d, double* e, double* tau,
                   double* work, const int* lwork, int* info) {
  sytrd_lapack<double>(uplo, n, a, lda, d, e, tau, work, lwork, info);
}
void xtbeig_ssytrd(const char* uplo, const int* n, float* a, const int* lda, float* d, float* e, float* tau,
                   float* work, const int* lwork, int* info) {
  sytrd_lapack<float>(uplo, n, a, lda, d, e, tau, work, lwork, info);
}

int xtbeig_dev_dsytrd_upper(xtbeig_handle_t handle, int n, double* a, int lda, double* d, double* e, double* tau,
                            int full) {
  return guarded([&] {
    Handle& h = handle->h;
    h.reserve(xtbeig::sytrd_workspace<double>(n) * sizeof(double) + 4096);
    xtbeig::sytrd_upper<double>(h.stream, n, a, lda, d, e, tau, static_cast<double*>(h.arena), full != 0);
  });
}
int xtbeig_dev_ssytrd_upper(xtbeig_handle_t handle, int n, float* a, int lda, float* d, float* e, float* tau,
                            int full) {
  return guarded([&] {
    Handle& h = handle->h;
    h.reserve(xtbeig::sytrd_workspace<float>(n) * sizeof(float) + 4096);
    xtbeig::sytrd_upper<float>(h.stream, n, a, lda, d, e, tau, static_cast<float*>(h.arena), full != 0);
  });
}
int xtbeig_dev_dormtr_upper(xtbeig_handle_t handle, int n, int m, const double* a, int lda, const double* tau,
                            double* z, int ldz) {
  return guarded([&] {
    Handle& h = handle->h;
    h.reserve(xtbeig::ormtr_workspace<double>(n, m) * sizeof(double) + 4096);
    xtbeig::ormtr_upper<double>(h.stream, n, m, a, lda, tau, z, ldz, static_cast<double*>(h.arena));
  });
}
int xtbeig_dev_sormtr_upper(xtbeig_handle_t handle, int n, int m, const float* a, int lda, const float* tau,
                            float* z, int ldz) {
  return guarded([&] {
    Handle& h = handle->h;
    h.reserve(xtbeig::ormtr_workspace<float>(n, m) * sizeof(float) + 4096);
    xtbeig::ormtr_upper<float>(h.stream, n, m, a, lda, tau, z, ldz, static_cast<float*>(h.arena));
  });
}

int xtbeig_dev_density(xtbeig_handle_t handle, int batch, int n, const double* w, long long sw, const double* c,
                       int ldc, long long sc, double kT, double nel, double* mu, int bisect, double* p, int ldp,
                       long long sp, double* occ) {
  return guarded([&] {
    Handle& h = handle->h;
    h.reserve_aux(xtbeig::density_workspace(batch, n) * sizeof(double));
    xtbeig::density_forward(h.stream, batch, n, 1, w, sw, c, ldc, sc, kT, nel, 0.0, bisect, mu, p, ldp, sp, occ,
                            static_cast<double*>(h.aux));
  });
}

int xtbeig_dev_density_backward(xtbeig_handle_t handle, int batch, int n, const double* w, long long sw,
                                const double* c, int ldc, long long sc, double kT, const double* mu, int mu_fixed,
                                const double* occ, const double* gbar, int ldg, long long sg, double* hbar, int ldhb,
                                long long shb, double* sbar, int ldsb, long long ssb) {
  return guarded([&] {
    Handle& h = handle->h;
    h.reserve_aux(xtbeig::density_workspace(batch, n) * sizeof(double));
    xtbeig::density_backward(h.stream, batch, n, 1, w, sw, c, ldc, sc, kT, mu, mu_fixed, occ, gbar, ldg, sg, hbar, ldhb,
                             shb, sbar, ldsb, ssb, static_cast<double*>(h.aux));
  });
}

int xtbeig_dev_density_spin(xtbeig_handle_t handle, int batch, int n, const double* w, long long sw, const double* c,
                            int ldc, long long sc, double kT, double nel_alpha, double nel_beta, double* mu, int bisect,
                            double* p, int ldp, long long sp, double* occ) {
  if (!handle) return set_error("xtbeig_dev_density_spin: null handle");
  if (!(nel_alpha >= 0) || !(nel_beta >= 0) || nel_alpha > n || nel_beta > n)
    return set_error("xtbeig_dev_density_spin: electron counts must be in [0, n]");
  return guarded([&] {
    Handle& h = handle->h;
    h.reserve_aux(xtbeig::density_workspace(batch, n) * sizeof(double));
    xtbeig::density_forward(h.stream, batch, n, 2, w, sw, c, ldc, sc, kT, nel_alpha, nel_beta, bisect, mu, p, ldp, sp,
                            occ, static_cast<double*>(h.aux));
  });
}

int xtbeig_dev_density_backward_spin(xtbeig_handle_t handle, int batch, int n, const double* w, long long sw,
                                     const double* c, int ldc, long long sc, double kT, const double* mu,
                                     int mu_fixed, const double* occ, const double* gbar, int ldg, long long sg,
                                     double* hbar, int ldhb, long long shb, double* sbar, int ldsb, long long ssb) {
  if (!handle) return set_error("xtbeig_dev_density_backward_spin: null handle");
  return guarded([&] {
    Handle& h = handle->h;
    h.reserve_aux(xtbeig::density_workspace(batch, n) * sizeof(double));
    xtbeig::density_backward(h.stream, batch, n, 2, w, sw, c, ldc, sc, kT, mu, mu_fixed, occ, gbar, ldg, sg, hbar,
                             ldhb, shb, sbar, ldsb, ssb, static_cast<double*>(h.aux));
  });
}

int xtbeig_dev_density_backward_recompute(xtbeig_handle_t handle, int batch, int n, const double* hm, int ldh,
                                          long long sh, const double* sm, int lds, long long ss, int precision,
                                          double kT, double nel, int mu_fixed, const double* gbar, int ldg,
                                          long long sg, double* hbar, int ldhb, long long shb, double* sbar, int ldsb,
                                          long long ssb) {
  if (precision != XTBEIG_FP64 && precision != XTBEIG_FP32_REFINE)
    return set_error("xtbeig_dev_density_backward_recompute: unknown precision");
  return guarded([&] {
    Handle& h = handle->h;
    // one problem at a time: eigenpairs, occupations (mu bisected to nel) and the backward in O(n^2) memory
    const size_t nn = (size_t)n * n;
    const size_t per = nn + 4 * (size_t)n + 8;
    h.reserve_aux((xtbeig::density_workspace(1, n) + per + 64) * sizeof(double));
    double* work = static_cast<double*>(h.aux);
    double* cbuf = work + xtbeig::density_workspace(1, n);
    double* wbuf = cbuf + nn;
    double* occ = wbuf + n;
    double* mu = occ + 3 * (size_t)n;
    for (int q = 0; q < batch; ++q) {
      xtbeig::sygvd_device(h, n, hm + (int64_t)q * sh, ldh, sm + (int64_t)q * ss, lds, wbuf, cbuf, n,
                           precision == XTBEIG_FP32_REFINE, h.band_override);
      // the occupations need mu: a forward pass without P (p = nullptr is not allowed; P goes to the backward's
      // first scratch matrix and is overwritten there)
      xtbeig::density_forward(h.stream, 1, n, 1, wbuf, n, cbuf, n, (int64_t)nn, kT, nel, 0.0, 1, mu, work + nn, n,
                              (int64_t)nn, occ, work);
      xtbeig::density_backward(h.stream, 1, n, 1, wbuf, n, cbuf, n, (int64_t)nn, kT, mu, mu_fixed, occ,
                               gbar + (int64_t)q * sg, ldg, sg, hbar + (int64_t)q * shb, ldhb, shb,
                               sbar + (int64_t)q * ssb, ldsb, ssb, work);
    }
  });
}

int xtbeig_set_scc_precision(xtbeig_handle_t handle, double fp32_rmsq, int fp32_maxiter, int fp32_refine) {
  if (!handle) return set_error("xtbeig_set_scc_precision: null handle");
  if (!std::isfinite(fp32_rmsq) || fp32_rmsq < 0)
    return set_error("xtbeig_set_scc_precision: fp32_rmsq must be finite and >= 0 (0: FP64 only)");
  if (fp32_maxiter < 1) return set_error("xtbeig_set_scc_precision: fp32_maxiter must be >= 1");
  if (fp32_refine != 0 && fp32_refine != 1) return set_error("xtbeig_set_scc_precision: fp32_refine must be 0 or 1");
  if (fp32_refine == 1 && fp32_rmsq == 0)
    return set_error("xtbeig_set_scc_precision: fp32_refine without FP32 (fp32_rmsq = 0)");
  Handle& h = handle->h;
  h.scc_fp32_rmsq = fp32_rmsq;
  h.scc_fp32_maxiter = fp32_rmsq > 0 ? fp32_maxiter : 0;
  h.scc_fp32_refine = fp32_refine;
  return 0;
}

int xtbeig_scc_begin(xtbeig_handle_t handle) {
  if (!handle) return set_error("xtbeig_scc_begin: null handle");
  Handle& h = handle->h;
  h.scc_fp32_cycles = h.scc_fp64_restart = h.scc_cycles = 0;
  h.scc_fp64_latched = false;
  return 0;
}

int xtbeig_dev_scc_sygvd(xtbeig_handle_t handle, int n, const double* hm, int ldh, const double* sm, int lds,
                         double* w, double* c, int ldc, double rmsq, int* used) {
  if (!handle) return set_error("xtbeig_dev_scc_sygvd: null handle");
  if (n < 0 || ldh < (n > 1 ? n : 1) || lds < (n > 1 ? n : 1) || ldc < (n > 1 ? n : 1))
    return set_error("xtbeig_dev_scc_sygvd: bad dimensions");
  Handle& h = handle->h;
  bool fp32 = h.scc_fp32_rmsq > 0 && !h.scc_fp64_latched && rmsq >= h.scc_fp32_rmsq &&
              h.scc_fp32_cycles < h.scc_fp32_maxiter;
  if (h.scc_fp32_rmsq > 0 && !fp32) h.scc_fp64_latched = true;  // FP64 for the rest of the sequence
  int st[xtbeig::NSTATUS] = {0, 0, 0, 0};
  int rc = guarded([&] {
    if (fp32) {
      const int keep = h.refine_iters;
      h.refine_iters = h.scc_fp32_refine ? (keep > 0 ? keep : 2) : 0;
      try {
        xtbeig::sygvd_device(h, n, hm, ldh, sm, lds, w, c, ldc, true, h.band_override);
      } catch (...) {
        h.refine_iters = keep;
        throw;
      }
      h.refine_iters = keep;
      read_status(h, st);
      if (st[1] || st[2] || st[3]) {
        // the FP32 solve failed (chase stall, D&C, or the refined pairs failed the check): redo the cycle in FP64,
        // FP64 from here on
        h.scc_fp64_restart = 1;
        h.scc_fp64_latched = true;
        fp32 = false;
      } else if (st[0]) {
        throw std::runtime_error("xtbeig_dev_scc_sygvd: S is not positive definite");
      }
    }
    if (!fp32) {
      xtbeig::sygvd_device(h, n, hm, ldh, sm, lds, w, c, ldc, false, h.band_override);
      read_status(h, st);
      if (st[0]) throw std::runtime_error("xtbeig_dev_scc_sygvd: S is not positive definite");
      if (st[1] || st[2]) throw std::runtime_error("xtbeig_dev_scc_sygvd: FP64 eigensolve failed");
    }
  });
  if (rc) return rc;
  h.scc_cycles += 1;
  if (fp32) h.scc_fp32_cycles += 1;
  if (used) *used = fp32 ? XTBEIG_FP32_REFINE : XTBEIG_FP64;
  return 0;
}

int xtbeig_get_scc_precision(xtbeig_handle_t handle, double* fp32_rmsq, int* fp32_maxiter, int* fp32_refine,
                             int* fp32_cycles, int* fp64_restart, int* scc_cycles) {
  if (!handle) return set_error("xtbeig_get_scc_precision: null handle");
  Handle& h = handle->h;
  if (fp32_rmsq) *fp32_rmsq = h.scc_fp32_rmsq;
  if (fp32_maxiter) *fp32_maxiter = h.scc_fp32_maxiter;
  if (fp32_refine) *fp32_refine = h.scc_fp32_refine;
  if (fp32_cycles) *fp32_cycles = h.scc_fp32_cycles;
  if (fp64_restart) *fp64_restart = h.scc_fp64_restart;
  if (scc_cycles) *scc_cycles = h.scc_cycles;
  if (h.scc_cycles == 0) return set_error("xtbeig_get_scc_precision: no SCC cycle has run since xtbeig_scc_begin");
  return 0;
}

int xtbeig_dev_sygvd_graph(xtbeig_handle_t handle, int precision, int n, const double* hm, int ldh,
                           const double* sm, int lds, double* w, double* c, int ldc) {
  if (!handle) return set_error("xtbeig_dev_sygvd_graph: null handle");
  if (n < 0 || ldh < (n > 1 ? n : 1) || lds < (n > 1 ? n : 1) || ldc < (n > 1 ? n : 1))
    return set_error("xtbeig_dev_sygvd_graph: bad dimensions");
  if (precision != XTBEIG_FP64 && precision != XTBEIG_FP32_REFINE)
    return set_error("xtbeig_dev_sygvd_graph: unknown precision");
  return guarded([&] {
    xtbeig::sygvd_graph(handle->h, n, hm, ldh, sm, lds, w, c, ldc, precision == XTBEIG_FP32_REFINE);
  });
}

int xtbeig_dev_sygvd_batched(xtbeig_handle_t handle, int precision, int count, const int* n, const double* const* hm,
                             const int* ldh, const double* const* sm, const int* lds, double* const* w,
                             double* const* c, const int* ldc, int slots, int use_graph, float* latency_ms) {
  if (!handle) return set_error("xtbeig_dev_sygvd_batched: null handle");
  if (precision != XTBEIG_FP64 && precision != XTBEIG_FP32_REFINE)
    return set_error("xtbeig_dev_sygvd_batched: unknown precision");
  if (count < 0 || slots < 1 || slots > 64) return set_error("xtbeig_dev_sygvd_batched: count >= 0, 1 <= slots <= 64");
  for (int q = 0; q < count; ++q)
    if (n[q] < 1 || ldh[q] < n[q] || lds[q] < n[q] || ldc[q] < n[q])
      return set_error("xtbeig_dev_sygvd_batched: bad dimensions");
  return guarded([&] {
    xtbeig::sygvd_batched(handle->h, count, n, hm, ldh, sm, lds, w, c, ldc, precision == XTBEIG_FP32_REFINE, slots,
                          use_graph != 0, latency_ms);
  });
}

int xtbeig_dev_sygvd_grouped(xtbeig_handle_t handle, int precision, int count, const int* n, const double* const* hm,
                             const int* ldh, const double* const* sm, const int* lds, double* const* w,
                             double* const* c, const int* ldc, int granule, int slots, int* status,
                             float* latency_ms) {
  if (!handle) return set_error("xtbeig_dev_sygvd_grouped: null handle");
  if (precision != XTBEIG_FP64 && precision != XTBEIG_FP32_REFINE)
    return set_error("xtbeig_dev_sygvd_grouped: unknown precision");
  if (count < 0 || granule < 0 || granule > 1024 || slots < 0 || slots > 64)
    return set_error("xtbeig_dev_sygvd_grouped: count >= 0, 0 <= granule <= 1024, 0 <= slots <= 64");
  for (int q = 0; q < count; ++q)
    if (n[q] < 1 || ldh[q] < n[q] || lds[q] < n[q] || ldc[q] < n[q])
      return set_error("xtbeig_dev_sygvd_grouped: bad dimensions");
  return guarded([&] {
    xtbeig::sygvd_grouped(handle->h, count, n, hm, ldh, sm, lds, w, c, ldc, precision == XTBEIG_FP32_REFINE, granule,
                          slots, status, latency_ms);
  });
}

int xtbeig_chase_in_lds(int n, int b, int fp32) {
  if (n < 1 || b < 1 || b > xtbeig::MAX_BAND) return 0;
  return xtbeig::chase_in_lds(n, b, fp32 != 0) ? 1 : 0;
}

int xtbeig_set_gemm_tile(int width) {
  if (width != 0 && width != 64 && width != 128) return set_error("xtbeig_set_gemm_tile: width must be 0, 64 or 128");
  xtbeig::set_gemm_tile(width);
  return 0;
}

int xtbeig_gemm_tile_width(int m, int n, int batch, int precision) {
  if (m < 1 || n < 1 || batch < 1) {
    set_error("xtbeig_gemm_tile_width: m, n and batch must be >= 1");
    return -1;
  }
  return xtbeig::gemm_tile_width(m, n, batch, precision != XTBEIG_FP32_REFINE);
}

int xtbeig_gemm_tile_buckets(void) { return xtbeig::GEMM_TILE_BUCKETS; }

int xtbeig_get_gemm_tile_table(int precision, int* bucket_tiles, int* width) {
  if (!width) return set_error("xtbeig_get_gemm_tile_table: null width");
  xtbeig::get_gemm_tile_table(precision != XTBEIG_FP32_REFINE, width);
  if (bucket_tiles)
    for (int q = 0; q < xtbeig::GEMM_TILE_BUCKETS; ++q) bucket_tiles[q] = (int)xtbeig::gemm_tile_bucket_start(q);
  return 0;
}

int xtbeig_set_gemm_tile_table(int precision, const int* width) {
  if (!width) return set_error("xtbeig_set_gemm_tile_table: null width");
  for (int q = 0; q < xtbeig::GEMM_TILE_BUCKETS; ++q)
    if (width[q] != 0 && width[q] != 64 && width[q] != 128)
      return set_error("xtbeig_set_gemm_tile_table: widths must be 0, 64 or 128");
  xtbeig::set_gemm_tile_table(precision != XTBEIG_FP32_REFINE, width);
  return 0;
}

extern "C++" {
namespace {
// The tile widths of one precision timed on the handle's device and stream (scratch of 3 nmax^2 elements allocated
// for the call), stored for gemm()'s choice. Returns the number of buckets tuned.
template <typename T>
int autotune_gemm_tiles_on(Handle& h, int nmin, int nmax, int* bucket_tiles, int* width) {
  hipStream_t s = h.stream;
  XTB_HIP(hipSetDevice(h.device));
  const size_t nn = (size_t)nmax * nmax;
  T* buf = nullptr;
  XTB_HIP(hipMalloc(&buf, 3 * nn * sizeof(T)));
  int count = 0;
  try {
    xtbeig::set_zero<T>(s, nmax, 3 * nmax, buf, nmax);  // operand values do not change the timing
    count = xtbeig::autotune_gemm_tiles<T>(s, nmin, nmax, buf, buf + nn, buf + 2 * nn, bucket_tiles, width);
  } catch (...) {
    (void)hipStreamSynchronize(s);
    (void)hipFree(buf);
    throw;
  }
  XTB_HIP(hipStreamSynchronize(s));
  XTB_HIP(hipFree(buf));
  return count;
}
}  // namespace
}  // extern "C++"

int xtbeig_autotune_gemm(xtbeig_handle_t handle, int precision, int nmin, int nmax, int* bucket_tiles, int* width,
                         int* count) {
  if (!handle) return set_error("xtbeig_autotune_gemm: null handle");
  if (precision != XTBEIG_FP64 && precision != XTBEIG_FP32_REFINE)
    return set_error("xtbeig_autotune_gemm: unknown precision");
  if (nmin < 1 || nmax < nmin || nmax > xtbeig::MAX_N)
    return set_error(("xtbeig_autotune_gemm: need 1 <= nmin <= nmax <= " + std::to_string(xtbeig::MAX_N)).c_str());
  return guarded([&] {
    const int c = precision == XTBEIG_FP64 ? autotune_gemm_tiles_on<double>(handle->h, nmin, nmax, bucket_tiles, width)
                                           : autotune_gemm_tiles_on<float>(handle->h, nmin, nmax, bucket_tiles, width);
    if (count) *count = c;
  });
}

int xtbeig_autotune(xtbeig_handle_t handle, int precision, int nmin, int nmax, int* bucket_start, int* band,
                    float* ms, int* count) {
  if (!handle) return set_error("xtbeig_autotune: null handle");
  if (precision != XTBEIG_FP64 && precision != XTBEIG_FP32_REFINE)
    return set_error("xtbeig_autotune: unknown precision");
  if (nmin < 2 || nmax < nmin || nmax > xtbeig::MAX_N)
    return set_error(("xtbeig_autotune: need 2 <= nmin <= nmax <= " + std::to_string(xtbeig::MAX_N)).c_str());
  int cnt = 0;
  for (int q = 0; q < Handle::NBUCKET; ++q)
    if (!(xtbeig::bucket_edge(q + 1) - 1 < nmin || xtbeig::bucket_edge(q) > nmax)) ++cnt;
  if (count) *count = cnt;
  return guarded([&] {
    // the GEMM tile costs first (both precisions for the mixed path, whose congruence and refinement are FP64), so
    // that the band widths are timed with the products as they will run
    autotune_gemm_tiles_on<double>(handle->h, nmin, nmax, nullptr, nullptr);
    if (precision == XTBEIG_FP32_REFINE) autotune_gemm_tiles_on<float>(handle->h, nmin, nmax, nullptr, nullptr);
    xtbeig::autotune_bands(handle->h, precision == XTBEIG_FP32_REFINE, nmin, nmax, bucket_start, band, ms);
  });
}

int xtbeig_get_band_table(xtbeig_handle_t handle, int precision, int* bucket_start, int* band) {
  if (!handle) return set_error("xtbeig_get_band_table: null handle");
  if (precision != XTBEIG_FP64 && precision != XTBEIG_FP32_REFINE)
    return set_error("xtbeig_get_band_table: unknown precision");
  for (int q = 0; q < Handle::NBUCKET; ++q) {
    if (bucket_start) bucket_start[q] = xtbeig::bucket_edge(q);
    if (band) band[q] = handle->h.tuned_band[precision == XTBEIG_FP32_REFINE ? 1 : 0][q];
  }
  return 0;
}

int xtbeig_set_band_table(xtbeig_handle_t handle, int precision, const int* band) {
  if (!handle || !band) return set_error("xtbeig_set_band_table: null argument");
  if (precision != XTBEIG_FP64 && precision != XTBEIG_FP32_REFINE)
    return set_error("xtbeig_set_band_table: unknown precision");
  for (int q = 0; q < Handle::NBUCKET; ++q)
    if (band[q] < 0 || band[q] > xtbeig::MAX_BAND)
      return set_error(("xtbeig_set_band_table: band widths must be in [0, " + std::to_string(xtbeig::MAX_BAND) +
                        "]").c_str());
  for (int q = 0; q < Handle::NBUCKET; ++q) handle->h.tuned_band[precision == XTBEIG_FP32_REFINE ? 1 : 0][q] = band[q];
  return 0;
}

int xtbeig_band_buckets(void) { return Handle::NBUCKET; }

}  // extern "C"
