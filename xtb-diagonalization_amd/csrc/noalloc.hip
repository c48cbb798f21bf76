// The allocation-free twin of xtbeig_dev_potrf_inverse, for callers that cannot allocate device memory (code under
// HIP graph capture, tight loops, harnesses that forbid hipMalloc). The Cholesky factorization with the inverse of
// its factor needs a scratch buffer (the products of the recursive inverse); here it lives in a file-scope device
// array of fixed size instead of the handle's arena. Each device holds its own copy of the array until the process
// ends. Its device address is looked up once per device (hipGetSymbolAddress, with that device current) when a
// handle is created on the device, and only that address is passed to kernels: the array's name in host code is a
// host-side stand-in, not a device address.
//
// Every call on one device shares that device's array: calls that overlap on one device (two streams, two threads)
// race. The array keeps the last call's contents; nothing in it needs to start at zero (every product writes its
// part of the scratch before reading it).
#include "../../include/xtbeig.h"
#include "solver.hpp"

#include <mutex>

namespace xtbeig {

namespace {

constexpr size_t SCRATCH_DOUBLES = (size_t)(NOALLOC_MAX_N / 2 + 64) * (NOALLOC_MAX_N / 2 + 64) +
                                   (size_t)NOALLOC_MAX_N * 64 + 64;

static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mutex;
void* scratch_ptr[MAX_DEVICES] = {};

}  // namespace

size_t noalloc_scratch_bytes() { return SCRATCH_DOUBLES * sizeof(double); }

// The device address of `scratch` on `device` (which must be current), looked up on the first call for the device.
void* noalloc_scratch(int device) {
  if (device < 0 || device >= MAX_DEVICES) throw std::runtime_error("noalloc_scratch: device index out of range");
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (!scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (rc != hipSuccess || p == nullptr)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(rc));
    scratch_ptr[device] = p;
  }
  return scratch_ptr[device];
}

void potrf_inverse_noalloc(Handle& h, int n, double* s, int lds, double* u, int ldu, double* x, int ldx) {
  if (n > NOALLOC_MAX_N)
    throw std::runtime_error("xtbeig_dev_potrf_inverse_noalloc: n = " + std::to_string(n) +
                             " exceeds the allocation-free limit n <= " + std::to_string(NOALLOC_MAX_N));
  if (n <= 0) return;
  if (potrf_workspace(n) > SCRATCH_DOUBLES) throw std::runtime_error("potrf_inverse_noalloc: scratch too small");
  void* tmp = h.scratch_dev;
  if (!tmp) throw std::runtime_error("potrf_inverse_noalloc: the handle has no scratch address");
  potrf_inverse(h.stream, n, s, lds, u, ldu, x, ldx, static_cast<double*>(tmp), h.dinfo);
}

}  // namespace xtbeig
