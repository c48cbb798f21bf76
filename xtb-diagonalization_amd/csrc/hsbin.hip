// The reference's dump format of one generalized eigenproblem, HS.bin (Sources/EigenBench/main.swift:151-175,
// scripts/kernels/refine.f90:64-83): int32 n, then H and S (n x n float64, column-major), optionally followed by the
// eigenvalues e (n) and the eigenvectors C (n x n, column-major); little-endian, no padding. And the FNV-1a hash the
// reference's eigen-bench prints of each variant's tridiagonal (d, e) bits (main.swift:253-266), so that variants
// that must give the same bits can be compared. Host code only: no device is touched.
#include "../../include/xtbeig.h"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

namespace {

thread_local std::string hs_error;

int fail(const std::string& what) {
  hs_error = what;
  return 1;
}

// File size in bytes, or -1.
long long file_size(FILE* f) {
  if (std::fseek(f, 0, SEEK_END) != 0) return -1;
  const long long sz = std::ftell(f);
  if (std::fseek(f, 0, SEEK_SET) != 0) return -1;
  return sz;
}

}  // namespace

extern "C" {

const char* xtbeig_hsbin_error(void) { return hs_error.c_str(); }

int xtbeig_hsbin_info(const char* path, int* n, int* has_solution) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(std::string("cannot open ") + path);
  const long long sz = file_size(f);
  int32_t n32 = 0;
  const size_t got = std::fread(&n32, 4, 1, f);
  std::fclose(f);
  if (got != 1 || sz < 4) return fail(std::string(path) + ": shorter than its header");
  if (n32 < 1) return fail(std::string(path) + ": n = " + std::to_string(n32));
  const long long nn = (long long)n32 * n32;
  const long long pair = 4 + 16 * nn, full = pair + 8 * (long long)n32 + 8 * nn;
  if (sz != pair && sz != full)
    return fail(std::string(path) + ": " + std::to_string(sz) + " bytes, expected " + std::to_string(pair) + " or " +
                std::to_string(full) + " for n = " + std::to_string(n32));
  if (n) *n = n32;
  if (has_solution) *has_solution = sz == full;
  return 0;
}

int xtbeig_hsbin_read(const char* path, int n, double* h, int ldh, double* s, int lds, double* e, double* c,
                      int ldc) {
  int nf = 0, sol = 0;
  if (xtbeig_hsbin_info(path, &nf, &sol)) return 1;
  if (nf != n) return fail(std::string(path) + ": holds n = " + std::to_string(nf) + ", not " + std::to_string(n));
  if (ldh < n || lds < n || (c && ldc < n)) return fail("xtbeig_hsbin_read: leading dimension below n");
  if ((e || c) && !sol) return fail(std::string(path) + ": holds no eigenpairs");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(std::string("cannot open ") + path);
  int rc = 0;
  auto read_matrix = [&](double* dst, int ld) {
    for (int j = 0; j < n && rc == 0; ++j) {
      if (dst) {
        if (std::fread(dst + (size_t)j * ld, 8, n, f) != (size_t)n) rc = fail(std::string(path) + ": short read");
      } else if (std::fseek(f, 8L * n, SEEK_CUR) != 0) {
        rc = fail(std::string(path) + ": seek failed");
      }
    }
  };
  std::fseek(f, 4, SEEK_SET);
  read_matrix(h, ldh);
  read_matrix(s, lds);
  if (rc == 0 && (e || c)) {
    if (e) {
      if (std::fread(e, 8, n, f) != (size_t)n) rc = fail(std::string(path) + ": short read");
    } else if (std::fseek(f, 8L * n, SEEK_CUR) != 0) {
      rc = fail(std::string(path) + ": seek failed");
    }
    if (rc == 0 && c) read_matrix(c, ldc);
  }
  std::fclose(f);
  return rc;
}

int xtbeig_hsbin_write(const char* path, int n, const double* h, int ldh, const double* s, int lds, const double* e,
                       const double* c, int ldc) {
  if (n < 1 || ldh < n || lds < n || (c && ldc < n)) return fail("xtbeig_hsbin_write: bad dimensions");
  if ((e == nullptr) != (c == nullptr)) return fail("xtbeig_hsbin_write: e and c go together");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(std::string("cannot create ") + path);
  int rc = 0;
  const int32_t n32 = n;
  if (std::fwrite(&n32, 4, 1, f) != 1) rc = fail("short write");
  auto write_matrix = [&](const double* src, int ld) {
    for (int j = 0; j < n && rc == 0; ++j)
      if (std::fwrite(src + (size_t)j * ld, 8, n, f) != (size_t)n) rc = fail("short write");
  };
  write_matrix(h, ldh);
  write_matrix(s, lds);
  if (rc == 0 && e) {
    if (std::fwrite(e, 8, n, f) != (size_t)n) rc = fail("short write");
    write_matrix(c, ldc);
  }
  if (std::fclose(f) != 0 && rc == 0) rc = fail("close failed");
  return rc;
}

unsigned long long xtbeig_fnv1a(const void* data, long long bytes, unsigned long long hash) {
  const unsigned char* p = static_cast<const unsigned char*>(data);
  for (long long k = 0; k < bytes; ++k) hash = (hash ^ p[k]) * 0x100000001b3ULL;
  return hash;
}

}  // extern "C"
