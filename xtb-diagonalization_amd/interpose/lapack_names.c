/* Opt-in interposing build: the LAPACK Fortran names dsyevd_, ssyevd_, dsygvd_, ssygvd_, dsytrd_ and ssytrd_ as thin
 * forwards to the library's xtbeig_* entry points, so that a caller that links this library ahead of the host LAPACK
 * (or preloads it) has its LAPACK eigensolves run on the GPU without a source change. The callers are xTB's
 * eigensolve (mctc_dsygvd: a full dsygvd per SCC cycle in release builds; mctc_dsygvd_factorized: dsygst + dsyevd +
 * dtrsm, the reference's BRIEF.md:101) and the reference's Fortran drivers (scripts/kernels/eigbench.f90:107 and
 * refine.f90:156 call dsyevd('V', 'U', ...)).
 *
 * gfortran passes the length of each CHARACTER argument as a hidden size_t after the last ordinary argument; the
 * forwards accept and ignore them (one character is read from each string, as LAPACK does).
 *
 * XTBEIG_DSYGVD_MIXED=1 in the environment (read at the first dsygvd_ call) routes dsygvd_ to the FP32 + FP64
 * refinement path (xtbeig_dsygvd_mixed), which returns FP64-accurate pairs or redoes the solve in FP64. */
#include <stddef.h>
#include <stdlib.h>

#include "../../include/xtbeig.h"
#include "../../include/xtbeig_lapack.h"

static int mixed_dsygvd(void) {
  static int mode = -1;
  if (mode < 0) {
    const char *v = getenv("XTBEIG_DSYGVD_MIXED");
    mode = (v && v[0] == '1') ? 1 : 0;
  }
  return mode;
}

void dsytrd_(const char *uplo, const int *n, double *a, const int *lda, double *d, double *e, double *tau,
             double *work, const int *lwork, int *info, size_t uplo_len) {
  (void)uplo_len;
  xtbeig_dsytrd(uplo, n, a, lda, d, e, tau, work, lwork, info);
}

void ssytrd_(const char *uplo, const int *n, float *a, const int *lda, float *d, float *e, float *tau, float *work,
             const int *lwork, int *info, size_t uplo_len) {
  (void)uplo_len;
  xtbeig_ssytrd(uplo, n, a, lda, d, e, tau, work, lwork, info);
}

void dsyevd_(const char *jobz, const char *uplo, const int *n, double *a, const int *lda, double *w, double *work,
             const int *lwork, int *iwork, const int *liwork, int *info, size_t jobz_len, size_t uplo_len) {
  (void)jobz_len;
  (void)uplo_len;
  xtbeig_dsyevd(jobz, uplo, n, a, lda, w, work, lwork, iwork, liwork, info);
}

void ssyevd_(const char *jobz, const char *uplo, const int *n, float *a, const int *lda, float *w, float *work,
             const int *lwork, int *iwork, const int *liwork, int *info, size_t jobz_len, size_t uplo_len) {
  (void)jobz_len;
  (void)uplo_len;
  xtbeig_ssyevd(jobz, uplo, n, a, lda, w, work, lwork, iwork, liwork, info);
}

void dsygvd_(const int *itype, const char *jobz, const char *uplo, const int *n, double *a, const int *lda, double *b,
             const int *ldb, double *w, double *work, const int *lwork, int *iwork, const int *liwork, int *info,
             size_t jobz_len, size_t uplo_len) {
  (void)jobz_len;
  (void)uplo_len;
  if (mixed_dsygvd())
    xtbeig_dsygvd_mixed(itype, jobz, uplo, n, a, lda, b, ldb, w, work, lwork, iwork, liwork, info);
  else
    xtbeig_dsygvd(itype, jobz, uplo, n, a, lda, b, ldb, w, work, lwork, iwork, liwork, info);
}

void ssygvd_(const int *itype, const char *jobz, const char *uplo, const int *n, float *a, const int *lda, float *b,
             const int *ldb, float *w, float *work, const int *lwork, int *iwork, const int *liwork, int *info,
             size_t jobz_len, size_t uplo_len) {
  (void)jobz_len;
  (void)uplo_len;
  xtbeig_ssygvd(itype, jobz, uplo, n, a, lda, b, ldb, w, work, lwork, iwork, liwork, info);
}
