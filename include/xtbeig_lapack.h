/* The LAPACK Fortran names exported by the opt-in interposing library libxtbeig_lapack.so (built beside
 * libxtbeig.so, linked against it). Each is a forward to the xtbeig_* entry point of the same arguments in xtbeig.h;
 * the trailing size_t arguments are the CHARACTER lengths gfortran passes and are ignored.
 *
 * Link libxtbeig_lapack.so ahead of the host LAPACK (or preload it) to interpose. The names collide with the host
 * LAPACK's by design, which is why they live in a separate library and are not exported by libxtbeig.so.
 *
 * Reference interfaces replaced: xTB's eigensolve calls dsygvd (mctc_dsygvd) or dsygst + dsyevd + dtrsm
 * (mctc_dsygvd_factorized), the reference's BRIEF.md:101; the reference's drivers call dsyevd('V', 'U', ...) at
 * scripts/kernels/eigbench.f90:107 and refine.f90:156, and dsytrd('U', ...) at scripts/kernels/tridiag.f90:74.
 * XTBEIG_DSYGVD_MIXED=1 routes dsygvd_ to xtbeig_dsygvd_mixed. */
#ifndef XTBEIG_LAPACK_H
#define XTBEIG_LAPACK_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* LAPACK dsytrd / ssytrd (uplo = 'U' only): xtbeig_dsytrd / xtbeig_ssytrd. */
void dsytrd_(const char *uplo, const int *n, double *a, const int *lda, double *d, double *e, double *tau,
             double *work, const int *lwork, int *info, size_t uplo_len);
void ssytrd_(const char *uplo, const int *n, float *a, const int *lda, float *d, float *e, float *tau, float *work,
             const int *lwork, int *info, size_t uplo_len);
/* LAPACK dsyevd / ssyevd: xtbeig_dsyevd / xtbeig_ssyevd. */
void dsyevd_(const char *jobz, const char *uplo, const int *n, double *a, const int *lda, double *w, double *work,
             const int *lwork, int *iwork, const int *liwork, int *info, size_t jobz_len, size_t uplo_len);
void ssyevd_(const char *jobz, const char *uplo, const int *n, float *a, const int *lda, float *w, float *work,
             const int *lwork, int *iwork, const int *liwork, int *info, size_t jobz_len, size_t uplo_len);
/* LAPACK dsygvd / ssygvd (itype = 1): xtbeig_dsygvd (or xtbeig_dsygvd_mixed under XTBEIG_DSYGVD_MIXED=1) /
 * xtbeig_ssygvd. */
void dsygvd_(const int *itype, const char *jobz, const char *uplo, const int *n, double *a, const int *lda, double *b,
             const int *ldb, double *w, double *work, const int *lwork, int *iwork, const int *liwork, int *info,
             size_t jobz_len, size_t uplo_len);
void ssygvd_(const int *itype, const char *jobz, const char *uplo, const int *n, float *a, const int *lda, float *b,
             const int *ldb, float *w, float *work, const int *lwork, int *iwork, const int *liwork, int *info,
             size_t jobz_len, size_t uplo_len);

#ifdef __cplusplus
}
#endif

#endif
