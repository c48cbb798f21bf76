/*
 * xtbeig: dense symmetric(-definite) eigensolver for one AMD MI355X (gfx950), a drop-in for the eigensolve in
 * xTB's SCC cycle. Plain C ABI: pointers and sizes only.
 *
 * Two groups of entry points:
 *  1. Host-memory, Fortran-by-reference entry points (every argument a pointer, 32-bit integers, column-major, no
 *     hidden string lengths), with the names and argument meaning of the library the reference builds into libxtb
 *     in place of xtb/src/mctc/lapack/sytrd.c (reference: Sources/Diagonalization/Exports.swift:1-138), plus
 *     LAPACK-convention sytrd / syevd / sygvd under the xtbeig_ prefix (workspace queries with lwork = -1). Each
 *     call copies its inputs to the GPU, runs there and copies the results back; a per-thread handle holds the
 *     device workspace between calls.
 *  2. Device-memory entry points on an explicit handle and HIP stream (xtbeig_dev_*), for callers that keep their
 *     matrices resident on the GPU. These enqueue work and return; xtbeig_synchronize waits and reports status.
 *
 * Matrices are column-major with leading dimensions; "upper" means the upper triangle is referenced.
 */
#ifndef XTBEIG_H
#define XTBEIG_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- 1a. The sytrd.c symbol set (reference: Sources/Diagonalization/Exports.swift) -------------------------- */

/* dstedc('I') semantics: eigenvalues of the symmetric tridiagonal (d, e) ascending in d, eigenvectors in z (n x n,
 * ldz). e is destroyed. info = 0 on success, > 0 when a secular equation fails to converge. work / iwork are not
 * used (the GPU holds its own workspace) but lwork / liwork must be at least 1 + 4n + n^2 / 3 + 5n as LAPACK's, or
 * -1 for a query answered in work[0] / iwork[0]. Replaces Exports.swift:88-101 (xtb_dstedc). */
void xtb_dstedc(const int *n, double *d, double *e, double *z, const int *ldz, double *work, const int *lwork,
                int *iwork, const int *liwork, int *info);
/* FP32 interface of the same (Exports.swift:103-115, xtb_sstedc); the tridiagonal problem is solved in FP64 on the
 * GPU and rounded to FP32. */
void xtb_sstedc(const int *n, float *d, float *e, float *z, const int *ldz, float *work, const int *lwork,
                int *iwork, const int *liwork, int *info);
/* dst(1:n, 1:n) = the symmetric matrix whose upper triangle is src's (src's lower triangle is not read)
 * (Exports.swift:117-138, xtb_dsymmetrize_pad). */
void xtb_dsymmetrize_pad(const int *n, const double *src, const int *lds, double *dst, const int *ldd);

/* One-stage reduction to tridiagonal form, LAPACK dsytrd('U') convention (nb = 32, nx = 32): on exit d (n), e (n-1),
 * tau (n-1), and the Householder vectors above the superdiagonal of a; the strictly lower triangle of a is not
 * touched. work (n * 32 in the reference) is not used. Replaces Exports.swift:6-16 (xtb_dsytrd_upper, the
 * reference's tridiagonalize, Tridiagonalization.swift:595-618). */
void xtb_dsytrd_upper(const int *n, double *a, const int *lda, double *d, double *e, double *tau, double *work);
/* FP32 twin (Exports.swift:18-27, xtb_ssytrd_upper). */
void xtb_ssytrd_upper(const int *n, float *a, const int *lda, float *d, float *e, float *tau, float *work);
/* The same reduction on the full symmetric matrix: both triangles of a are read and both are updated and returned
 * (Exports.swift:29-38, xtb_dsytrd_full; tridiagonalize with symmetric = false). */
void xtb_dsytrd_full(const int *n, double *a, const int *lda, double *d, double *e, double *tau, double *work);
/* z = Q z for the first m columns (dormtr 'L', 'U', 'N'), Q from xtb_dsytrd_upper's a and tau; work (32 (2n + 2m +
 * 32) in the reference) is not used. Replaces Exports.swift:56-65 (xtb_dormtr_upper, backTransformTridiagonal,
 * TridiagonalBackTransformation.swift:16-45). */
void xtb_dormtr_upper(const int *n, const int *m, const double *a, const int *lda, const double *tau, double *z,
                      const int *ldz, double *work);
/* FP32 twin (Exports.swift:67-75, xtb_sormtr_upper). */
void xtb_sormtr_upper(const int *n, const int *m, const float *a, const int *lda, const float *tau, float *z,
                      const int *ldz, float *work);

/* The two steps either side of the reduction in xTB's per-cycle eigensolve (eigensolve.F90's congruence and
 * back-transformation, restated in the reference's scripts/kernels/congruence.c:1-8 and eigbench.f90:101-168): the
 * upper triangle of Z = X^T H X with X = U^-1 upper triangular (S = U^T U), Z's strictly lower triangle untouched
 * (trmm 'R','U','N','N' + gemm + trmm 'L','U','T','N' in blocks of 96 in the reference; here two GEMMs on the GPU). */
void xtb_dcongruence_upper(const int *n, const double *h, const int *ldh, const double *x, const int *ldx, double *z,
                           const int *ldz);
/* C(:, 1:m) = X C(:, 1:m), X upper triangular (n x n): the back-solve trmm('L','U','N','N', n, m) of the occupied
 * columns (congruence.c:6, eigbench.f90:166; xTB passes m = occupied orbitals, the reference's LOG.md:310-312). */
void xtb_dtrmm_back(const int *n, const int *m, const double *x, const int *ldx, double *c, const int *ldc);

/* ---- 1b. LAPACK-convention entry points ------------------------------------------------------------------------ */

/* dsytrd(uplo, ...): uplo = 'U' only (info = -1 otherwise), as xtb_dsytrd_upper; lwork >= 1, lwork = -1 queries
 * the optimal n * 32 into work[0]. d/s sytrd with LAPACK dsytrd's interface. info = n + 1 when the GPU work failed
 * (HIP error, allocation or launch failure; the reason in xtbeig_last_error()); a, d, e and tau are then not valid.
 * The void xtb_* entry points of section 1a report such failures only through xtbeig_last_error(). */
void xtbeig_dsytrd(const char *uplo, const int *n, double *a, const int *lda, double *d, double *e, double *tau,
                   double *work, const int *lwork, int *info);
void xtbeig_ssytrd(const char *uplo, const int *n, float *a, const int *lda, float *d, float *e, float *tau,
                   float *work, const int *lwork, int *info);

/* dsyevd(jobz, uplo, ...): all eigenvalues (w, ascending) and, for jobz = 'V', eigenvectors in a. lwork >=
 * 1 + 6n + 2n^2, liwork >= 3 + 5n ('V'); lwork = -1 or liwork = -1 is a workspace query. The reference's
 * eigensolve (tridiagonalize/sbrReduce + stedc + back-transformation, Sources/EigenBench/main.swift:291-420).
 * info > 0: a secular equation of the divide and conquer did not converge (its index), or n + 1 when the bulge
 * chase stalled or the GPU work failed or n exceeds 2048 + band width (the reason in xtbeig_last_error()). */
void xtbeig_dsyevd(const char *jobz, const char *uplo, const int *n, double *a, const int *lda, double *w,
                   double *work, const int *lwork, int *iwork, const int *liwork, int *info);
void xtbeig_ssyevd(const char *jobz, const char *uplo, const int *n, float *a, const int *lda, float *w, float *work,
                   const int *lwork, int *iwork, const int *liwork, int *info);
/* dsygvd(itype = 1, jobz, uplo, ...): A x = lambda B x with B symmetric positive definite. On exit a holds the
 * B-orthonormal eigenvectors ('V') and b the Cholesky factor U (B = U^T U, upper) as LAPACK leaves it. info > n:
 * B's leading minor of order info - n is not positive definite. The xTB fact_solve generalized solve (reference
 * spec: scripts/kernels/eigbench.f90:101-168). Sizes up to n = 2048 + band width (the built-in band is 8);
 * a larger n returns info = 2 n + 1 with the reason in xtbeig_last_error(), before any work is done. */
void xtbeig_dsygvd(const int *itype, const char *jobz, const char *uplo, const int *n, double *a, const int *lda,
                   double *b, const int *ldb, double *w, double *work, const int *lwork, int *iwork,
                   const int *liwork, int *info);
/* As xtbeig_dsygvd, but computed in FP32 on the matrix cores and refined in FP64 (Ogita-Aishima with Rayleigh-Ritz
 * on clusters, the reference's fp32_refine scheme, scripts/kernels/fp32_error.c:95-158, refine.f90:254-325); same
 * arguments and results as the FP64 call. When the refined pairs fail the a-posteriori check (status word 3 of
 * xtbeig_synchronize), or the FP32 chase or divide and conquer fails, the call solves the problem again in FP64 and
 * returns that result. */
void xtbeig_dsygvd_mixed(const int *itype, const char *jobz, const char *uplo, const int *n, double *a,
                         const int *lda, double *b, const int *ldb, double *w, double *work, const int *lwork,
                         int *iwork, const int *liwork, int *info);
/* ssygvd: FP32 interface of the generalized solve (LAPACK ssygvd's arguments); computed by the FP32 + FP64-refinement
 * path and rounded to FP32. */
void xtbeig_ssygvd(const int *itype, const char *jobz, const char *uplo, const int *n, float *a, const int *lda,
                   float *b, const int *ldb, float *w, float *work, const int *lwork, int *iwork, const int *liwork,
                   int *info);

/* ---- 2. Device-memory API ------------------------------------------------------------------------------------- */

typedef struct xtbeig_handle *xtbeig_handle_t;

/* Precision of a generalized solve. */
enum { XTBEIG_FP64 = 0, XTBEIG_FP32_REFINE = 1 };

/* Create a handle on a HIP device with its own stream (an ordinary blocking stream: ordered with the legacy default
 * stream, so that a caller's default-stream work on the inputs precedes the solve). Returns 0 on success. */
int xtbeig_create(xtbeig_handle_t *handle, int device);
int xtbeig_destroy(xtbeig_handle_t handle);
/* Enqueue on the caller's HIP stream (hipStream_t passed as void*) instead of the handle's own. */
int xtbeig_set_stream(xtbeig_handle_t handle, void *stream);
void *xtbeig_get_stream(xtbeig_handle_t handle);
/* Band width of the two-stage reduction (0: the built-in table per n; at most 64) and the refinement iterations
 * (0 to 8; 0 returns the FP32 eigenpairs unrefined and unchecked). */
int xtbeig_set_tuning(xtbeig_handle_t handle, int band, int refine_iterations);
/* Test setting: cap the divide and conquer's secular-equation iterations at secular_maxit (0 gives up after the
 * starting guess; -1 restores the solver's own limit), so that the failure path (info > 0 from stedc / syevd /
 * sygvd, status word 2) can be exercised without faulting anything. handle NULL: the calling thread's handle of the
 * host-array entry points. No reference counterpart (the reference propagates a failed dlaed4 as info,
 * Stages/ParallelDivideAndConquer.swift:70-71). */
int xtbeig_set_test_limits(xtbeig_handle_t handle, int secular_maxit);
/* The built-in band width for an n x n solve (fp32 nonzero: the FP32 reduction of the mixed path). No reference
 * counterpart beyond SBROptions.recommended (SBROptions.swift:73-82), which picks b per n the same way. */
int xtbeig_default_band(int n, int fp32);
/* 1 when the bulge chase of an n x n band of width b keeps the band in one CU's LDS (else it works on the band in
 * device memory, several times slower), 0 otherwise or for invalid sizes. No reference counterpart. */
int xtbeig_chase_in_lds(int n, int b, int fp32);
/* Wait for the handle's stream; status[0..3] = Cholesky info, bulge-chase stall flag, divide-and-conquer info, and
 * the refinement check of an FP32 + refinement solve (nonzero: the refined eigenpairs failed it, or a cluster's
 * Rayleigh-Ritz step met a Gram block that is not positive definite) (status: 4 ints, may be NULL). Returns 0 when all
 * four are zero. Clears them. */
int xtbeig_synchronize(xtbeig_handle_t handle, int *status);
/* The last refinement check on the handle (after xtbeig_synchronize or with the stream idle): ratios[0] = the
 * residual estimate ||(H C - S C Lambda) V||_F / sqrt(8) over 8 random sign vectors V, divided by 1e-13 n ||H||_F /
 * sqrt(n); ratios[1] = the estimate of ||C^T S C - I||_F divided by 1e-12 n. A solve passes with both <= 1. No
 * reference counterpart (the reference's gates need the FP64 oracle: eigbench.f90:189-204). */
int xtbeig_refine_check(xtbeig_handle_t handle, double *ratios);
/* Message of the last failed call on this thread. */
const char *xtbeig_last_error(void);

/* Generalized solve of device-resident H, S (n x n, full symmetric storage, not modified): eigenvalues w (n,
 * ascending) and S-orthonormal eigenvectors C (n x n, ldc). precision: XTBEIG_FP64 or XTBEIG_FP32_REFINE. */
int xtbeig_dev_sygvd(xtbeig_handle_t handle, int precision, int n, const double *h, int ldh, const double *s,
                     int lds, double *w, double *c, int ldc);
/* Standard solve of a device-resident symmetric A (full storage, not modified). */
int xtbeig_dev_dsyevd(xtbeig_handle_t handle, int n, const double *a, int lda, double *w, double *z, int ldz);
/* FP32 a and z; the eigenvalues w are FP64 (n doubles: the divide and conquer runs in FP64). */
int xtbeig_dev_ssyevd(xtbeig_handle_t handle, int n, const float *a, int lda, double *w, float *z, int ldz);
/* All n eigenvalues and the first m eigenvectors (1 <= m <= n) of the symmetric a: the last product of the
 * back-transformation costs 2 n^2 m instead of 2 n^3 (xTB needs the occupied columns only, LOG.md:310-312). */
int xtbeig_dev_dsyevd_partial(xtbeig_handle_t handle, int n, int m, const double *a, int lda, double *w, double *z,
                              int ldz);

/* The geometry of an SCC sequence, factored once and kept on the handle (xTB's X = U^-1 is fixed across the cycles of
 * one geometry: congruence.c:7): S (n x n, full storage) is copied, S = U^T U and X = U^-1 are formed on the GPU and
 * kept until the next call (Cholesky info in status word 0). */
int xtbeig_dev_set_geometry(xtbeig_handle_t handle, int n, const double *s, int lds);
/* Z's upper triangle = X^T H X with the handle's X (Z's strictly lower triangle untouched); congruence.c:1-5. */
int xtbeig_dev_congruence(xtbeig_handle_t handle, int n, const double *h, int ldh, double *z, int ldz);
/* C(:, 0:m) = X C(:, 0:m) with the handle's X (0 <= m <= n); congruence.c:6. */
int xtbeig_dev_back_solve(xtbeig_handle_t handle, int n, int m, double *c, int ldc);
/* One cycle's generalized solve on the handle's geometry: as xtbeig_dev_sygvd with the cached S, U, X (no
 * factorization per cycle), all n eigenvalues and the first m eigenvectors C (n x m, ldc >= n), 1 <= m <= n. The FP32
 * + refinement path refines all n pairs (the refinement needs the whole basis) and returns m. */
int xtbeig_dev_sygvd_geometry(xtbeig_handle_t handle, int precision, int n, int m, const double *h, int ldh,
                              double *w, double *c, int ldc);
/* Tridiagonal (d, e) -> eigenvalues ascending in d, eigenvectors in z; e destroyed (FP64). */
int xtbeig_dev_dstedc(xtbeig_handle_t handle, int n, double *d, double *e, double *z, int ldz);

/* ---- 2b. Stage entry points (device memory), for callers composing their own pipelines and for testing ------ */

/* C = alpha op(A) op(B) + beta C on the matrix cores (op: 'N' or 'T'). */
int xtbeig_dev_dgemm(xtbeig_handle_t handle, char transa, char transb, int m, int n, int k, double alpha,
                     const double *a, int lda, const double *b, int ldb, double beta, double *c, int ldc);
int xtbeig_dev_sgemm(xtbeig_handle_t handle, char transa, char transb, int m, int n, int k, float alpha,
                     const float *a, int lda, const float *b, int ldb, float beta, float *c, int ldc);
/* S = U^T U and X = U^{-1} (upper; zero below the diagonal). s (full storage) is destroyed. */
int xtbeig_dev_potrf_inverse(xtbeig_handle_t handle, int n, double *s, int lds, double *u, int ldu, double *x,
                             int ldx);
/* The same as xtbeig_dev_potrf_inverse with no device allocation: no hipMalloc or hipFree, no allocating library
 * call; its scratch buffer (10 MiB) is a static device array that each device holds until the process ends. No
 * counterpart in the reference: for callers that cannot allocate (code under HIP graph capture, tight loops,
 * harnesses that forbid hipMalloc). n <= 2048 (a larger n returns nonzero with the reason in xtbeig_last_error()).
 * All calls on one device share its scratch array: calls that overlap on one device race. Cholesky info in status
 * word 0 (xtbeig_synchronize), as for xtbeig_dev_potrf_inverse. */
int xtbeig_dev_potrf_inverse_noalloc(xtbeig_handle_t handle, int n, double *s, int lds, double *u, int ldu,
                                     double *x, int ldx);
/* Dense -> band (reference: Stages/BandForm.swift:28-142): a (full symmetric, destroyed) -> lower band of width b
 * in band storage ab (element (i, j), 0 <= i - j < 2b, at i + (2b - 1) j; xtbeig_band_elems(n, b) elements, which
 * include b padding columns past n that the chase clears and uses), the panels' Householder vectors v1 (n x n) and
 * T factors t1 (b x n). */
long long xtbeig_band_elems(int n, int b);
int xtbeig_dev_dreduce_to_band(xtbeig_handle_t handle, int n, int b, double *a, int lda, double *v1, double *t1,
                               double *ab);
int xtbeig_dev_sreduce_to_band(xtbeig_handle_t handle, int n, int b, float *a, int lda, float *v1, float *t1,
                               float *ab);
/* Band -> tridiagonal by bulge chasing (reference: Stages/BulgeChasing.swift:123-321): ab (destroyed) -> d, e;
 * reflectors w (H = I - w w^T) in v2 (n x n, sweep s in column s). */
int xtbeig_dev_dchase(xtbeig_handle_t handle, int n, int b, double *ab, double *v2, double *d, double *e);
int xtbeig_dev_schase(xtbeig_handle_t handle, int n, int b, float *ab, float *v2, float *d, float *e);
/* z = Q2 z (bulge reflectors) and z = Q1 z (band panels) for m columns (reference:
 * Stages/BulgeBackTransformation.swift:11-30, Stages/BandFormBackTransformation.swift:10-21). */
int xtbeig_dev_dback_bulges(xtbeig_handle_t handle, int n, int b, const double *v2, double *z, int ldz, int m);
int xtbeig_dev_sback_bulges(xtbeig_handle_t handle, int n, int b, const float *v2, float *z, int ldz, int m);
int xtbeig_dev_dback_band(xtbeig_handle_t handle, int n, int b, const double *v1, const double *t1, double *z,
                          int ldz, int m);
int xtbeig_dev_sback_band(xtbeig_handle_t handle, int n, int b, const float *v1, const float *t1, float *z, int ldz,
                          int m);

/* Device twins of xtb_{d,s}sytrd_upper (full = 0) and xtb_dsytrd_full (full = 1): a (n x n, lda) is reduced in
 * place and left in full storage (both triangles updated); d, e, tau device arrays. */
int xtbeig_dev_dsytrd_upper(xtbeig_handle_t handle, int n, double *a, int lda, double *d, double *e, double *tau,
                            int full);
int xtbeig_dev_ssytrd_upper(xtbeig_handle_t handle, int n, float *a, int lda, float *d, float *e, float *tau,
                            int full);
/* Device twins of xtb_{d,s}ormtr_upper. */
int xtbeig_dev_dormtr_upper(xtbeig_handle_t handle, int n, int m, const double *a, int lda, const double *tau,
                            double *z, int ldz);
int xtbeig_dev_sormtr_upper(xtbeig_handle_t handle, int n, int m, const float *a, int lda, const float *tau,
                            float *z, int ldz);

/* The generalized solve replayed from a HIP graph captured for (n, precision, band) on this handle (captured on the
 * first call after one uncaptured solve; recaptured when a larger solve has since regrown the handle's workspace).
 * Same arguments and results as xtbeig_dev_sygvd; H and S are copied into the graph's buffers and w, C out of them.
 * For the repeated same-n solves of an SCC sequence. No reference counterpart (the reference solves on the CPU). */
int xtbeig_dev_sygvd_graph(xtbeig_handle_t handle, int precision, int n, const double *h, int ldh, const double *s,
                           int lds, double *w, double *c, int ldc);
/* count independent generalized solves of uneven n (arrays of count device pointers and sizes, in host memory) on up
 * to `slots` concurrent streams of sub-handles, longest first onto the least-loaded slot; every workspace and plan
 * of an eager batch is set up before the first solve is enqueued. use_graph: each solve through the slot's graph
 * cache. Waits for completion; latency_ms (count, may be NULL) receives each solve's time on its stream. Status words
 * of the slots are merged into the handle's (xtbeig_synchronize). For batches of independent problems (256 of n in
 * [124, 1204], say); no reference counterpart. */
int xtbeig_dev_sygvd_batched(xtbeig_handle_t handle, int precision, int count, const int *n, const double *const *h,
                             const int *ldh, const double *const *s, const int *lds, double *const *w,
                             double *const *c, const int *ldc, int slots, int use_graph, float *latency_ms);
/* The same count independent solves grouped inside the kernels: each problem padded to n rounded up to a multiple of
 * `granule` (0: 64; 1: no padding) with an exactly decoupled block (H: a diagonal inside the problem's spectral
 * range, S: the identity) that the results leave out, problems of one padded size solved together, every kernel of
 * the solve launched once over the whole group (the problem in the grid's z extent, workspaces in equal slices of
 * an arena). Groups (or chunks of a large group) run concurrently on up to `slots` sub-handle streams (0: 4).
 * Results as xtbeig_dev_sygvd's for each problem (bit-identical to it with granule 1). Waits for completion. status
 * (count x 4 ints, may be NULL): each problem's status words as xtbeig_synchronize reports them (word 3 = 2: the
 * padded block did not separate; the problem is not solved), also merged into the handle's; latency_ms (count, may
 * be NULL): each problem's time from the call's start to the end of its chunk. The sub-handles' workspaces grow to
 * their largest chunks' slices (chunks are cut to fit 60 % of free device memory between the slots). For batches
 * of independent problems (256 of n in [124, 1204], say); no reference counterpart (the reference solves one
 * problem per thread). */
int xtbeig_dev_sygvd_grouped(xtbeig_handle_t handle, int precision, int count, const int *n, const double *const *h,
                             const int *ldh, const double *const *s, const int *lds, double *const *w,
                             double *const *c, const int *ldc, int granule, int slots, int *status,
                             float *latency_ms);
/* Autotuner: first the GEMM tile widths over [nmin, nmax] (xtbeig_autotune_gemm, FP64, and FP32 too for precision
 * XTBEIG_FP32_REFINE), then for every n bucket of the handle's table that meets [nmin, nmax] (2 <= nmin <= nmax <=
 * 2048), the candidate band widths (also the panel width of the dense-to-band reduction) 8, 16 and 32 timed on a
 * synthetic pair of the bucket's middle size, the fastest kept in the table, which later solves on the handle use
 * (an explicit xtbeig_set_tuning band overrides it).
 * Outputs per tuned bucket (arrays of xtbeig_band_buckets() entries, may be NULL): its first n, the band, the time
 * in ms; *count the number of buckets tuned. Reference counterpart: SBROptions.recommended's fixed table
 * (SBROptions.swift:73-82). */
int xtbeig_autotune(xtbeig_handle_t handle, int precision, int nmin, int nmax, int *bucket_start, int *band,
                    float *ms, int *count);
/* The tuned table (xtbeig_band_buckets() entries; band 0: not tuned, the built-in choice applies), to save and
 * restore it. */
int xtbeig_get_band_table(xtbeig_handle_t handle, int precision, int *bucket_start, int *band);
int xtbeig_set_band_table(xtbeig_handle_t handle, int precision, const int *band);
int xtbeig_band_buckets(void);

/* GEMM tile shape. Every product runs in tiles of 64 rows by 64 or 128 columns. The width of a product (m x n,
 * batched b times) is the tuned width of its bucket of 64 x 64 tile counts (b ceil(m / 64) ceil(n / 64), half-octave
 * buckets: the count sets how the grid fills the CUs), or, for an untuned bucket, a round model (ceil(tiles / CUs)
 * rounds of one tile each). Results do not depend on the width (each element sums the same products in the same
 * order). Process-wide, per precision. No reference counterpart (the reference calls the host BLAS).
 * xtbeig_set_gemm_tile: force one width for every product (64 or 128); 0 restores the choice above.
 * xtbeig_gemm_tile_width: the width a product of m x n, batch times, gets (-1 on invalid sizes).
 * xtbeig_gemm_tile_buckets: the number of buckets; xtbeig_get/set_gemm_tile_table: the first tile count of each
 * bucket and its width (0: not tuned), to save and restore them.
 * xtbeig_autotune_gemm: for each bucket that the square products of n in [nmin, nmax] reach (1 <= nmin <= nmax <=
 * 2048), times an n x n x n product of about the bucket's middle tile count in both widths on the handle's device
 * and stores the faster (scratch of 3 nmax^2 elements allocated for the call); the buckets tuned (first tile count,
 * width; arrays of xtbeig_gemm_tile_buckets() entries, may be NULL) and their count (may be NULL) are returned. */
int xtbeig_set_gemm_tile(int width);
int xtbeig_gemm_tile_width(int m, int n, int batch, int precision);
int xtbeig_gemm_tile_buckets(void);
int xtbeig_get_gemm_tile_table(int precision, int *bucket_tiles, int *width);
int xtbeig_set_gemm_tile_table(int precision, const int *width);
int xtbeig_autotune_gemm(xtbeig_handle_t handle, int precision, int nmin, int nmax, int *bucket_tiles, int *width,
                         int *count);

/* SCC precision schedule, per handle (mirrors xTB's xtb_setSCCPrecision / xtb_getSCCPrecision contract, reference
 * LOG.md:356-370 and scripts/test_sccprecision.py:79-161): fp32_rmsq = 0 runs every cycle in FP64 (the default);
 * fp32_rmsq > 0 runs a cycle in FP32 while the caller's rms charge change is >= fp32_rmsq, for at most fp32_maxiter
 * FP32 cycles, and in FP64 from the first cycle that is not, to the end of the sequence. fp32_refine = 1 refines
 * the FP32 eigenpairs in FP64 (Ogita-Aishima), 0 uses them as they are. Invalid values (negative or non-finite
 * fp32_rmsq, fp32_maxiter < 1, fp32_refine not 0 or 1, fp32_refine without FP32) return nonzero with the message in
 * xtbeig_last_error and leave the settings unchanged. Handles are independent: two threads with their own handles
 * never affect each other. */
int xtbeig_set_scc_precision(xtbeig_handle_t handle, double fp32_rmsq, int fp32_maxiter, int fp32_refine);
/* Start an SCC sequence: counters zero, FP32 allowed again. */
int xtbeig_scc_begin(xtbeig_handle_t handle);
/* One cycle's generalized eigensolve under the schedule (device memory, as xtbeig_dev_sygvd), rmsq the cycle's rms
 * charge change; waits for the result. An FP32 solve that fails (chase stall, divide-and-conquer info) is redone in
 * FP64 and the sequence stays FP64 (fp64_restart = 1). *used (may be NULL): the precision the cycle ran in. */
int xtbeig_dev_scc_sygvd(xtbeig_handle_t handle, int n, const double *h, int ldh, const double *s, int lds, double *w,
                         double *c, int ldc, double rmsq, int *used);
/* The settings in effect and the sequence's counters (any pointer may be NULL). Nonzero when no cycle has run since
 * xtbeig_scc_begin (the counters are still written). */
int xtbeig_get_scc_precision(xtbeig_handle_t handle, double *fp32_rmsq, int *fp32_maxiter, int *fp32_refine,
                             int *fp32_cycles, int *fp64_restart, int *scc_cycles);

/* Fermi-smeared density matrix of batches of equal-n generalized eigensystems (device memory; problem p at base +
 * p * stride). f_i = 1 / (1 + exp((w_i - mu) / kT)); bisect != 0: mu[p] is bisected so that sum_i f_i = nel, else
 * mu[p] is used as given. Outputs P = C diag(f) C^T (n x n, ldp) and occ (3 n per problem: f, df/deps,
 * log cosh((w - mu) / 2kT)), which the backward takes. The reference forms P this way from the eigenpairs
 * (scripts/kernels/refine.f90:85-132); no reference backward exists. */
int xtbeig_dev_density(xtbeig_handle_t handle, int batch, int n, const double *w, long long sw, const double *c,
                       int ldc, long long sc, double kT, double nel, double *mu, int bisect, double *p, int ldp,
                       long long sp, double *occ);
/* Backward of the density matrix: given gbar = dL/dP (symmetric), hbar = dL/dH = C (F o G) C^T and sbar = dL/dS =
 * -C (L o G) C^T with G = C^T gbar C, F and L the Loewner matrices of f and eps f (their derivatives on the
 * diagonal, formed stably for any gap), plus the chemical-potential term unless mu_fixed. */
int xtbeig_dev_density_backward(xtbeig_handle_t handle, int batch, int n, const double *w, long long sw,
                                const double *c, int ldc, long long sc, double kT, const double *mu, int mu_fixed,
                                const double *occ, const double *gbar, int ldg, long long sg, double *hbar, int ldhb,
                                long long shb, double *sbar, int ldsb, long long ssb);
/* Open-shell occupations, as the reference forms them (scripts/kernels/refine.f90:85-93): alpha and beta electrons
 * Fermi-smeared separately, f = f_alpha + f_beta, with chemical potentials mu[2 p] (sum f_alpha = nel_alpha) and
 * mu[2 p + 1] (sum f_beta = nel_beta), each bisected when bisect != 0; P = C diag(f) C^T. occ: 6 n per problem (the
 * alpha channel's f | f' | log cosh, then the beta channel's), for the backward below. */
int xtbeig_dev_density_spin(xtbeig_handle_t handle, int batch, int n, const double *w, long long sw, const double *c,
                            int ldc, long long sc, double kT, double nel_alpha, double nel_beta, double *mu, int bisect,
                            double *p, int ldp, long long sp, double *occ);
/* Backward of the open-shell density matrix: the sums over the two channels of the closed-shell terms (one chemical
 * potential term per channel unless mu_fixed); mu and occ as xtbeig_dev_density_spin left them. */
int xtbeig_dev_density_backward_spin(xtbeig_handle_t handle, int batch, int n, const double *w, long long sw,
                                     const double *c, int ldc, long long sc, double kT, const double *mu,
                                     int mu_fixed, const double *occ, const double *gbar, int ldg, long long sg,
                                     double *hbar, int ldhb, long long shb, double *sbar, int ldsb, long long ssb);
/* The same backward without stored eigenvectors: each problem's eigenpairs are recomputed from (h, s) with the
 * given precision (XTBEIG_FP64 or XTBEIG_FP32_REFINE) and mu re-bisected to nel, one problem at a time in O(n^2)
 * device memory (for batches whose eigenvectors do not fit). */
int xtbeig_dev_density_backward_recompute(xtbeig_handle_t handle, int batch, int n, const double *h, int ldh,
                                          long long sh, const double *s, int lds, long long ss, int precision,
                                          double kT, double nel, int mu_fixed, const double *gbar, int ldg,
                                          long long sg, double *hbar, int ldhb, long long shb, double *sbar, int ldsb,
                                          long long ssb);

/* One divide-and-conquer merge (dlaed1 semantics; reference: Stages/ParallelDivideAndConquer.swift:77-124) of two
 * solved halves: d (n) ascending within rows 0 ..< n1 and n1 ..< n, q (n x n, ldq) block diagonal with their
 * eigenvectors, rho the coupling. On return d and q hold the merged eigensystem (ascending). dlaed2's outputs are
 * exposed for checking: k_ctot = K and the four column-type counts, d_deflated = d after deflation, q2 (n^2),
 * dlamda (K), w (K), indxc (n, 1-based), rho_out = |2 rho|. */
int xtbeig_dev_dc_merge(xtbeig_handle_t handle, int n, int n1, double rho, double *d, double *q, int ldq,
                        double *d_deflated, double *q2, int *k_ctot, double *dlamda, double *w, int *indxc,
                        double *rho_out);

/* ---- 3. Dumps and hashes (host code; no device) ------------------------------------------------------------- */

/* HS.bin, the reference's dump of one generalized eigenproblem (Sources/EigenBench/main.swift:151-175,
 * scripts/kernels/refine.f90:64-83): int32 n, H, S (n x n float64, column-major), optionally e (n) and C (n x n).
 * info: n and whether the eigenpairs follow (the file size must match one of the two layouts). Returns nonzero on
 * error, with the message in xtbeig_hsbin_error(). */
int xtbeig_hsbin_info(const char *path, int *n, int *has_solution);
/* Read H and S (and e, C when not NULL; the file must hold them) into caller arrays with leading dimensions; h or s
 * may be NULL to skip them. */
int xtbeig_hsbin_read(const char *path, int n, double *h, int ldh, double *s, int lds, double *e, double *c, int ldc);
/* Write H, S and, when e and c are both given, the eigenpairs. */
int xtbeig_hsbin_write(const char *path, int n, const double *h, int ldh, const double *s, int lds, const double *e,
                       const double *c, int ldc);
const char *xtbeig_hsbin_error(void);
/* FNV-1a (64-bit) of `bytes` bytes, continuing from `hash` (start from 0xcbf29ce484222325): the reference's
 * fingerprint of a variant's tridiagonal (d, e) bits, main.swift:253-266. */
unsigned long long xtbeig_fnv1a(const void *data, long long bytes, unsigned long long hash);

#ifdef __cplusplus
}
#endif

#endif /* XTBEIG_H */
