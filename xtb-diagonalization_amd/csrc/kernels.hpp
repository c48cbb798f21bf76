// Internal launchers (host side) of the HIP kernels. All take the stream they run on and device pointers.
#pragma once
#include "common.hpp"

#include <vector>

namespace xtbeig {

// gemm.hip ---------------------------------------------------------------------------------------------------------
template <typename T>
void gemm(hipStream_t s, char ta, char tb, int m, int n, int k, T alpha, const T* A, int lda, const T* B, int ldb,
          T beta, T* C, int ldc, int batch = 1, int64_t sA = 0, int64_t sB = 0, int64_t sC = 0);

// The tile width (64 or 128 columns, 64 rows) gemm() uses for an m x n product batched `batch` times, and its
// controls: a fixed width for every product (0: chosen per product), and the tuned width per bucket of the product's
// count of 64 x 64 tiles (half-octave buckets; 0: not tuned, a round model decides), per precision, process-wide.
constexpr int GEMM_TILE_BUCKETS = 24;
int gemm_tile_bucket(int64_t tiles64);
int64_t gemm_tile_bucket_start(int q);
int gemm_tile_width(int m, int n, int batch, bool fp64);
void set_gemm_tile(int width);
int get_gemm_tile();
void set_gemm_tile_table(bool fp64, const int* width);
void get_gemm_tile_table(bool fp64, int* width);
template <typename T>
int autotune_gemm_tiles(hipStream_t s, int nmin, int nmax, T* a, T* b, T* c, int* bucket_tiles, int* widths);

// C = alpha A^T B + beta C (C n x n) over C's 64 x 64 tiles on and above the diagonal only; the tiles below it are
// not written (a syrk-shaped update whose consumer reads the upper triangle).
template <typename T>
void gemm_upper_tn(hipStream_t s, int n, int k, T alpha, const T* A, int lda, const T* B, int ldb, T beta, T* C,
                   int ldc, int batch = 1, int64_t sA = 0, int64_t sB = 0, int64_t sC = 0);

template <typename T>
int gemm_splitk(hipStream_t s, int m, int n, int k, const T* A, int lda, const T* B, int ldb, T* C, int kc);

// C (m x n, n <= 64) = op(A) B, op(A) m x k ('N' or 'T'), B k x n: K split in chunks of THIN_KC so that a thin
// product still fills the GPU, partials summed in a fixed order. part: gemm_thin_parts(m, n, k) elements of scratch
// (none when k <= THIN_KC: then one plain product).
constexpr int THIN_KC = 64;
size_t gemm_thin_parts(int m, int n, int k);
template <typename T>
void gemm_thin(hipStream_t s, char ta, int m, int n, int k, const T* A, int lda, const T* B, int ldb, T* C, int ldc,
               T* part);

// grouped.hip ------------------------------------------------------------------------------------------------------
// A problem of size n_p < n solved padded to n: H = [H_p 0; 0 D], S = [S_p 0; 0 I] with D diagonal. The padded block
// stays exactly decoupled through every stage (its reflectors are the identity, its merges deflate it), so its
// eigenpairs are (D_kk, e_k) and the problem's own eigenvectors keep exact zeros in the padded rows.
struct PadSpec {
  const int* n;  // device, one per problem of the batch (not moved by Zb): each problem's own size
  double* h;     // H as the solve holds it (moved by Zb), or null
  int ldh;
};
// hp (n x n, symmetric, the problem's block in its leading n_p rows and columns and zeros elsewhere): the padded
// diagonal D_kk (k >= n_p) set to distinct values inside the block's spectral range, spread evenly between its smallest
// and largest diagonal entries (diagonal entries are Rayleigh quotients, so the scale of the matrix is unchanged);
// the same values into h's padded diagonal when h is given.
void pad_spectrum(hipStream_t s, int n, const int* nvec, double* hp, int ldhp, double* h, int ldh);

// util.hip ---------------------------------------------------------------------------------------------------------
// dst(m x n) = src (with conversion); both column-major.
template <typename S, typename D>
void copy_matrix(hipStream_t s, int m, int n, const S* src, int lds, D* dst, int ldd);
// dst (n x m) = src (m x n)^T (with conversion).
template <typename S, typename D>
void transpose_matrix(hipStream_t s, int m, int n, const S* src, int lds, D* dst, int ldd);
// dst's upper triangle (i <= j) = src's; dst's strictly lower triangle untouched.
template <typename T>
void copy_upper(hipStream_t s, int n, const T* src, int lds, T* dst, int ldd);
// Mirror the referenced triangle of a (uplo 'U': upper is read) into the other one, in place.
template <typename T>
void symmetrize(hipStream_t s, char uplo, int n, T* a, int lda);
template <typename T>
void set_zero(hipStream_t s, int m, int n, T* a, int lda);
template <typename T>
void set_identity(hipStream_t s, int n, T* a, int lda);
// The FP32 copy of an FP64 matrix for the mixed path, kept inside FP32's range: *amax (one 64-bit word, cleared here)
// receives max |src|; while that lies in [2^-60, 2^60] (squares and sums of squares stay normal FP32 numbers) dst =
// src rounded, else dst = 2^-e src with 2^e the power of two just above max |src| (|e| <= 1000; a power of two
// commutes with every rounding of the FP32 reduction, chase and products, so the solve is that of src up to the
// factor). A non-finite src gives dst = 0 and *status = 1. unscale_values multiplies w (n) by the same 2^e.
void copy_to_fp32_ranged(hipStream_t s, int n, const double* src, int lds, float* dst, int ldd,
                         unsigned long long* amax, int* status);
void unscale_values(hipStream_t s, int n, double* w, const unsigned long long* amax);
// dst = the symmetric matrix of src's upper triangle (xtb_dsymmetrize_pad).
void symmetrize_pad(hipStream_t s, int n, const double* src, int lds, double* dst, int ldd);

// chol.hip ---------------------------------------------------------------------------------------------------------
// S = U^T U (upper Cholesky) of the symmetric positive definite a (n x n, full storage, destroyed); writes U (zero
// below the diagonal) and X = U^{-1} (zero below). info (device int): 0, or the LAPACK potrf index j > 0 of the
// first non-positive pivot. tmp: potrf_workspace(n) doubles. With a second stream s2 and two events, X's block
// columns are formed on s2 while s factors (left-looking: X(0:k, k) = -X(0:k, 0:k) U(0:k, k) X_kk once block k's
// diagonal is inverted), and s waits for s2 before it returns; without, X is assembled at the end by halving.
size_t potrf_workspace(int n);
void potrf_inverse(hipStream_t s, int n, double* a, int lda, double* u, int ldu, double* x, int ldx, double* tmp,
                   int* info, hipStream_t s2 = nullptr, hipEvent_t ev = nullptr, hipEvent_t ev_join = nullptr);

// band.hip ---------------------------------------------------------------------------------------------------------
struct BandWork;
template <typename T>
size_t band_workspace(int n, int b);  // elements of T
// Dense symmetric a (n x n, full storage) -> lower band b stored in ab (band storage, element (i, j), i >= j at
// ab[(i - j) + ... ] see band.hip), Householder vectors v1 (n x n, explicit unit lower trapezoids per panel) and
// T factors t1 (b x n).
template <typename T>
void reduce_to_band(hipStream_t s, int n, int b, T* a, int lda, T* v1, T* t1, T* ab, T* work);

// chase.hip --------------------------------------------------------------------------------------------------------
// Band storage of the chase: element (i, j), 0 <= i - j <= 2b - 1, at i + band_ld(b) j, for columns j < n + b (the
// b columns past the matrix, and the rows past n in every column's slot, are zero padding the chase may read).
__host__ __device__ inline int band_ld(int b) { return b > 1 ? 2 * b - 1 : 1; }
__host__ __device__ inline int64_t band_elems(int n, int b) { return (int64_t)band_ld(b) * (n + b) + (n + b) + 1; }
// Band (ab, leading dimension 2b, (2b + 1) n elements) -> tridiagonal d, e; reflectors w into v2 (n x n, column s
// holds sweep s, rows rs ..< re).
template <typename T>
void chase_bulges(hipStream_t s, int n, int b, T* ab, T* v2, T* d, T* e, int* flags);
// Whether the chase of an n x n band of width b (8 or 16) keeps the band in LDS.
bool chase_in_lds(int n, int b, bool fp32);

// dc.hip -----------------------------------------------------------------------------------------------------------
// One merge of the divide-and-conquer tree: rows/columns off ..< off + sz, torn after n1; q2off: its Q2 / S offset.
struct DcNode {
  int off, sz, n1, q2off;
};
// The tree for one n (tear at sz / 2 down to 1 x 1 leaves), uploaded once.
struct DcPlan {
  struct Level { int start = 0, count = 0, pre_start = 0, maxsz = 0; };
  int n = 0, nmerge = 0;
  std::vector<Level> lvls;  // by depth, root first
  DcNode* nodes_dev = nullptr;
  int* pre_index_dev = nullptr;
  int* split_pre_dev = nullptr;  // n + 1: pre-order index of the merge that tears at row m, or -1
  explicit DcPlan(int n);
  ~DcPlan();
  DcPlan(const DcPlan&) = delete;
  DcPlan& operator=(const DcPlan&) = delete;
};
size_t dc_workspace_doubles(int n);
size_t dc_workspace_ints(int n);
// Eigenvalues (ascending, into d) and eigenvectors (z, n x n, ldz) of the tridiagonal (d, e), FP64; e is destroyed.
// info: device int, set to a positive value when a secular equation fails to converge. secular_maxit >= 0 caps the
// secular solver's iterations below its own limit (a test setting that makes it give up; -1: the solver's limit).
void stedc(hipStream_t s, const DcPlan& plan, double* d, double* e, double* z, int ldz, double* work, int* iwork,
           int* info, int secular_maxit = -1);
void dc_merge_probe(hipStream_t s, int n, int n1, double rho, double* d, double* q, int ldq, double* work,
                    int* iwork, int* info, double* d_defl_out, double* q2_out, int* kctot_out, double* dlamda_out,
                    double* w_out, int* indxc_out, double* rho_out);

}  // namespace xtbeig

namespace xtbeig {
// backtransform.hip ------------------------------------------------------------------------------------------------
// Both back-transformations of the two-stage reduction act on zt = Z^T (m x n, column-major, ld ldzt: row i of Z is
// the contiguous column i of zt), so that a wavefront's lanes hold 64 consecutive eigenvectors.
// Bulge reflectors: the chase's reflectors of g consecutive sweeps s0 ..< s0 + g at one bulge index k (rows
// row0 ..< row0 + height) form one block; blocks are grouped into stages of row-disjoint blocks, each stage after
// every block it overlaps.
struct BulgeBlock {
  int row0, height, count, s0;
};
struct BackPlan {
  int n = 0, b = 0, g = 0;
  int nbulge = 0, nstages = 0, max_stage = 0;
  BulgeBlock* blocks_dev = nullptr;
  int* stage_off_dev = nullptr;  // nstages + 1
  int* order_dev = nullptr;      // block indices by stage
  BulgeBlock* sorted_dev = nullptr;  // the blocks in stage order (blocks_dev[order_dev[q]] at q)
  int rows = 0;                      // rows the blocks' windows reach (at most n + 63)
  // band panels j = 0, b, ..., last, merged ppg at a time into groups of up to ppg * b reflectors
  int last = -1, ppg = 1, ngroups = 0;
  BackPlan(int n, int b);
  ~BackPlan();
  BackPlan(const BackPlan&) = delete;
  BackPlan& operator=(const BackPlan&) = delete;
};
template <typename T>
size_t back_band_workspace(int n, int b, int m);  // elements of T
// zt's shape for m eigenvectors of size n: leading dimension zt_ld(m) (m rounded up to 64) and zt_rows(n) columns,
// the last zt_rows(n) - n of them zero (the bulge kernel's windows run into them).
int zt_ld(int m);
int zt_rows(int n);
// zt <- zt Q2^T, i.e. Z <- Q2 Z (v2: the chase's reflectors w, H = I - w w^T, column s = sweep s); transposed:
// zt <- zt Q2, i.e. Z <- Q2^T Z.
template <typename T>
void back_transform_bulges(hipStream_t s, const BackPlan& p, const T* v2, T* zt, int ldzt, int m,
                           bool transposed = false);
// zt <- zt Q1^T, i.e. Z <- Q1 Z (v1, t1: the panels of reduce_to_band; v1 zero outside the panels' V).
template <typename T>
void back_transform_band(hipStream_t s, const BackPlan& p, const T* v1, const T* t1, T* zt, int ldzt, int m, T* work);

// sytrd.hip --------------------------------------------------------------------------------------------------------
// LAPACK dsytrd('U') (nb = 32, nx = 32) of the symmetric a (n x n; full_input: both triangles hold the matrix, else
// the upper one, mirrored first). a is kept in full storage and returned with both triangles updated (the reference's
// full variant); its upper triangle is dsytrd('U')'s: Householder vectors above the superdiagonal, d, e, tau.
template <typename T>
size_t sytrd_workspace(int n);  // elements of T
template <typename T>
void sytrd_upper(hipStream_t s, int n, T* a, int lda, T* d, T* e, T* tau, T* work, bool full_input);
// z = Q z (dormtr 'L', 'U', 'N') for m columns; a, tau from sytrd_upper.
template <typename T>
size_t ormtr_workspace(int n, int m);
template <typename T>
void ormtr_upper(hipStream_t s, int n, int m, const T* a, int lda, const T* tau, T* z, int ldz, T* work);

// density.hip ------------------------------------------------------------------------------------------------------
// Batches of equal-n problems, problem p at base + p * stride. occ: 3 n per problem (f | f' | log cosh(x / 2)).
size_t density_workspace(int batch, int n);  // doubles
// nspin spin channels (1, or 2: alpha and beta smeared separately, f = f_alpha + f_beta); mu and occ hold nspin
// entries / 3 n blocks per problem. mu (bisect: sum f_channel = nel_a, nel_b; else the given mu), occ,
// P = C diag(f) C^T.
void density_forward(hipStream_t s, int batch, int n, int nspin, const double* w, int64_t sw, const double* c,
                     int ldc, int64_t sc, double kT, double nel_a, double nel_b, int bisect, double* mu, double* p,
                     int ldp, int64_t sp, double* occ, double* work);
// Hbar, Sbar from Gbar = dL/dP (mu_fixed: no re-bisection term).
void density_backward(hipStream_t s, int batch, int n, int nspin, const double* w, int64_t sw, const double* c,
                      int ldc, int64_t sc, double kT, const double* mu, int mu_fixed, const double* occ,
                      const double* gbar, int ldg, int64_t sg, double* hbar, int ldhb, int64_t shb, double* sbar,
                      int ldsb, int64_t ssb, double* work);

// refine.hip -------------------------------------------------------------------------------------------------------
size_t refine_workspace(int n);  // doubles
size_t refine_iwork(int n);      // ints
// iterations Ogita-Aishima steps (with Rayleigh-Ritz on clusters) on the FP32 solve's (w, C), the eigenpairs sorted,
// then the a-posteriori check: status (device int) |= 1 when the result fails it; ratios (device, 2 doubles, may be
// null) receive the check's residual and orthogonality ratios (<= 1 passes). iterations = 0: nothing is done.
void refine_oa(hipStream_t s, int n, const double* h, int ldh, const double* sm, int lds, double* c, int ldc,
               double* w, int iterations, double* work, int* iwork, int* status, double* ratios);
}  // namespace xtbeig
