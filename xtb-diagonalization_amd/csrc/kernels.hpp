// Internal launchers (host side) of the HIP kernels. All take the stream they run on and device pointers.
#pragma once
#include "common.hpp"

#include <vector>

namespace xtbeig {

// gemm.hip ---------------------------------------------------------------------------------------------------------
template <typename T>
void gemm(hipStream_t s, char ta, char tb, int m, int n, int k, T alpha, const T* A, int lda, const T* B, int ldb,
          T beta, T* C, int ldc, int batch = 1, int64_t sA = 0, int64_t sB = 0, int64_t sC = 0);

// util.hip ---------------------------------------------------------------------------------------------------------
// dst(m x n) = src (with conversion); both column-major.
template <typename S, typename D>
void copy_matrix(hipStream_t s, int m, int n, const S* src, int lds, D* dst, int ldd);
// Mirror the referenced triangle of a (uplo 'U': upper is read) into the other one, in place.
template <typename T>
void symmetrize(hipStream_t s, char uplo, int n, T* a, int lda);
template <typename T>
void set_zero(hipStream_t s, int m, int n, T* a, int lda);
template <typename T>
void set_identity(hipStream_t s, int n, T* a, int lda);
// dst = the symmetric matrix of src's upper triangle (xtb_dsymmetrize_pad).
void symmetrize_pad(hipStream_t s, int n, const double* src, int lds, double* dst, int ldd);

// chol.hip ---------------------------------------------------------------------------------------------------------
// S = U^T U (upper Cholesky) of the symmetric positive definite a (n x n, full storage, destroyed); writes U (zero
// below the diagonal) and X = U^{-1} (zero below). info (device int): 0, or the LAPACK potrf index j > 0 of the
// first non-positive pivot. tmp: n * nb elements.
void potrf_inverse(hipStream_t s, int n, double* a, int lda, double* u, int ldu, double* x, int ldx, double* tmp,
                   int* info);

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
// Band storage of the chase: element (i, j), 0 <= i - j <= 2b - 1, at i + band_ld(b) j.
__host__ __device__ inline int band_ld(int b) { return b > 1 ? 2 * b - 1 : 1; }
__host__ __device__ inline int64_t band_elems(int n, int b) { return (int64_t)band_ld(b) * n + n + 1; }
// Band (ab, leading dimension 2b, (2b + 1) n elements) -> tridiagonal d, e; reflectors w into v2 (n x n, column s
// holds sweep s, rows rs ..< re).
template <typename T>
void chase_bulges(hipStream_t s, int n, int b, T* ab, T* v2, T* d, T* e, int* flags);

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
  explicit DcPlan(int n);
  ~DcPlan();
  DcPlan(const DcPlan&) = delete;
  DcPlan& operator=(const DcPlan&) = delete;
};
size_t dc_workspace_doubles(int n);
size_t dc_workspace_ints(int n);
// Eigenvalues (ascending, into d) and eigenvectors (z, n x n, ldz) of the tridiagonal (d, e), FP64; e is destroyed.
// info: device int, set to a positive value when a secular equation fails to converge.
void stedc(hipStream_t s, const DcPlan& plan, double* d, double* e, double* z, int ldz, double* work, int* iwork,
           int* info);
void dc_merge_probe(hipStream_t s, int n, int n1, double rho, double* d, double* q, int ldq, double* work,
                    int* iwork, int* info, double* d_defl_out, double* q2_out, int* kctot_out, double* dlamda_out,
                    double* w_out, int* indxc_out, double* rho_out);

}  // namespace xtbeig

namespace xtbeig {
// wy.hip -----------------------------------------------------------------------------------------------------------
// One compact-WY block Z(row0 .. row0 + height, :) <- (I - V T V^T) Z(...): V height x count (ld ldv, at v_off),
// T count x count upper (ld ldt, at t_off). s0, k: the bulge block's first sweep and bulge index.
struct WyBlock {
  int row0, height, count, s0, k, ldv, ldt, pad;
  int64_t v_off, t_off;
};
// The block lists of the two back-transformations for one (n, b), uploaded once.
struct WyPlan {
  int n = 0, b = 0, g = 0, nbulge = 0, npanel = 0;
  WyBlock* bulge_dev = nullptr;
  WyBlock* panel_dev = nullptr;
  WyPlan(int n, int b);
  ~WyPlan();
  WyPlan(const WyPlan&) = delete;
  WyPlan& operator=(const WyPlan&) = delete;
  size_t v_elems() const;  // workspace for the bulge blocks' V and T
  size_t t_elems() const;
};
int bulge_group(int b);
template <typename T>
void back_transform_bulges_wy(hipStream_t s, const WyPlan& p, const T* v2, T* z, int ldz, int m, T* vwork,
                              T* twork);
template <typename T>
void back_transform_band_wy(hipStream_t s, const WyPlan& p, const T* v1, const T* t1, T* z, int ldz, int m);

// refine.hip -------------------------------------------------------------------------------------------------------
size_t refine_workspace(int n);  // doubles; iwork n ints
void refine_oa(hipStream_t s, int n, const double* h, int ldh, const double* sm, int lds, double* c, int ldc,
               double* w, int iterations, double* work, int* iwork);
}  // namespace xtbeig
