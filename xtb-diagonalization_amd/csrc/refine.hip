// FP64 refinement of approximate generalized eigenpairs (H C = S C Lambda, C^T S C = I), Ogita-Aishima style with
// the S inner product, an a-posteriori check of the result, and the eigenpair sort. The reference refines FP32
// eigenpairs in FP64 (scripts/kernels/fp32_error.c:95-158; the S-inner-product variant and the Rayleigh-Ritz step on
// clusters in scripts/kernels/refine.f90:254-325):
//     R = I - C^T S C,  A = C^T H C,  lambda_i = A_ii / (1 - R_ii),
//     E_ij = (A_ij + lambda_j R_ij) / (lambda_j - lambda_i) for resolved pairs i != j,
//     C <- C (I + E).
// The four products and the update run as FP64 GEMMs on the matrix cores; R and A are formed on their upper
// triangles and mirrored, so both are exactly symmetric (E_ij + E_ji = R_ij then holds to rounding for every
// resolved pair, which is what keeps the new C S-orthonormal to second order).
//
// Resolved pairs and clusters. A pair i != j is resolved when both of its Newton corrections are small,
// |E_ij| <= TAU and |E_ji| <= TAU (the decision is symmetric, so a pair is never half resolved). Larger corrections
// mean the gap is not large against the current error: the linearized step would leave O(E^2) errors behind. Such
// pairs are grouped into clusters, the smallest runs of consecutive columns that contain every unresolved pair
// (columns are in ascending eigenvalue order), and each cluster gets a Rayleigh-Ritz step: the k x k generalized
// problem (A_cc, I - R_cc) is solved exactly (Cholesky of the Gram block, cyclic Jacobi), Y its G-orthonormal
// eigenvectors, and the cluster's columns of I + E become (I + E_off) Y (the Newton terms of the cluster's vectors
// towards the other columns rotated with them). Clusters larger than CLUSTER_MAX keep E_ij = R_ij / 2 inside: their
// subspace converges and the basis inside it is not rotated.
//
// The diagonal of the other columns carries the second-order terms of the new Gram matrix,
// diag((I + E)^T (I - R) (I + E)) = 1 - R_ii + 2 E_ii + sum_k E_ki^2 - 2 sum_k R_ik E_ki + O(3):
//     E_ii = (R_ii - sum_{k != i} E_ki^2 + 2 sum_{k != i} R_ik E_ki) / 2.
//
// Check. After the last step the result is verified without the oracle: for 8 random sign vectors v (fixed
// pseudo-random), ||(H C - S C Lambda) v|| and ||(C^T S C - I) v|| estimate the Frobenius norms of the residual and
// of the orthogonality error (E ||M v||^2 = ||M||_F^2), from four skinny products. The solve fails the check when
// the residual estimate exceeds 1e-13 n ||H||_F / sqrt(n) (||H||_F / sqrt(n) <= ||H||_2, so this is at least as
// strict as the residual gate 1e-13 n ||H||_2) or the orthogonality estimate exceeds 1e-12 n (the Frobenius norm
// bounds the largest element). A failure sets the caller's status word, which the SCC schedule and the host-array
// mixed solve turn into an FP64 solve of the same problem.
#include "common.hpp"
#include "kernels.hpp"

namespace xtbeig {

namespace {

// Newton corrections above 1e-3 go to the cluster step: a resolved pair's linearized update leaves O(E^2) behind,
// and with TAU = 1e-2 pairs at gaps near 1e-5 left C^T S C - I at 1e-9 after two steps from FP32 (above the
// orthogonality gate at n = 1204), against 1e-11 with 1e-3.
constexpr double TAU = 1e-3;
constexpr int CLUSTER_MAX = 32;
constexpr int CHECK_VECS = 8;
constexpr int SCAN_THREADS = 1024;
constexpr int SCAN_CHUNK = 4;  // n <= SCAN_THREADS * SCAN_CHUNK
constexpr int CHECK_BLOCKS = 128;

// R <- I - R (R = C^T S C on entry), lambda_i = A_ii / (C^T S C)_ii, reach_i = i. Only the upper triangles of R and
// A are formed (the products are symmetric), so each lower element (j, i) is set from (i, j). One workgroup per
// 64 x 64 tile (I, J), I <= J, of the upper triangle: the tile is read and written down its columns, and its mirror
// image, staged transposed in LDS, is written down the columns of tile (J, I).
__global__ __launch_bounds__(256) void oa_prepare_kernel(int n, double* __restrict__ r, int ldr,
                                                         double* __restrict__ a, int lda, double* __restrict__ lam,
                                                         int* __restrict__ reach, Zb zb) {
  r = zb.at(r);
  a = zb.at(a);
  lam = zb.at(lam);
  reach = zb.at(reach);
  const int I = blockIdx.x, J = blockIdx.y;
  if (I > J) return;
  __shared__ double tr[64][65], ta[64][65];
  const int t = threadIdx.x, i0 = I * 64, j0 = J * 64;
#pragma unroll 4
  for (int e = t; e < 64 * 64; e += 256) {
    const int ii = e % 64, jj = e / 64, i = i0 + ii, j = j0 + jj;
    if (i >= n || j >= n || i > j) continue;
    const double v = r[idx(i, j, ldr)];
    if (i == j) {
      lam[i] = a[idx(i, i, lda)] / v;
      r[idx(i, i, ldr)] = 1.0 - v;
      reach[i] = i;
    } else {
      r[idx(i, j, ldr)] = -v;
      tr[jj][ii] = -v;
      ta[jj][ii] = a[idx(i, j, lda)];
    }
  }
  __syncthreads();
#pragma unroll 4
  for (int e = t; e < 64 * 64; e += 256) {
    const int jj = e % 64, ii = e / 64, i = i0 + ii, j = j0 + jj;  // lower element (j, i) = upper (i, j)
    if (i >= n || j >= n || i >= j) continue;
    r[idx(j, i, ldr)] = tr[jj][ii];
    a[idx(j, i, lda)] = ta[jj][ii];
  }
}

// E_ij for every pair; an unresolved pair (i, j), i < j, extends reach_i to j.
__global__ void oa_update_kernel(int n, const double* __restrict__ r, int ldr, const double* __restrict__ a, int lda,
                                 const double* __restrict__ lam, double* __restrict__ e, int lde,
                                 int* __restrict__ reach, Zb zb) {
  r = zb.at(r);
  a = zb.at(a);
  lam = zb.at(lam);
  e = zb.at(e);
  reach = zb.at(reach);
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= n) return;
  const double rij = r[idx(i, j, ldr)];
  if (i == j) {
    e[idx(i, j, lde)] = 0.5 * rij;  // replaced by the diagonal kernel or the cluster step
    return;
  }
  const double aij = a[idx(i, j, lda)];
  const double gap = lam[j] - lam[i];
  const double vij = (aij + lam[j] * rij) / gap;   // correction of column j along column i
  const double vji = (aij + lam[i] * rij) / -gap;  // and of column i along column j (A, R symmetric)
  const bool resolved = fabs(vij) <= TAU && fabs(vji) <= TAU;  // false for a zero gap (inf or NaN)
  e[idx(i, j, lde)] = resolved ? vij : 0.5 * rij;
  if (!resolved && j > i) atomicMax(&reach[i], j);
}

// Clusters from reach: the boundary after column k lies where max_{i <= k} reach_i = k. Writes, per column, the
// first column of its cluster and the cluster's size (1: not in a cluster). One workgroup; each thread owns
// SCAN_CHUNK consecutive columns.
__global__ __launch_bounds__(SCAN_THREADS) void cluster_scan_kernel(int n, const int* __restrict__ reach,
                                                                    int* __restrict__ cstart,
                                                                    int* __restrict__ csize, Zb zb) {
  reach = zb.at(reach);
  cstart = zb.at(cstart);
  csize = zb.at(csize);
  __shared__ int buf[SCAN_THREADS];
  const int t = threadIdx.x, i0 = t * SCAN_CHUNK;
  // 1. prefix maximum of reach -> boundary flags
  int loc = -1;
  for (int q = 0; q < SCAN_CHUNK; ++q)
    if (i0 + q < n) loc = max(loc, reach[i0 + q]);
  buf[t] = loc;
  __syncthreads();
  for (int off = 1; off < SCAN_THREADS; off <<= 1) {
    const int v = t >= off ? buf[t - off] : -1;
    __syncthreads();
    buf[t] = max(buf[t], v);
    __syncthreads();
  }
  int run = t > 0 ? buf[t - 1] : -1;
  __syncthreads();
  bool bnd[SCAN_CHUNK];
  int lastb = -1;  // last boundary column in this chunk
  for (int q = 0; q < SCAN_CHUNK; ++q) {
    const int i = i0 + q;
    bnd[q] = false;
    if (i < n) {
      run = max(run, reach[i]);
      bnd[q] = run == i;
      if (bnd[q]) lastb = i;
    }
  }
  // 2. start_i = 1 + the last boundary before i (prefix maximum of boundary positions)
  buf[t] = lastb;
  __syncthreads();
  for (int off = 1; off < SCAN_THREADS; off <<= 1) {
    const int v = t >= off ? buf[t - off] : -1;
    __syncthreads();
    buf[t] = max(buf[t], v);
    __syncthreads();
  }
  int prevb = t > 0 ? buf[t - 1] : -1;
  __syncthreads();
  int start[SCAN_CHUNK];
  int firstb = n;  // first boundary column in this chunk
  for (int q = 0; q < SCAN_CHUNK; ++q) {
    start[q] = prevb + 1;
    if (i0 + q < n && bnd[q]) {
      prevb = i0 + q;
      if (firstb == n) firstb = i0 + q;
    }
  }
  // 3. end_i = the first boundary at or after i (suffix minimum)
  buf[t] = firstb;
  __syncthreads();
  for (int off = 1; off < SCAN_THREADS; off <<= 1) {
    const int v = t + off < SCAN_THREADS ? buf[t + off] : n;
    __syncthreads();
    buf[t] = min(buf[t], v);
    __syncthreads();
  }
  int nextb = t + 1 < SCAN_THREADS ? buf[t + 1] : n;
  for (int q = SCAN_CHUNK - 1; q >= 0; --q) {
    const int i = i0 + q;
    if (i >= n) continue;
    if (bnd[q]) nextb = i;
    cstart[i] = start[q];
    csize[i] = nextb - start[q] + 1;
  }
}

// Rayleigh-Ritz on one cluster (one wavefront per candidate first column a0): Y with Y^T G Y = I and Y^T A Y =
// diag(mu) ascending, G = I - R_cc; Y into ybuf rows a0 ..< a0 + k (ld n, k columns), mu into lam[a0 ..< a0 + k].
__global__ __launch_bounds__(64) void cluster_solve_kernel(int n, const double* __restrict__ r, int ldr,
                                                           const double* __restrict__ a, int lda,
                                                           const int* __restrict__ cstart,
                                                           const int* __restrict__ csize, double* __restrict__ ybuf,
                                                           double* __restrict__ lam, int* __restrict__ status, Zb zb) {
  r = zb.at(r);
  a = zb.at(a);
  cstart = zb.at(cstart);
  csize = zb.at(csize);
  ybuf = zb.at(ybuf);
  lam = zb.at(lam);
  status = zb.at(status);
  const int a0 = blockIdx.x, t = threadIdx.x;
  if (cstart[a0] != a0) return;
  const int k = csize[a0];
  if (k < 2 || k > CLUSTER_MAX) return;
  __shared__ double G[CLUSTER_MAX][CLUSTER_MAX + 1], B[CLUSTER_MAX][CLUSTER_MAX + 1], V[CLUSTER_MAX][CLUSTER_MAX + 1];
  __shared__ double rot[2];
  __shared__ int flag;
  for (int q = t; q < k * k; q += 64) {
    const int ii = q % k, jj = q / k;
    G[ii][jj] = (ii == jj ? 1.0 : 0.0) - r[idx(a0 + ii, a0 + jj, ldr)];
    B[ii][jj] = a[idx(a0 + ii, a0 + jj, lda)];
    V[ii][jj] = ii == jj ? 1.0 : 0.0;
  }
  if (t == 0) flag = 0;
  __syncthreads();
  // Cholesky G = L L^T (L in the lower triangle of G), right-looking
  for (int j = 0; j < k; ++j) {
    const double d = G[j][j];
    __syncthreads();
    if (!(d > 0.0)) {
      if (t == 0) flag = 1;
      break;
    }
    const double sd = sqrt(d);
    if (t == 0) G[j][j] = sd;
    if (t > j && t < k) G[t][j] /= sd;
    __syncthreads();
    for (int q = t; q < (k - j - 1) * (k - j - 1); q += 64) {
      const int ii = j + 1 + q % (k - j - 1), ll = j + 1 + q / (k - j - 1);
      if (ll <= ii) G[ii][ll] -= G[ii][j] * G[ll][j];
    }
    __syncthreads();
  }
  __syncthreads();
  if (flag) {
    if (t == 0) atomicOr(status, 1);
    return;
  }
  // B <- L^-1 A L^-T: W = L^-1 A (lane t solves column t), then B = L^-1 W^T (W^T = A^T L^-T = (L^-1 A)^T, and
  // L^-1 (L^-1 A)^T = L^-1 A L^-T since A is symmetric)
  for (int pass = 0; pass < 2; ++pass) {
    if (t < k) {
      for (int i = 0; i < k; ++i) {
        double v = B[i][t];
        for (int p = 0; p < i; ++p) v -= G[i][p] * B[p][t];
        B[i][t] = v / G[i][i];
      }
    }
    __syncthreads();
    for (int q0 = 0; q0 < k * k; q0 += 64) {  // transpose in place (uniform trip count around the barrier)
      const int q = q0 + t, ii = q % k, jj = q / k;
      const bool mine = q < k * k && ii < jj;
      const double x = mine ? B[ii][jj] : 0.0, y = mine ? B[jj][ii] : 0.0;
      __syncthreads();
      if (mine) {
        B[ii][jj] = y;
        B[jj][ii] = x;
      }
      __syncthreads();
    }
  }
  // cyclic Jacobi on B, rotations accumulated in V
  for (int sweep = 0; sweep < 30; ++sweep) {
    double off = 0.0, dia = 0.0;
    for (int q = t; q < k * k; q += 64) {
      const int ii = q % k, jj = q / k;
      if (ii < jj) off += B[ii][jj] * B[ii][jj];
      if (ii == jj) dia += B[ii][jj] * B[ii][jj];
    }
    off = wave_sum(off);
    dia = wave_sum(dia);
    if (!(off > 1e-34 * dia) || off == 0.0) break;
    for (int p = 0; p < k - 1; ++p)
      for (int q = p + 1; q < k; ++q) {
        if (t == 0) {
          const double bpq = B[p][q];
          double c = 1.0, s = 0.0;
          if (bpq != 0.0) {
            const double th = (B[q][q] - B[p][p]) / (2.0 * bpq);
            const double tt = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
            c = 1.0 / sqrt(tt * tt + 1.0);
            s = tt * c;
          }
          rot[0] = c;
          rot[1] = s;
        }
        __syncthreads();
        const double c = rot[0], s = rot[1];
        if (t < k) {  // columns p, q of B and V
          const double bp = B[t][p], bq = B[t][q];
          B[t][p] = c * bp - s * bq;
          B[t][q] = s * bp + c * bq;
          const double vp = V[t][p], vq = V[t][q];
          V[t][p] = c * vp - s * vq;
          V[t][q] = s * vp + c * vq;
        }
        __syncthreads();
        if (t < k) {  // rows p, q of B
          const double bp = B[p][t], bq = B[q][t];
          B[p][t] = c * bp - s * bq;
          B[q][t] = s * bp + c * bq;
        }
        __syncthreads();
      }
  }
  // Y = L^-T V (lane t solves column t), columns ordered by ascending mu
  if (t < k) {
    for (int i = k - 1; i >= 0; --i) {
      double v = V[i][t];
      for (int p = i + 1; p < k; ++p) v -= G[p][i] * V[p][t];
      V[i][t] = v / G[i][i];
    }
  }
  __syncthreads();
  if (t < k) {
    const double mu = B[t][t];
    int rank = 0;
    for (int q = 0; q < k; ++q) {
      const double mq = B[q][q];
      rank += (mq < mu) || (mq == mu && q < t);
    }
    lam[a0 + rank] = mu;
    for (int i = 0; i < k; ++i) ybuf[idx(a0 + i, rank, n)] = V[i][t];
  }
}

// 1 + E_ii from column i of E and R (one workgroup per column; R symmetric); columns of a solved cluster are set by
// the cluster step instead.
__global__ __launch_bounds__(256) void oa_diag_kernel(int n, const double* __restrict__ r, int ldr,
                                                      double* __restrict__ e, int lde,
                                                      const int* __restrict__ csize, Zb zb) {
  r = zb.at(r);
  e = zb.at(e);
  csize = zb.at(csize);
  __shared__ double red[16];
  const int i = blockIdx.x, t = threadIdx.x;
  if (csize[i] >= 2 && csize[i] <= CLUSTER_MAX) return;
  const double* rc = r + (int64_t)i * ldr;
  const double* ec = e + (int64_t)i * lde;
  double s1 = 0, s2 = 0;
  for (int k = t; k < n; k += blockDim.x) {
    if (k == i) continue;
    const double ek = ec[k];
    s1 += ek * ek;
    s2 += rc[k] * ek;
  }
  s1 = block_sum(s1, red);
  s2 = block_sum(s2, red);
  if (t == 0) e[idx(i, i, lde)] = 1.0 + 0.5 * (rc[i] - s1 + 2.0 * s2);  // I + E: the update is one product C (I + E)
}

// The cluster columns of I + E: rows inside the cluster Y, rows outside (E_off Y) (one thread per row, grid.y over
// candidate first columns).
__global__ __launch_bounds__(256) void cluster_apply_kernel(int n, double* __restrict__ e, int lde,
                                                            const int* __restrict__ cstart,
                                                            const int* __restrict__ csize,
                                                            const double* __restrict__ ybuf, Zb zb) {
  e = zb.at(e);
  cstart = zb.at(cstart);
  csize = zb.at(csize);
  ybuf = zb.at(ybuf);
  const int a0 = blockIdx.y;
  if (cstart[a0] != a0) return;
  const int k = csize[a0];
  if (k < 2 || k > CLUSTER_MAX) return;
  __shared__ double Y[CLUSTER_MAX][CLUSTER_MAX + 1];
  for (int q = threadIdx.x; q < k * k; q += blockDim.x) Y[q % k][q / k] = ybuf[idx(a0 + q % k, q / k, n)];
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i >= a0 && i < a0 + k) {
    for (int j = 0; j < k; ++j) e[idx(i, a0 + j, lde)] = Y[i - a0][j];
    return;
  }
  double ev[CLUSTER_MAX];
#pragma unroll
  for (int l = 0; l < CLUSTER_MAX; ++l) ev[l] = l < k ? e[idx(i, a0 + l, lde)] : 0.0;
  for (int j = 0; j < k; ++j) {
    double s = 0.0;
#pragma unroll
    for (int l = 0; l < CLUSTER_MAX; ++l)
      if (l < k) s += ev[l] * Y[l][j];
    e[idx(i, a0 + j, lde)] = s;
  }
}

// rank of lambda_i in ascending order (ties by index): RANK_G lanes per i, each counting every RANK_G-th j
constexpr int RANK_G = 16;
__global__ __launch_bounds__(256) void rank_kernel(int n, const double* __restrict__ lam, int* __restrict__ rank,
                                                   Zb zb) {
  lam = zb.at(lam);
  rank = zb.at(rank);
  const int g = (blockIdx.x * blockDim.x + threadIdx.x) / RANK_G, gl = threadIdx.x % RANK_G;
  const int i = min(g, n - 1);  // lanes past n count for a real i and store nothing
  const double li = lam[i];
  int r = 0;
  for (int j = gl; j < n; j += RANK_G) {
    const double lj = lam[j];
    r += (lj < li) || (lj == li && j < i);
  }
#pragma unroll
  for (int m = RANK_G / 2; m > 0; m >>= 1) r += __shfl_xor(r, m, RANK_G);
  if (g < n && gl == 0) rank[g] = r;
}

// dst(:, rank[j]) = src(:, j), w[rank[j]] = lam[j]
template <typename T>
__global__ void scatter_columns_kernel(int n, const int* __restrict__ rank, const T* __restrict__ src, int lds,
                                       T* __restrict__ dst, int ldd, const double* __restrict__ lam,
                                       double* __restrict__ w, Zb zb) {
  rank = zb.at(rank);
  src = zb.at(src);
  dst = zb.at(dst);
  lam = zb.at(lam);
  w = zb.at(w);
  const int j = blockIdx.y;
  const int r = rank[j];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    dst[idx(i, r, ldd)] = src[idx(i, j, lds)];
  if (blockIdx.x == 0 && threadIdx.x == 0 && lam != nullptr) w[r] = lam[j];
}

// The check's probe vectors: pv = [V | Lambda V], V (n x CHECK_VECS) of signs from a hash of (i, k).
__global__ void check_vectors_kernel(int n, const double* __restrict__ w, double* __restrict__ pv, Zb zb) {
  w = zb.at(w);
  pv = zb.at(pv);
  const int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
  if (i >= n) return;
  uint32_t x = (uint32_t)i * 0x9E3779B1u + (uint32_t)k * 0x85EBCA77u + 0x165667B1u;
  x ^= x >> 15;
  x *= 0x2C1B3C6Du;
  x ^= x >> 12;
  x *= 0x297A2D39u;
  x ^= x >> 15;
  const double v = (x & 1u) ? 1.0 : -1.0;
  pv[idx(i, k, n)] = v;
  pv[idx(i, CHECK_VECS + k, n)] = w[i] * v;
}

// Partial sums of ||H Y1 - S Y2||^2, ||C^T S Y1 - V||^2 (ld n, CHECK_VECS columns each) and ||H||_F^2 per block.
__global__ __launch_bounds__(256) void check_partial_kernel(int n, const double* __restrict__ hy,
                                                            const double* __restrict__ sy,
                                                            const double* __restrict__ o,
                                                            const double* __restrict__ pv,
                                                            const double* __restrict__ h, int ldh,
                                                            double* __restrict__ part, Zb zb) {
  hy = zb.at(hy);
  sy = zb.at(sy);
  o = zb.at(o);
  pv = zb.at(pv);
  h = zb.at(h);
  part = zb.at(part);
  __shared__ double red[16];
  const int64_t nv = (int64_t)n * CHECK_VECS;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x, t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double sr = 0, so = 0, sh = 0;
  for (int64_t q = t0; q < nv; q += stride) {
    const double dr = hy[q] - sy[q + nv];  // S Y2 is the second half of sy
    const double d0 = o[q] - pv[q];
    sr += dr * dr;
    so += d0 * d0;
  }
  for (int j = blockIdx.x; j < n; j += gridDim.x)
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const double x = h[idx(i, j, ldh)];
      sh += x * x;
    }
  sr = block_sum(sr, red);
  so = block_sum(so, red);
  sh = block_sum(sh, red);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = sr;
    part[3 * blockIdx.x + 1] = so;
    part[3 * blockIdx.x + 2] = sh;
  }
}

// The check's verdict: ratios[0] = residual estimate / (1e-13 n ||H||_F / sqrt(n)), ratios[1] = orthogonality
// estimate / (1e-12 n); the status word is set when either is above 1 (or not a number).
__global__ void check_final_kernel(int n, int nblocks, const double* __restrict__ part, double* __restrict__ ratios,
                                   int* __restrict__ status, Zb zb) {
  part = zb.at(part);
  ratios = zb.at(ratios);
  status = zb.at(status);
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double sr = 0, so = 0, sh = 0;
  for (int b = 0; b < nblocks; ++b) {
    sr += part[3 * b];
    so += part[3 * b + 1];
    sh += part[3 * b + 2];
  }
  const double res = sqrt(sr / CHECK_VECS), orth = sqrt(so / CHECK_VECS), hn = sqrt(sh / n);
  const double q0 = res / (1e-13 * n * hn), q1 = orth / (1e-12 * n);
  if (ratios) {
    ratios[0] = q0;
    ratios[1] = q1;
  }
  if (!(q0 <= 1.0) || !(q1 <= 1.0)) atomicOr(status, 1);
}

}  // namespace

size_t refine_workspace(int n) {
  return 8 * (size_t)n * n + 2 * (size_t)n + 16 + (size_t)CLUSTER_MAX * n + 56 * (size_t)n + 3 * CHECK_BLOCKS + 64;
}
size_t refine_iwork(int n) { return 4 * (size_t)n; }

// Steps on C (n x n, ldc), then eigenvalues w (ascending) and C's columns sorted to match, then the check (when
// at least one step ran).
void refine_oa(hipStream_t s, int n, const double* h, int ldh, const double* sm, int lds, double* c, int ldc,
               double* w, int iterations, double* work, int* iwork, int* status, double* ratios) {
  // no step: the eigenpairs stay as the FP32 solve left them (already in ascending order), unchecked
  if (iterations <= 0) return;
  if (n > SCAN_THREADS * SCAN_CHUNK) throw std::runtime_error("refine_oa: n too large for the cluster scan");
  if (gemm_thin_parts(n, 2 * CHECK_VECS, n) > (size_t)n * n) throw std::runtime_error("refine_oa: check scratch");
  const size_t nn = (size_t)n * n;
  double* sc = work;
  double* hc = sc + nn;
  double* r = hc + nn;
  double* a = r + nn;
  double* e = a + nn;
  double* cnew = e + nn;
  double* shm = cnew + nn;  // [S | H], each n x n (ld n): the paired products run as one batched launch
  double* lam = shm + 2 * nn;
  double* ybuf = lam + 2 * (size_t)n + 16;
  double* pv = ybuf + (size_t)CLUSTER_MAX * n;  // n x 16
  double* py = pv + 16 * (size_t)n;             // n x 16
  double* psy = py + 16 * (size_t)n;            // n x 16
  double* phy = psy + 16 * (size_t)n;           // n x 8
  double* po = sc;                              // n x 8 (sc is free once the steps are done)
  double* part = phy + 8 * (size_t)n;
  int* rank = iwork;
  int* reach = iwork + n;
  int* cstart = reach + n;
  int* csize = cstart + n;
  const dim3 eg(cdiv(n, 256), n, zgrid());
  const dim3 tg(cdiv(n, 64), cdiv(n, 64), zgrid());
  copy_matrix<double, double>(s, n, n, sm, lds, shm, n);
  copy_matrix<double, double>(s, n, n, h, ldh, shm + nn, n);
  // the current C alternates between c (ldc) and cnew (ld n)
  double* cur = c;
  int ldcur = ldc;
  for (int it = 0; it < iterations; ++it) {
    gemm<double>(s, 'N', 'N', n, n, n, 1.0, shm, n, cur, ldcur, 0.0, sc, n, 2, nn, 0, nn);  // S C, H C
    gemm_upper_tn<double>(s, n, n, 1.0, cur, ldcur, sc, n, 0.0, r, n, 2, 0, nn, nn);  // C^T S C, C^T H C (upper)
    hipLaunchKernelGGL(oa_prepare_kernel, tg, dim3(256), 0, s, n, r, n, a, n, lam, reach, zb());
    XTB_LAUNCH_CHECK();
    hipLaunchKernelGGL(oa_update_kernel, eg, dim3(256), 0, s, n, r, n, a, n, lam, e, n, reach, zb());
    XTB_LAUNCH_CHECK();
    hipLaunchKernelGGL(cluster_scan_kernel, dim3(1, 1, zgrid()), dim3(SCAN_THREADS), 0, s, n, reach, cstart, csize,
                       zb());
    XTB_LAUNCH_CHECK();
    hipLaunchKernelGGL(cluster_solve_kernel, dim3(n, 1, zgrid()), dim3(64), 0, s, n, r, n, a, n, cstart, csize, ybuf,
                       lam, status, zb());
    XTB_LAUNCH_CHECK();
    hipLaunchKernelGGL(oa_diag_kernel, dim3(n, 1, zgrid()), dim3(256), 0, s, n, r, n, e, n, csize, zb());
    XTB_LAUNCH_CHECK();
    hipLaunchKernelGGL(cluster_apply_kernel, eg, dim3(256), 0, s, n, e, n, cstart, csize, ybuf, zb());
    XTB_LAUNCH_CHECK();
    double* nxt = (cur == c) ? cnew : c;
    const int ldnxt = (cur == c) ? n : ldc;
    gemm<double>(s, 'N', 'N', n, n, n, 1.0, cur, ldcur, e, n, 0.0, nxt, ldnxt);  // C (I + E)
    cur = nxt;
    ldcur = ldnxt;
  }
  // sort: columns of the current C scattered into the other buffer in ascending order of lambda, ending in c
  hipLaunchKernelGGL(rank_kernel, dim3(cdiv(n * RANK_G, 256), 1, zgrid()), dim3(256), 0, s, n, lam, rank, zb());
  XTB_LAUNCH_CHECK();
  if (cur == cnew) {
    hipLaunchKernelGGL(scatter_columns_kernel<double>, dim3(cdiv(n, 256), n, zgrid()), dim3(256), 0, s, n, rank, cnew,
                       n, c, ldc, lam, w, zb());
    XTB_LAUNCH_CHECK();
  } else {
    hipLaunchKernelGGL(scatter_columns_kernel<double>, dim3(cdiv(n, 256), n, zgrid()), dim3(256), 0, s, n, rank, c,
                       ldc, cnew, n, lam, w, zb());
    XTB_LAUNCH_CHECK();
    copy_matrix<double, double>(s, n, n, cnew, n, c, ldc);
  }
  // check: Y = C [V | Lambda V], S Y, H Y1, C^T (S Y1)
  hipLaunchKernelGGL(check_vectors_kernel, dim3(cdiv(n, 256), CHECK_VECS, zgrid()), dim3(256), 0, s, n, w, pv, zb());
  XTB_LAUNCH_CHECK();
  // thin products, K split across the GPU; the partials go to r (n x n, free once the steps are done)
  gemm_thin<double>(s, 'N', n, 2 * CHECK_VECS, n, c, ldc, pv, n, py, n, r);
  gemm_thin<double>(s, 'N', n, 2 * CHECK_VECS, n, shm, n, py, n, psy, n, r);
  gemm_thin<double>(s, 'N', n, CHECK_VECS, n, shm + nn, n, py, n, phy, n, r);
  gemm_thin<double>(s, 'T', n, CHECK_VECS, n, c, ldc, psy, n, po, n, r);
  hipLaunchKernelGGL(check_partial_kernel, dim3(CHECK_BLOCKS, 1, zgrid()), dim3(256), 0, s, n, phy, psy, po, pv,
                     shm + nn, n, part, zb());
  XTB_LAUNCH_CHECK();
  hipLaunchKernelGGL(check_final_kernel, dim3(1, 1, zgrid()), dim3(64), 0, s, n, CHECK_BLOCKS, part, ratios, status,
                     zb());
  XTB_LAUNCH_CHECK();
}

}  // namespace xtbeig
