// Divide and conquer for the symmetric tridiagonal eigenproblem, FP64, all eigenpairs (dstedc('I') semantics:
// eigenvalues ascending, eigenvectors of T). The reference tears at n1 = n / 2 (a rank-one change, as dlaed0), solves
// both halves and merges them as dlaed1: deflation by LAPACK's dlaed2, a port of dlaed3 (one dlaed4 per root,
// Gu-Eisenstat z, merge GEMMs), dlamrg and a column sort (Stages/ParallelDivideAndConquer.swift:14-211).
//
// Here the same tear is applied recursively down to 1 x 1 leaves, so that every subproblem is a merge and each level
// of the tree is one batch of GPU launches over all its merges:
//   deflate_kernel   one workgroup per merge: z from the halves' boundary rows, dlaed2 (sort, small-z and close-pole
//                    deflation with Givens rotations, the four column groups, Q2) staged in LDS
//   secular_kernel   one lane per secular-equation root (dlaed4 / dlaed5), outputs (origin pole, shift)
//   gu_kernel        Gu-Eisenstat z_i from the products over all roots (dlaed3's order)
//   vectors_kernel   one wavefront per root: the normalized eigenvector of the rank-one problem, rows permuted
//   merge_gemm       Q(0:n1, 0:K) = Q2a S(0:n12), Q(n1:n, 0:K) = Q2b S(c0:c0+n23), MFMA FP64 with device-side sizes
//   order_kernel     dlamrg of the K new and n - K deflated eigenvalues, then gather / copy-back of the columns
// Deflation bookkeeping (K, the index arrays, column types) follows dlaed2 step for step and is compared with the host
// LAPACK's dlaed2 on the same inputs by the tests.
//
// The deflation (deflate_kernel) and the secular solver follow LAPACK's dlaed2 / dlaed4 / dlaed5 / dlaed6 closely
// (their variable names and branch structure), so LAPACK's license applies to those parts:
//   Copyright (c) 1992-2017 The University of Tennessee and The University of Tennessee Research Foundation.
//   Copyright (c) 2000-2017 The University of California Berkeley. Copyright (c) 2006-2017 The University of
//   Colorado Denver. All rights reserved.
//   Redistribution and use in source and binary forms, with or without modification, are permitted provided that the
//   following conditions are met: (1) Redistributions of source code must retain the above copyright notice, this
//   list of conditions and the following disclaimer. (2) Redistributions in binary form must reproduce the above
//   copyright notice, this list of conditions and the following disclaimer listed in this license in the
//   documentation and/or other materials provided with the distribution. (3) Neither the name of the copyright
//   holders nor the names of its contributors may be used to endorse or promote products derived from this software
//   without specific prior written permission.
//   The copyright holders provide no reassurances that the source code provided does not infringe any patent,
//   copyright, or any other intellectual property rights of third parties. The copyright holders disclaim any
//   liability to any recipient for claims brought against recipient by any third party for infringement of that
//   party's intellectual property rights.
//   THIS SOFTWARE IS PROVIDED BY THE COPYRIGHT HOLDERS AND CONTRIBUTORS "AS IS" AND ANY EXPRESS OR IMPLIED
//   WARRANTIES, INCLUDING, BUT NOT LIMITED TO, THE IMPLIED WARRANTIES OF MERCHANTABILITY AND FITNESS FOR A PARTICULAR
//   PURPOSE ARE DISCLAIMED. IN NO EVENT SHALL THE COPYRIGHT OWNER OR CONTRIBUTORS BE LIABLE FOR ANY DIRECT, INDIRECT,
//   INCIDENTAL, SPECIAL, EXEMPLARY, OR CONSEQUENTIAL DAMAGES (INCLUDING, BUT NOT LIMITED TO, PROCUREMENT OF SUBSTITUTE
//   GOODS OR SERVICES; LOSS OF USE, DATA, OR PROFITS; OR BUSINESS INTERRUPTION) HOWEVER CAUSED AND ON ANY THEORY OF
//   LIABILITY, WHETHER IN CONTRACT, STRICT LIABILITY, OR TORT (INCLUDING NEGLIGENCE OR OTHERWISE) ARISING IN ANY WAY
//   OUT OF THE USE OF THIS SOFTWARE, EVEN IF ADVISED OF THE POSSIBILITY OF SUCH DAMAGE.
#include "common.hpp"
#include "gemm_tile.hpp"
#include "kernels.hpp"
#include "secular.hpp"

#include <vector>

namespace xtbeig {

namespace {

constexpr double EPS = 1.1102230246251565e-16;

using Node = DcNode;
// per merge results: K, ctot[4], Givens rotations recorded by the deflation
struct NodeOut {
  int k, ctot[4], nrot, pad[2];
};

__device__ inline double lapy2(double x, double y) {
  const double xa = fabs(x), ya = fabs(y);
  const double w = fmax(xa, ya), z = fmin(xa, ya);
  if (z == 0 || w > 1.79e308) return w;
  const double r = z / w;
  return w * sqrt(1 + r * r);
}

// dlamrg: index (1-based values) merging a[0:n1] (stride s1) with a[n1:n1+n2] (stride s2) into ascending order.
__device__ inline void lamrg(int n1, int n2, const double* a, int s1, int s2, int* index) {
  int n1sv = n1, n2sv = n2;
  int ind1 = s1 > 0 ? 1 : n1;
  int ind2 = s2 > 0 ? 1 + n1 : n1 + n2;
  int i = 0;
  while (n1sv > 0 && n2sv > 0) {
    if (a[ind1 - 1] <= a[ind2 - 1]) { index[i++] = ind1; ind1 += s1; --n1sv; }
    else { index[i++] = ind2; ind2 += s2; --n2sv; }
  }
  if (n1sv == 0) { for (int q = 0; q < n2sv; ++q) { index[i++] = ind2; ind2 += s2; } }
  else { for (int q = 0; q < n1sv; ++q) { index[i++] = ind1; ind1 += s1; } }
}

// Scale by the max norm, apply every tear (pre-order, parent before children), Q = I.
__global__ __launch_bounds__(1024) void prepare_kernel(int n, double* d, double* e, const int* __restrict__ split_pre,
                                                       double* rho, double* scale, Zb zb) {
  d = zb.at(d);
  e = zb.at(e);
  rho = zb.at(rho);
  scale = zb.at(scale);
  __shared__ double red[16];
  const int t = threadIdx.x;
  double m = 0;
  for (int i = t; i < n; i += blockDim.x) m = fmax(m, fabs(d[i]));
  for (int i = t; i < n - 1; i += blockDim.x) m = fmax(m, fabs(e[i]));
  m = wave_max(m);
  if ((t & 63) == 0) red[t >> 6] = m;
  __syncthreads();
  double nrm = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) nrm = fmax(nrm, red[w]);
  if (nrm == 0) nrm = 1;
  if (t == 0) *scale = nrm;
  for (int i = t; i < n - 1; i += blockDim.x) e[i] /= nrm;
  __syncthreads();
  // every tear at row m (merge p in pre-order) subtracts |e[m - 1]| from d[m - 1] and d[m]: row i meets the tears at
  // i (as d[m]) and at i + 1 (as d[m - 1]), applied in pre-order, as the tree's recursion applies them
  for (int i = t; i < n; i += blockDim.x) {
    double di = d[i] / nrm;
    const int pa = (i >= 1) ? split_pre[i] : -1, pb = (i + 1 <= n - 1) ? split_pre[i + 1] : -1;
    const double va = (pa >= 0) ? fabs(e[i - 1]) : 0.0, vb = (pb >= 0) ? fabs(e[i]) : 0.0;
    if (pa >= 0 && pb >= 0) {
      if (pa < pb) { di -= va; di -= vb; }
      else { di -= vb; di -= va; }
    } else if (pa >= 0) {
      di -= va;
    } else if (pb >= 0) {
      di -= vb;
    }
    d[i] = di;
    if (pa >= 0) rho[pa] = e[i - 1];
  }
}

__global__ void unscale_kernel(int n, double* d, const double* scale, Zb zb) {
  d = zb.at(d);
  scale = zb.at(scale);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] *= *scale;
}

// Exclusive scan of one int per thread over a 256-thread workgroup in thread order: sums (MAX false, identity 0) or
// maxima (MAX true, identity -1). Within a wavefront by lane shifts, then the four wavefront totals through part
// (LDS, 4 ints); *total gets the whole workgroup's. Every thread calls it (two barriers).
template <bool MAX>
__device__ inline int block_exclusive_scan(int v, int* part, int* total = nullptr) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  constexpr int ID = MAX ? -1 : 0;
  int x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(x, d, 64);
    if (lane >= d) x = MAX ? max(x, y) : x + y;
  }
  if (lane == 63) part[wid] = x;
  __syncthreads();
  int pre = ID, tot = ID;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const int pw = part[w];
    if (w < wid) pre = MAX ? max(pre, pw) : pre + pw;
    tot = MAX ? max(tot, pw) : tot + pw;
  }
  int ex = __shfl_up(x, 1, 64);
  if (lane == 0) ex = ID;
  if (total != nullptr) *total = tot;
  __syncthreads();  // part free again
  return MAX ? max(pre, ex) : pre + ex;
}

// dlaed2 for one merge (one workgroup). LDS: d, z, dlamda, w (doubles) and indxq, indx, indxc, indxp, coltyp (ints).
__global__ __launch_bounds__(256) void deflate_kernel(const Node* nodes, const int* node_pre_index,
                                                      double* __restrict__ d_all, const double* __restrict__ q_all,
                                                      int ldq, const double* __restrict__ rho_pre,
                                                      double* __restrict__ rho_lvl, double* __restrict__ dlamda_all,
                                                      double* __restrict__ w_all, int* __restrict__ indx_all,
                                                      int* __restrict__ indxc_all, int* __restrict__ rotp_all,
                                                      double* __restrict__ rotcs_all, NodeOut* __restrict__ out,
                                                      Zb zb) {
#pragma clang fp contract(off)  // dlaed2's scalar recurrences rounded as written (no fused multiply-add)
  d_all = zb.at(d_all);
  q_all = zb.at(q_all);
  rho_pre = zb.at(rho_pre);
  rho_lvl = zb.at(rho_lvl);
  dlamda_all = zb.at(dlamda_all);
  w_all = zb.at(w_all);
  indx_all = zb.at(indx_all);
  indxc_all = zb.at(indxc_all);
  rotp_all = zb.at(rotp_all);
  rotcs_all = zb.at(rotcs_all);
  out = zb.at(out);
  extern __shared__ char smem[];
  const Node nd = nodes[blockIdx.x];
  const int n = nd.sz, n1 = nd.n1, n2 = n - n1, off = nd.off;
  double* D = reinterpret_cast<double*>(smem);
  double* Z = D + n;
  double* DL = Z + n;
  double* W = DL + n;
  int* INDXQ = reinterpret_cast<int*>(W + n);
  int* INDX = INDXQ + n;
  int* INDXC = INDX + n;
  int* INDXP = INDXC + n;
  int* COLTYP = INDXP + n;
  int* ROT_P = COLTYP + n;  // Givens rotations recorded by the scan (at most n)
  int* ROT_N = ROT_P + n;
  double* ROT_C = reinterpret_cast<double*>(ROT_N + n + (n & 1));
  double* ROT_S = ROT_C + n;
  __shared__ int sh_k, sh_nrot, sh_ctot[4], sh_early, sh_serial;
  __shared__ double sh_rho, sh_tol;
  __shared__ double red_v[256];
  __shared__ int red_i[256], part[4];
  const int t = threadIdx.x, nt = blockDim.x;
  const double* q = q_all + idx(off, off, ldq);
  const double rho_in = rho_pre[node_pre_index[blockIdx.x]];
  const double tt = 1.0 / sqrt(2.0);
  // z from the halves' boundary rows (sign of rho folded into the second half), scaled by 1 / sqrt(2); the halves
  // are sorted, so dlaed2's INDXQ is the identity and DLAMDA = D before the merge
  for (int i = t; i < n; i += nt) {
    const double di = d_all[off + i];
    double zi = (i < n1) ? q[idx(n1 - 1, i, ldq)] : q[idx(n1, i, ldq)];
    if (rho_in < 0 && i >= n1) zi = -zi;
    D[i] = di;
    DL[i] = di;
    Z[i] = zi * tt;
    INDXQ[i] = i + 1;
  }
  __syncthreads();
  const double rho = fabs(2 * rho_in);
  // dlamrg of the two ascending halves. By ranks when both halves are in order (on ties the first half's element
  // comes first, so each rank is the position dlamrg's sequential merge gives the element); a half that rounding left
  // a few ulps out of order (or a NaN) would make ranks collide, and then thread 0 runs dlamrg's sequential merge,
  // which yields a permutation whatever the order.
  if (t == 0) sh_serial = 0;
  __syncthreads();
  for (int i = t; i + 1 < n; i += nt)
    if (i + 1 != n1 && !(DL[i] <= DL[i + 1])) atomicOr(&sh_serial, 1);
  __syncthreads();
  if (sh_serial) {
    if (t == 0) {
      int i1 = 0, i2 = n1, o = 0;
      while (i1 < n1 && i2 < n) {
        const int src = (DL[i1] <= DL[i2]) ? i1++ : i2++;
        INDXC[o] = src + 1;
        INDX[o++] = src + 1;
      }
      while (i1 < n1) {
        INDXC[o] = i1 + 1;
        INDX[o++] = ++i1;
      }
      while (i2 < n) {
        INDXC[o] = i2 + 1;
        INDX[o++] = ++i2;
      }
    }
  } else {
    for (int i = t; i < n; i += nt) {
      int r;
      const double v = DL[i];
      if (i < n1) {
        int lo = 0, hi = n2;  // # second-half elements < v
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (DL[n1 + mid] < v) lo = mid + 1;
          else hi = mid;
        }
        r = i + lo;
      } else {
        int lo = 0, hi = n1;  // # first-half elements <= v
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (DL[mid] <= v) lo = mid + 1;
          else hi = mid;
        }
        r = (i - n1) + lo;
      }
      INDXC[r] = i + 1;
      INDX[r] = i + 1;
    }
  }
  __syncthreads();
  // idamax of z and of d (first index of the largest magnitude)
  {
    double bz = -1, bd = -1;
    int iz = 0, id = 0;
    for (int i = t; i < n; i += nt) {
      if (fabs(Z[i]) > bz) { bz = fabs(Z[i]); iz = i; }
      if (fabs(D[i]) > bd) { bd = fabs(D[i]); id = i; }
    }
    red_v[t] = bz;
    red_i[t] = iz;
    __syncthreads();
    for (int st = nt / 2; st > 0; st >>= 1) {
      if (t < st && (red_v[t + st] > red_v[t] || (red_v[t + st] == red_v[t] && red_i[t + st] < red_i[t]))) {
        red_v[t] = red_v[t + st];
        red_i[t] = red_i[t + st];
      }
      __syncthreads();
    }
    const int imax = red_i[0];
    __syncthreads();
    red_v[t] = bd;
    red_i[t] = id;
    __syncthreads();
    for (int st = nt / 2; st > 0; st >>= 1) {
      if (t < st && (red_v[t + st] > red_v[t] || (red_v[t + st] == red_v[t] && red_i[t + st] < red_i[t]))) {
        red_v[t] = red_v[t + st];
        red_i[t] = red_i[t + st];
      }
      __syncthreads();
    }
    const int jmax = red_i[0];
    if (t == 0) {
      sh_tol = 8 * EPS * fmax(fabs(D[jmax]), fabs(Z[imax]));
      sh_early = (rho * fabs(Z[imax]) <= sh_tol) ? 1 : 0;
      sh_rho = rho;
      sh_serial = 0;
    }
    __syncthreads();
  }
  const double tol = sh_tol;
  // small z (sorted position j) into ROT_P as flags, and whether any two consecutive large-z poles pass dlaed2's
  // close-pole test on the original values: if none does, the scan makes no rotation and its result follows from the
  // flags alone (a rotation changes only the values of the pair it merges, so the first rotation of the serial scan
  // is at the first pair that passes here).
  int* SMALL = ROT_P;
  if (!sh_early) {
    const int per = (n + nt - 1) / nt, j0 = min(n, t * per), j1 = min(n, j0 + per);
    int last = -1;
    for (int j = j0; j < j1; ++j) {
      const int sm = (rho * fabs(Z[INDX[j] - 1]) <= tol) ? 1 : 0;
      SMALL[j] = sm;
      if (!sm) last = j;
    }
    // the last large-z position before this chunk (positions grow with the chunk: the maximum of the earlier ones)
    int prev = block_exclusive_scan<true>(last, part);
    bool close = false;
    for (int j = j0; j < j1; ++j) {
      if (!SMALL[j]) {
        if (prev >= 0) {
          const int pj = INDX[prev], nj = INDX[j];
          double s2 = Z[pj - 1], c2 = Z[nj - 1];
          const double tau = lapy2(c2, s2);
          const double tdiff = D[nj - 1] - D[pj - 1];
          c2 = c2 / tau;
          s2 = -s2 / tau;
          if (fabs(tdiff * c2 * s2) <= tol) close = true;
        }
        prev = j;
      }
    }
    if (close) atomicOr(&sh_serial, 1);
    __syncthreads();
  }
  if (!sh_early && sh_serial) {
    // dlaed2's scan, serial: close poles are merged by Givens rotations, which change D and Z as the scan goes
    if (t == 0) {
      for (int i = 0; i < n1; ++i) COLTYP[i] = 1;
      for (int i = n1; i < n; ++i) COLTYP[i] = 3;
      int k = 0, nrot = 0;
      int k2 = n + 1, j = 1, pj = 0;
      // first non-deflated column (one exists: z[imax] is not small)
      for (; j <= n; ++j) {
        const int nj = INDX[j - 1];
        if (rho * fabs(Z[nj - 1]) <= tol) {
          --k2; COLTYP[nj - 1] = 4; INDXP[k2 - 1] = nj;
        } else {
          pj = nj;
          break;
        }
      }
      if (pj == 0) pj = INDX[n - 1];  // unreachable: z[imax] is not small
      for (++j; j <= n; ++j) {
        const int nj = INDX[j - 1];
        if (rho * fabs(Z[nj - 1]) <= tol) {
          --k2; COLTYP[nj - 1] = 4; INDXP[k2 - 1] = nj;
          continue;
        }
        double s = Z[pj - 1], c = Z[nj - 1];
        const double tau = lapy2(c, s);
        double tdiff = D[nj - 1] - D[pj - 1];
        c = c / tau;
        s = -s / tau;
        if (fabs(tdiff * c * s) <= tol) {
          Z[nj - 1] = tau;
          Z[pj - 1] = 0;
          if (COLTYP[nj - 1] != COLTYP[pj - 1]) COLTYP[nj - 1] = 2;
          COLTYP[pj - 1] = 4;
          ROT_P[nrot] = pj - 1; ROT_N[nrot] = nj - 1; ROT_C[nrot] = c; ROT_S[nrot] = s; ++nrot;
          const double tp = D[pj - 1] * c * c + D[nj - 1] * s * s;
          D[nj - 1] = D[pj - 1] * s * s + D[nj - 1] * c * c;
          D[pj - 1] = tp;
          --k2;
          int i = 1;
          while (true) {
            if (k2 + i <= n) {
              if (D[pj - 1] < D[INDXP[k2 + i - 1] - 1]) {
                INDXP[k2 + i - 2] = INDXP[k2 + i - 1];
                INDXP[k2 + i - 1] = pj;
                ++i;
                continue;
              }
              INDXP[k2 + i - 2] = pj;
            } else {
              INDXP[k2 + i - 2] = pj;
            }
            break;
          }
          pj = nj;
        } else {
          ++k;
          DL[k - 1] = D[pj - 1];
          W[k - 1] = Z[pj - 1];
          INDXP[k - 1] = pj;
          pj = nj;
        }
      }
      ++k;
      DL[k - 1] = D[pj - 1];
      W[k - 1] = Z[pj - 1];
      INDXP[k - 1] = pj;
      sh_nrot = nrot;
    }
    __syncthreads();
  } else if (!sh_early) {
    // no rotation: large-z poles keep their sorted order at 0 ..< K, small-z ones fill n - 1 down to K in the order
    // met (ranks from per-chunk counts)
    const int per = (n + nt - 1) / nt, j0 = min(n, t * per), j1 = min(n, j0 + per);
    int nl = 0;
    for (int j = j0; j < j1; ++j) nl += !SMALL[j];
    const int lb = block_exclusive_scan<false>(nl, part);
    if (t == 0) sh_nrot = 0;
    int kl = lb, ks = j0 - lb;  // large / small met before this chunk
    for (int j = j0; j < j1; ++j) {
      const int nj = INDX[j];
      if (!SMALL[j]) {
        DL[kl] = D[nj - 1];
        W[kl] = Z[nj - 1];
        INDXP[kl] = nj;
        ++kl;
      } else {
        INDXP[n - 1 - ks] = nj;
        ++ks;
      }
    }
    for (int i = t; i < n; i += nt) COLTYP[i] = (i < n1) ? 1 : 3;
    __syncthreads();
    for (int j = j0; j < j1; ++j)
      if (SMALL[j]) COLTYP[INDX[j] - 1] = 4;
    __syncthreads();
  }
  if (!sh_early) {
    // the four column groups: INDX / INDXC by type, stable (dlaed2's PSM pass), from per-chunk counts
    const int per = (n + nt - 1) / nt, j0 = min(n, t * per), j1 = min(n, j0 + per);
    int cnt4[4] = {0, 0, 0, 0};
    for (int j = j0; j < j1; ++j) cnt4[COLTYP[INDXP[j] - 1] - 1]++;
    int ex4[4], tot4[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) ex4[c] = block_exclusive_scan<false>(cnt4[c], part, &tot4[c]);
    const int base[4] = {0, tot4[0], tot4[0] + tot4[1], tot4[0] + tot4[1] + tot4[2]};
    int pos[4];
    for (int c = 0; c < 4; ++c) pos[c] = base[c] + ex4[c];
    for (int j = j0; j < j1; ++j) {
      const int js = INDXP[j], ct = COLTYP[js - 1] - 1;
      INDX[pos[ct]] = js;
      INDXC[pos[ct]] = j + 1;
      pos[ct]++;
    }
    if (t == 0) {
      for (int c = 0; c < 4; ++c) sh_ctot[c] = tot4[c];
      sh_k = n - tot4[3];
    }
    __syncthreads();
  }
  if (sh_early) {
    // K = 0: every column deflates, in sorted order (Q2 = Q(:, INDX), d = D(INDX))
    if (t == 0) {
      sh_k = 0;
      sh_nrot = 0;
      sh_ctot[0] = sh_ctot[1] = sh_ctot[2] = 0;
      sh_ctot[3] = n;
    }
    __syncthreads();
  }
  const int k = sh_k;
  for (int j = t; j < n; j += blockDim.x) {
    indx_all[off + j] = INDX[j];
    indxc_all[off + j] = INDXC[j];
    if (j >= k) d_all[off + j] = D[INDX[j] - 1];
  }
  for (int j = t; j < k; j += blockDim.x) {
    dlamda_all[off + j] = DL[j];
    w_all[off + j] = W[j];
  }
  for (int r = t; r < sh_nrot; r += blockDim.x) {
    rotp_all[2 * (off + r)] = ROT_P[r];
    rotp_all[2 * (off + r) + 1] = ROT_N[r];
    rotcs_all[2 * (off + r)] = ROT_C[r];
    rotcs_all[2 * (off + r) + 1] = ROT_S[r];
  }
  if (t == 0) {
    out[blockIdx.x].k = k;
    for (int c = 0; c < 4; ++c) out[blockIdx.x].ctot[c] = sh_ctot[c];
    out[blockIdx.x].nrot = sh_nrot;
    rho_lvl[blockIdx.x] = sh_rho;
  }
}

// The deflation's Givens rotations on Q's columns, in order; each thread owns one row of the merge's block.
__global__ void rotate_kernel(const Node* nodes, const NodeOut* out, const int* __restrict__ rotp_all,
                              const double* __restrict__ rotcs_all, double* __restrict__ q_all, int ldq, Zb zb) {
  out = zb.at(out);
  rotp_all = zb.at(rotp_all);
  rotcs_all = zb.at(rotcs_all);
  q_all = zb.at(q_all);
  const Node nd = nodes[blockIdx.y];
  const int nrot = out[blockIdx.y].nrot;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (nrot == 0 || i >= nd.sz) return;
  double* row = q_all + nd.off + i + (int64_t)nd.off * ldq;
  for (int r = 0; r < nrot; ++r) {
    const int cp = rotp_all[2 * (nd.off + r)], cn = rotp_all[2 * (nd.off + r) + 1];
    const double c = rotcs_all[2 * (nd.off + r)], sn = rotcs_all[2 * (nd.off + r) + 1];
    const double x = row[(int64_t)cp * ldq], y = row[(int64_t)cn * ldq];
    row[(int64_t)cp * ldq] = c * x + sn * y;
    row[(int64_t)cn * ldq] = c * y - sn * x;
  }
}

// Q2 = [ n1 x n12 | n2 x n23 | n x ctot3 ]: column pos of the grouped order from Q's column INDX(pos) (dlaed2's
// copies). One workgroup per column.
__global__ void q2_gather_kernel(const Node* nodes, const NodeOut* out, const int* __restrict__ indx_all,
                                 const double* __restrict__ q_all, int ldq, double* __restrict__ q2_all, int* info,
                                 Zb zb) {
  info = zb.at(info);
  out = zb.at(out);
  indx_all = zb.at(indx_all);
  q_all = zb.at(q_all);
  q2_all = zb.at(q2_all);
  const Node nd = nodes[blockIdx.y];
  const int pos = blockIdx.x;
  if (pos >= nd.sz) return;
  const NodeOut o = out[blockIdx.y];
  const int n = nd.sz, n1 = nd.n1, n2 = n - n1, k = o.k;
  const int c0 = o.ctot[0], c1 = o.ctot[1], c2 = o.ctot[2];
  const int n12 = c0 + c1, n23 = c1 + c2;
  const double* q = q_all + idx(nd.off, nd.off, ldq);
  double* q2a = q2_all + nd.q2off;
  double* q2b = q2a + (size_t)n1 * n12;
  double* q2c = q2b + (size_t)n2 * n23;
  const int js = indx_all[nd.off + pos] - 1;
  if (js < 0 || js >= n) {  // not a column of this merge (a broken permutation): read nothing, report
    if (threadIdx.x == 0) atomicCAS(info, 0, nd.off + pos + 1);
    return;
  }
  const double* src = q + (int64_t)js * ldq;
  const int t = threadIdx.x;
  if (pos < c0) {
    for (int i = t; i < n1; i += blockDim.x) q2a[idx(i, pos, n1)] = src[i];
  } else if (pos < c0 + c1) {
    for (int i = t; i < n1; i += blockDim.x) q2a[idx(i, pos, n1)] = src[i];
    for (int i = t; i < n2; i += blockDim.x) q2b[idx(i, pos - c0, n2)] = src[n1 + i];
  } else if (pos < c0 + c1 + c2) {
    for (int i = t; i < n2; i += blockDim.x) q2b[idx(i, pos - c0, n2)] = src[n1 + i];
  } else {
    for (int i = t; i < n; i += blockDim.x) q2c[idx(i, pos - k, n)] = src[i];
  }
}

// One lane per root. S (the K x K eigenvector matrix of the rank-one problem, ld sz) is written here only for K <= 2.
constexpr int SEC_G = 32;  // lanes per secular root: each sums every SEC_G-th pole

__global__ __launch_bounds__(64) void secular_kernel(const Node* nodes, const NodeOut* out, const double* rho_lvl,
                                                     const double* __restrict__ dlamda_all,
                                                     const double* __restrict__ w_all,
                                                     const int* __restrict__ indxc_all, double* __restrict__ d_all,
                                                     int* __restrict__ orig_all,
                                                     double* __restrict__ tau_all, double* __restrict__ s_all,
                                                     int* info, int maxit, Zb zb) {
  out = zb.at(out);
  rho_lvl = zb.at(rho_lvl);
  dlamda_all = zb.at(dlamda_all);
  w_all = zb.at(w_all);
  indxc_all = zb.at(indxc_all);
  d_all = zb.at(d_all);
  orig_all = zb.at(orig_all);
  tau_all = zb.at(tau_all);
  s_all = zb.at(s_all);
  info = zb.at(info);
  extern __shared__ double poles[];  // dlamda (k), then z (k): every root reads all of them each iteration
  constexpr int RPB = 64 / SEC_G;    // roots per workgroup
  const Node nd = nodes[blockIdx.y];
  const int k = out[blockIdx.y].k;
  if (blockIdx.x * RPB >= k) return;
  for (int i = threadIdx.x; i < k; i += 64) {
    poles[i] = dlamda_all[nd.off + i];
    poles[k + i] = w_all[nd.off + i];
  }
  __syncthreads();
  const int j = blockIdx.x * RPB + threadIdx.x / SEC_G, gl = threadIdx.x % SEC_G;
  if (j >= k) return;  // whole groups leave together
  const double* dl = poles;
  const double* w = poles + k;
  const double rho = rho_lvl[blockIdx.y];
  double* S = s_all + nd.q2off;
  if (k == 1) {
    if (gl == 0) {
      d_all[nd.off] = dl[0] + rho * w[0] * w[0];
      S[0] = 1;
    }
    return;
  }
  if (k == 2) {
    if (gl != 0) return;
    double v[2], lam;
    const double dd[2] = {dl[0], dl[1]}, zz[2] = {w[0], w[1]};
    secular::laed5(j, dd, zz, rho, v, lam);
    d_all[nd.off + j] = lam;
    const int* ix = indxc_all + nd.off;
    S[idx(0, j, nd.sz)] = v[ix[0] - 1];
    S[idx(1, j, nd.sz)] = v[ix[1] - 1];
    return;
  }
  int orig = 0;
  double tau = 0;
  const int bad = secular::laed4<SEC_G>(k, j, dl, w, rho, orig, tau, gl, maxit);
  if (gl != 0) return;
  if (bad) atomicCAS(info, 0, nd.off + j + 1);
  orig_all[nd.off + j] = orig;
  tau_all[nd.off + j] = tau;
  d_all[nd.off + j] = dl[orig] + tau;
}

__device__ inline double delta_of(const double* dl, int i, int orig, double tau) { return (dl[i] - dl[orig]) - tau; }

// Gu-Eisenstat: w_i = sign(z_i) sqrt(-delta(i, i) prod_{j != i} delta(i, j) / (d_i - d_j)), same product order as
// dlaed3 (j ascending).
template <int G>
__device__ inline double group_prod(double v) {
  if (G >= 2) v *= dpp_mov<DPP_XOR1>(v);
  if (G >= 4) v *= dpp_mov<DPP_XOR2>(v);
  if (G >= 8) v *= dpp_mov<DPP_HALF_MIRROR>(v);
  if (G >= 16) v *= dpp_mov<DPP_MIRROR>(v);
  static_assert(G <= 16, "group_prod: groups within a row of 16 lanes");
  return v;
}
constexpr int GU_G = 16;  // lanes per z_i: each multiplies every GU_G-th factor

__global__ void gu_kernel(const Node* nodes, const NodeOut* out, const double* __restrict__ dlamda_all,
                          const double* __restrict__ w_all, const int* __restrict__ orig_all,
                          const double* __restrict__ tau_all, double* __restrict__ wnew_all, Zb zb) {
  out = zb.at(out);
  dlamda_all = zb.at(dlamda_all);
  w_all = zb.at(w_all);
  orig_all = zb.at(orig_all);
  tau_all = zb.at(tau_all);
  wnew_all = zb.at(wnew_all);
  const Node nd = nodes[blockIdx.y];
  const int k = out[blockIdx.y].k;
  const int i = (blockIdx.x * blockDim.x + threadIdx.x) / GU_G, gl = threadIdx.x % GU_G;
  if (k <= 2 || i >= k) return;  // whole groups leave together
  const double* dl = dlamda_all + nd.off;
  const int* og = orig_all + nd.off;
  const double* ta = tau_all + nd.off;
  double prod = (gl == 0) ? delta_of(dl, i, og[i], ta[i]) : 1.0;
  for (int j = gl; j < k; j += GU_G) {
    if (j == i) continue;
    prod *= delta_of(dl, i, og[j], ta[j]) / (dl[i] - dl[j]);
  }
  prod = group_prod<GU_G>(prod);
  if (gl != 0) return;
  const double mag = sqrt(-prod);
  wnew_all[nd.off + i] = (w_all[nd.off + i] < 0) ? -mag : mag;
}

// Column j of S: s_i = w_i / delta(i, j), normalized, rows permuted by indxc. One wavefront per column.
__global__ __launch_bounds__(256) void vectors_kernel(const Node* nodes, const NodeOut* out,
                                                      const double* __restrict__ dlamda_all,
                                                      const int* __restrict__ orig_all,
                                                      const double* __restrict__ tau_all,
                                                      const double* __restrict__ wnew_all,
                                                      const int* __restrict__ indxc_all, double* __restrict__ s_all,
                                                      Zb zb) {
  out = zb.at(out);
  dlamda_all = zb.at(dlamda_all);
  orig_all = zb.at(orig_all);
  tau_all = zb.at(tau_all);
  wnew_all = zb.at(wnew_all);
  indxc_all = zb.at(indxc_all);
  s_all = zb.at(s_all);
  const Node nd = nodes[blockIdx.y];
  const int k = out[blockIdx.y].k;
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (k <= 2 || j >= k) return;
  const double* dl = dlamda_all + nd.off;
  const double* wn = wnew_all + nd.off;
  const int* ix = indxc_all + nd.off;
  const int oj = orig_all[nd.off + j];
  const double tj = tau_all[nd.off + j];
  double sum = 0;
  for (int i = lane; i < k; i += 64) {
    const double s = wn[i] / delta_of(dl, i, oj, tj);
    sum += s * s;
  }
  sum = wave_sum(sum);
  const double scale = 1 / sqrt(sum);
  double* S = s_all + nd.q2off;
  for (int i = lane; i < k; i += 64) {
    const int src = ix[i] - 1;
    S[idx(i, j, nd.sz)] = wn[src] / delta_of(dl, src, oj, tj) * scale;
  }
}

// Merge products with device-side sizes: blockIdx.z = 2 * merge + half.
__global__ __launch_bounds__(256) void merge_gemm_kernel(const Node* nodes, const NodeOut* out,
                                                         const double* __restrict__ q2_all,
                                                         const double* __restrict__ s_all, double* __restrict__ q_all,
                                                         int ldq, Zb zb) {
  out = zb.at(out);
  q2_all = zb.at(q2_all);
  s_all = zb.at(s_all);
  q_all = zb.at(q_all);
  using namespace gemm_detail;
  __shared__ __attribute__((aligned(16))) double lds[tile_lds_elems<double>()];
  const int node = zb.z() >> 1, half = zb.z() & 1;
  const Node nd = nodes[node];
  const NodeOut o = out[node];
  const int k = o.k;
  if (k == 0) return;
  const int n1 = nd.n1, n2 = nd.sz - nd.n1;
  const int n12 = o.ctot[0] + o.ctot[1], n23 = o.ctot[1] + o.ctot[2];
  const int rows = half ? n2 : n1, inner = half ? n23 : n12;
  const int m0 = blockIdx.x * BM, c0 = blockIdx.y * BN;
  if (m0 >= rows || c0 >= k) return;
  const double* A = q2_all + nd.q2off + (half ? (size_t)n1 * n12 : 0);
  const double* B = s_all + nd.q2off + (half ? o.ctot[0] : 0);
  double* C = q_all + idx(nd.off + (half ? n1 : 0), nd.off, ldq);
  gemm_tile<double, false, false>(rows, k, inner, 1.0, A, rows, B, nd.sz, 0.0, C, ldq, m0, c0, lds);
}

// The same products for a level whose merges are all at most SMALL_MERGE wide: direct dot products in registers
// (lane = row, one wavefront per column, the sum over the inner index in order), no LDS, so that these small
// workgroups find room beside other kernels' LDS-heavy ones (in a batch, the bulge back-transformation's strips run
// beside the divide and conquer). Grid: (cdiv(maxsz, 4), 1, 2 * merges).
constexpr int SMALL_MERGE = 128;
__global__ __launch_bounds__(256) void merge_small_kernel(const Node* nodes, const NodeOut* out,
                                                          const double* __restrict__ q2_all,
                                                          const double* __restrict__ s_all,
                                                          double* __restrict__ q_all, int ldq, Zb zb) {
  out = zb.at(out);
  q2_all = zb.at(q2_all);
  s_all = zb.at(s_all);
  q_all = zb.at(q_all);
  const int node = zb.z() >> 1, half = zb.z() & 1;
  const Node nd = nodes[node];
  const NodeOut o = out[node];
  const int k = o.k;
  const int n1 = nd.n1, n2 = nd.sz - nd.n1;
  const int n12 = o.ctot[0] + o.ctot[1], n23 = o.ctot[1] + o.ctot[2];
  const int rows = half ? n2 : n1, inner = half ? n23 : n12;
  const double* A = q2_all + nd.q2off + (half ? (size_t)n1 * n12 : 0);
  const double* B = s_all + nd.q2off + (half ? o.ctot[0] : 0);
  double* C = q_all + idx(nd.off + (half ? n1 : 0), nd.off, ldq);
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= k) return;
  const double* Bj = B + (size_t)j * nd.sz;
  for (int i = threadIdx.x & 63; i < rows; i += 64) {
    double acc = 0;
#pragma unroll 8
    for (int l = 0; l < inner; ++l) acc += A[i + (size_t)l * rows] * Bj[l];
    C[idx(i, j, ldq)] = acc;
  }
}

// dlamrg of the K merged (ascending) and n - K deflated (descending) eigenvalues.
// dlamrg(k, sz - k, d, 1, -1) per merge (one workgroup each): the K new eigenvalues ascending, the deflated ones
// stored descending (read backwards), merged by ranks; on ties the first list's element comes first, as dlamrg.
// K = 0 (every column deflated at once) leaves them ascending: the identity.
// The deflated values are not always in exact order: a Givens deflation of close poles recombines two values, and
// LAPACK's dlaed2 itself can leave neighbours a few ulps out of order (with a secular root between them, as
// degenerate spectra produce, ranks would collide). Then thread 0 runs dlamrg's sequential merge on an LDS copy,
// which yields a permutation whatever the order. Dynamic LDS: sz doubles.
__global__ __launch_bounds__(256) void lamrg_kernel(int count, const Node* nodes, const NodeOut* out,
                                                    const double* __restrict__ d_all, int* __restrict__ perm_all,
                                                    Zb zb) {
  out = zb.at(out);
  d_all = zb.at(d_all);
  perm_all = zb.at(perm_all);
  extern __shared__ double s_a[];
  __shared__ int unsorted;
  const Node nd = nodes[blockIdx.x];
  const int k = out[blockIdx.x].k, sz = nd.sz, n2 = sz - k;
  const double* a = d_all + nd.off;
  int* perm = perm_all + nd.off;
  (void)count;
  if (k == 0) {
    // every column deflated at once: they are already ascending
    for (int i = threadIdx.x; i < sz; i += blockDim.x) perm[i] = i + 1;
    return;
  }
  if (threadIdx.x == 0) unsorted = 0;
  __syncthreads();
  for (int i = threadIdx.x; i + 1 < sz; i += blockDim.x) {
    const bool ok = (i + 1 < k) ? (a[i] <= a[i + 1]) : (i < k || a[i] >= a[i + 1]);
    if (!ok) atomicOr(&unsorted, 1);
  }
  __syncthreads();
  if (unsorted) {
    for (int i = threadIdx.x; i < sz; i += blockDim.x) s_a[i] = a[i];
    __syncthreads();
    if (threadIdx.x == 0) {
      int i1 = 0, i2 = sz - 1, o = 0;
      while (i1 < k && i2 >= k) perm[o++] = (s_a[i1] <= s_a[i2]) ? ++i1 : 1 + i2--;
      while (i1 < k) perm[o++] = ++i1;
      while (i2 >= k) perm[o++] = 1 + i2--;
    }
    return;
  }
  for (int i = threadIdx.x; i < sz; i += blockDim.x) {
    if (i < k) {
      const double v = a[i];
      int lo = 0, hi = n2;  // # second-list elements (b_q = a[sz - 1 - q]) < v
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[sz - 1 - mid] < v) lo = mid + 1;
        else hi = mid;
      }
      perm[i + lo] = i + 1;
    } else {
      const int q = sz - 1 - i;  // traversal position in the second list
      const double v = a[i];
      int lo = 0, hi = k;  // # first-list elements <= v
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] <= v) lo = mid + 1;
        else hi = mid;
      }
      perm[q + lo] = i + 1;
    }
  }
}

// temp(:, i) = column perm[i] of the merged system: the merge products (Q(:, 0 ..< K)) or a deflated eigenvector
// (Q2's last group); dtemp(i) = d(perm[i]). A source outside the merge (a permutation that is not one) reads
// nothing and sets info.
__global__ void gather_kernel(const Node* nodes, const NodeOut* out, const int* __restrict__ perm_all,
                              const double* __restrict__ q_all, int ldq, const double* __restrict__ q2_all,
                              const double* __restrict__ d_all, double* __restrict__ tmp_all,
                              double* __restrict__ dtmp_all, int* info, Zb zb) {
  info = zb.at(info);
  out = zb.at(out);
  perm_all = zb.at(perm_all);
  q_all = zb.at(q_all);
  q2_all = zb.at(q2_all);
  d_all = zb.at(d_all);
  tmp_all = zb.at(tmp_all);
  dtmp_all = zb.at(dtmp_all);
  const Node nd = nodes[zb.z()];
  const int col = blockIdx.y;
  if (col >= nd.sz) return;
  const NodeOut o = out[zb.z()];
  const int n = nd.sz, n1 = nd.n1, n2 = n - n1, k = o.k;
  const int src = perm_all[nd.off + col] - 1;
  if (src < 0 || src >= n) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicCAS(info, 0, nd.off + col + 1);
    return;
  }
  const double* from;
  if (src < k) {
    from = q_all + idx(nd.off, nd.off + src, ldq);
  } else {
    const int n12 = o.ctot[0] + o.ctot[1], n23 = o.ctot[1] + o.ctot[2];
    from = q2_all + nd.q2off + (size_t)n1 * n12 + (size_t)n2 * n23 + (size_t)(src - k) * n;
  }
  double* tmp = tmp_all + nd.q2off;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    tmp[idx(i, col, n)] = from[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) dtmp_all[nd.off + col] = d_all[nd.off + src];
}

__global__ void scatter_kernel(const Node* nodes, const double* __restrict__ tmp_all,
                               const double* __restrict__ dtmp_all, double* __restrict__ q_all, int ldq,
                               double* __restrict__ d_all, Zb zb) {
  tmp_all = zb.at(tmp_all);
  dtmp_all = zb.at(dtmp_all);
  q_all = zb.at(q_all);
  d_all = zb.at(d_all);
  const Node nd = nodes[zb.z()];
  const int col = blockIdx.y;
  if (col >= nd.sz) return;
  const double* tmp = tmp_all + nd.q2off;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nd.sz; i += gridDim.x * blockDim.x)
    q_all[idx(nd.off + i, nd.off + col, ldq)] = tmp[idx(i, col, nd.sz)];
  if (blockIdx.x == 0 && threadIdx.x == 0) d_all[nd.off + col] = dtmp_all[nd.off + col];
}

void build(std::vector<Node>& pre, std::vector<std::vector<int>>& levels, int off, int sz, int depth) {
  if (sz <= 1) return;
  const int n1 = sz / 2;
  const int id = (int)pre.size();
  pre.push_back(Node{off, sz, n1, 0});
  if ((int)levels.size() <= depth) levels.resize(depth + 1);
  levels[depth].push_back(id);
  build(pre, levels, off, n1, depth + 1);
  build(pre, levels, off + n1, sz - n1, depth + 1);
}

}  // namespace

DcPlan::DcPlan(int n_) : n(n_) {
  if (n <= 1) return;
  std::vector<Node> pre;
  std::vector<std::vector<int>> levels;
  build(pre, levels, 0, n, 0);
  nmerge = (int)pre.size();
  // device node array: [pre-order (for the tears)] then every level's merges with their Q2 / S offsets
  std::vector<Node> nodes(pre);
  std::vector<int> pre_index;
  for (size_t L = 0; L < levels.size(); ++L) {
    Level lv;
    lv.start = (int)nodes.size();
    lv.count = (int)levels[L].size();
    lv.pre_start = (int)pre_index.size();
    int q2off = 0;
    for (int id : levels[L]) {
      Node nd = pre[id];
      nd.q2off = q2off;
      q2off += nd.sz * nd.sz;
      lv.maxsz = std::max(lv.maxsz, nd.sz);
      nodes.push_back(nd);
      pre_index.push_back(id);
    }
    lvls.push_back(lv);
  }
  std::vector<int> split_pre(n + 1, -1);
  for (size_t p = 0; p < pre.size(); ++p) split_pre[pre[p].off + pre[p].n1] = (int)p;
  XTB_HIP(hipMalloc(&split_pre_dev, sizeof(int) * split_pre.size()));
  XTB_HIP(hipMemcpy(split_pre_dev, split_pre.data(), sizeof(int) * split_pre.size(), hipMemcpyHostToDevice));
  XTB_HIP(hipMalloc(&nodes_dev, sizeof(Node) * nodes.size()));
  XTB_HIP(hipMalloc(&pre_index_dev, sizeof(int) * pre_index.size()));
  XTB_HIP(hipMemcpy(nodes_dev, nodes.data(), sizeof(Node) * nodes.size(), hipMemcpyHostToDevice));
  XTB_HIP(hipMemcpy(pre_index_dev, pre_index.data(), sizeof(int) * pre_index.size(), hipMemcpyHostToDevice));
}

DcPlan::~DcPlan() {
  if (nodes_dev) (void)hipFree(nodes_dev);
  if (pre_index_dev) (void)hipFree(pre_index_dev);
  if (split_pre_dev) (void)hipFree(split_pre_dev);
}

// Workspace (doubles): dlamda, w, tau, wnew (n each), dtmp (2n), rho_pre, rho_lvl (n each), scale (16), rotation
// cosines/sines (2n), Q2 (n^2), S (n^2). Ints: orig, indxc, indx, perm (n each), rotation columns (2n), out (8n),
// probe scratch (16).
size_t dc_workspace_doubles(int n) { return 10 * (size_t)n + 16 + 2 * (size_t)n * n; }
size_t dc_workspace_ints(int n) { return 14 * (size_t)n + 32; }

namespace {

struct DcWork {
  double *dlamda, *w, *tau, *wnew, *dtmp, *rho_pre, *rho_lvl, *scale, *rotcs, *q2, *sbuf;
  int *orig, *indxc, *indx, *perm, *rotp;
  NodeOut* outs;
  int* scratch;
};

DcWork carve(int n, double* work, int* iwork) {
  DcWork w;
  w.dlamda = work;
  w.w = w.dlamda + n;
  w.tau = w.w + n;
  w.wnew = w.tau + n;
  w.dtmp = w.wnew + n;
  w.rho_pre = w.dtmp + 2 * (size_t)n;
  w.rho_lvl = w.rho_pre + n;
  w.scale = w.rho_lvl + n;
  w.rotcs = w.scale + 16;
  w.q2 = w.rotcs + 2 * (size_t)n;
  w.sbuf = w.q2 + (size_t)n * n;
  w.orig = iwork;
  w.indxc = w.orig + n;
  w.indx = w.indxc + n;
  w.perm = w.indx + n;
  w.rotp = w.perm + n;
  w.outs = reinterpret_cast<NodeOut*>(w.rotp + 2 * (size_t)n);
  w.scratch = reinterpret_cast<int*>(w.outs + n);
  return w;
}

// Every kernel of one tree level (cnt merges, the largest of size maxsz). probe: optional snapshot of dlaed2's
// outputs right after the deflation.
struct Probe {
  double *d_defl, *q2, *dlamda, *w, *rho;
  int *kctot, *indxc;
};

void run_level(hipStream_t s, int n, const Node* lvl, const int* lvl_pre, int cnt, int maxsz, double* d, double* z,
               int ldz, const DcWork& W, int* info, const Probe* probe, int maxit) {
  const size_t lds = (size_t)maxsz * (6 * sizeof(double) + 7 * sizeof(int)) + 64;
  if (lds > 160 * 1024) throw std::runtime_error("stedc: n too large for the LDS-staged deflation");
  allow_lds(deflate_kernel, lds);
  hipLaunchKernelGGL(deflate_kernel, dim3(cnt, 1, zgrid()), dim3(256), lds, s, lvl, lvl_pre, d, z, ldz, W.rho_pre,
                     W.rho_lvl, W.dlamda, W.w, W.indx, W.indxc, W.rotp, W.rotcs, W.outs, zb());
  XTB_LAUNCH_CHECK();
  hipLaunchKernelGGL(rotate_kernel, dim3(cdiv(maxsz, 256), cnt, zgrid()), dim3(256), 0, s, lvl, W.outs, W.rotp,
                     W.rotcs, z, ldz, zb());
  XTB_LAUNCH_CHECK();
  hipLaunchKernelGGL(q2_gather_kernel, dim3(maxsz, cnt, zgrid()), dim3(256), 0, s, lvl, W.outs, W.indx, z, ldz, W.q2,
                     info, zb());
  XTB_LAUNCH_CHECK();
  if (probe) {
    XTB_HIP(hipMemcpyAsync(probe->d_defl, d, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    XTB_HIP(hipMemcpyAsync(probe->q2, W.q2, sizeof(double) * n * n, hipMemcpyDeviceToDevice, s));
    XTB_HIP(hipMemcpyAsync(probe->kctot, W.outs, sizeof(int) * 5, hipMemcpyDeviceToDevice, s));
    XTB_HIP(hipMemcpyAsync(probe->dlamda, W.dlamda, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    XTB_HIP(hipMemcpyAsync(probe->w, W.w, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    XTB_HIP(hipMemcpyAsync(probe->indxc, W.indxc, sizeof(int) * n, hipMemcpyDeviceToDevice, s));
    XTB_HIP(hipMemcpyAsync(probe->rho, W.rho_lvl, sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  allow_lds(secular_kernel, 2 * sizeof(double) * maxsz);
  hipLaunchKernelGGL(secular_kernel, dim3(cdiv(maxsz, 64 / SEC_G), cnt, zgrid()), dim3(64), 2 * sizeof(double) * maxsz,
                     s, lvl, W.outs, W.rho_lvl, W.dlamda, W.w, W.indxc, d, W.orig, W.tau, W.sbuf, info, maxit, zb());
  XTB_LAUNCH_CHECK();
  hipLaunchKernelGGL(gu_kernel, dim3(cdiv(maxsz * GU_G, 64), cnt, zgrid()), dim3(64), 0, s, lvl, W.outs, W.dlamda, W.w,
                     W.orig, W.tau, W.wnew, zb());
  XTB_LAUNCH_CHECK();
  hipLaunchKernelGGL(vectors_kernel, dim3(cdiv(maxsz, 4), cnt, zgrid()), dim3(256), 0, s, lvl, W.outs, W.dlamda,
                     W.orig, W.tau, W.wnew, W.indxc, W.sbuf, zb());
  XTB_LAUNCH_CHECK();
  if (maxsz <= SMALL_MERGE)
    hipLaunchKernelGGL(merge_small_kernel, dim3(cdiv(maxsz, 4), 1, zgrid(2 * cnt)), dim3(256), 0, s, lvl, W.outs, W.q2, W.sbuf, z,
                       ldz, zb(2 * cnt));
  else
    hipLaunchKernelGGL(merge_gemm_kernel, dim3(cdiv(maxsz, 64), cdiv(maxsz, 64), zgrid(2 * cnt)), dim3(256), 0, s,
                       lvl, W.outs, W.q2, W.sbuf, z, ldz, zb(2 * cnt));
  XTB_LAUNCH_CHECK();
  hipLaunchKernelGGL(lamrg_kernel, dim3(cnt, 1, zgrid()), dim3(256), sizeof(double) * maxsz, s, cnt, lvl, W.outs, d,
                     W.perm, zb());
  XTB_LAUNCH_CHECK();
  hipLaunchKernelGGL(gather_kernel, dim3(cdiv(maxsz, 256), maxsz, zgrid(cnt)), dim3(256), 0, s, lvl, W.outs, W.perm,
                     z, ldz, W.q2, d, W.sbuf, W.dtmp, info, zb(cnt));
  XTB_LAUNCH_CHECK();
  hipLaunchKernelGGL(scatter_kernel, dim3(cdiv(maxsz, 256), maxsz, zgrid(cnt)), dim3(256), 0, s, lvl, W.sbuf, W.dtmp,
                     z, ldz, d, zb(cnt));
  XTB_LAUNCH_CHECK();
}

}  // namespace

void stedc(hipStream_t s, const DcPlan& plan, double* d, double* e, double* z, int ldz, double* work, int* iwork,
           int* info, int secular_maxit) {
  const int maxit = secular_maxit >= 0 ? std::min(secular_maxit, secular::MAXIT) : secular::MAXIT;
  const int n = plan.n;
  if (n <= 0) return;
  if (n == 1) {
    set_identity(s, 1, z, ldz);
    return;
  }
  const DcWork W = carve(n, work, iwork);
  const Node* nodes = plan.nodes_dev;
  set_identity(s, n, z, ldz);
  hipLaunchKernelGGL(prepare_kernel, dim3(1, 1, zgrid()), dim3(1024), 0, s, n, d, e, plan.split_pre_dev, W.rho_pre,
                     W.scale, zb());
  XTB_LAUNCH_CHECK();
  for (int L = (int)plan.lvls.size() - 1; L >= 0; --L) {
    const DcPlan::Level& lv = plan.lvls[L];
    run_level(s, n, nodes + lv.start, plan.pre_index_dev + lv.pre_start, lv.count, lv.maxsz, d, z, ldz, W, info,
              nullptr, maxit);
  }
  hipLaunchKernelGGL(unscale_kernel, dim3(cdiv(n, 256), 1, zgrid()), dim3(256), 0, s, n, d, W.scale, zb());
  XTB_LAUNCH_CHECK();
}

// One merge (dlaed1 semantics) of two solved halves on the device, exposing dlaed2's outputs for comparison with the
// host LAPACK: d (n) ascending within each half, q (n x n, ldq) block diagonal with the halves' eigenvectors.
// kctot_out: K, ctot[0..3]; d_defl_out: d after deflation (the deflated eigenvalues in K ..< n); q2_out: Q2 (n^2).
void dc_merge_probe(hipStream_t s, int n, int n1, double rho, double* d, double* q, int ldq, double* work,
                    int* iwork, int* info, double* d_defl_out, double* q2_out, int* kctot_out, double* dlamda_out,
                    double* w_out, int* indxc_out, double* rho_out) {
  if (n < 2 || n1 < 1 || n1 >= n) throw std::runtime_error("dc_merge_probe: need 1 <= n1 < n");
  const DcWork W = carve(n, work, iwork);
  int* pre_index = W.scratch;
  Node* node = reinterpret_cast<Node*>(W.scratch + 4);
  const Node host_node{0, n, n1, 0};
  const int zero = 0;
  XTB_HIP(hipMemcpy(node, &host_node, sizeof(Node), hipMemcpyHostToDevice));
  XTB_HIP(hipMemcpy(pre_index, &zero, sizeof(int), hipMemcpyHostToDevice));
  XTB_HIP(hipMemcpy(W.rho_pre, &rho, sizeof(double), hipMemcpyHostToDevice));
  const Probe pr{d_defl_out, q2_out, dlamda_out, w_out, rho_out, kctot_out, indxc_out};
  run_level(s, n, node, pre_index, 1, n, d, q, ldq, W, info, &pr, secular::MAXIT);
}

}  // namespace xtbeig
