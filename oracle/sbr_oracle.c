/*
 * ORACLE — test infrastructure only. CPU restatement of the reference's eigensolver stages, used by the tests (and
 * the benchmark's CPU-baseline leg) as the checker. Nothing in the product links or calls this file.
 *
 * Restated from the reference (philipturner/xtb-diagonalization, Swift sources; sytrd.c itself is absent from the
 * reference checkout):
 *   orc_reduce_to_band      Sources/Diagonalization/Stages/BandForm.swift:28-142 (combinedUpdate, full storage,
 *                           gemm variant), geqrf + larft restated as Householder QR (LAPACK dgeqr2/dlarfg) and the
 *                           forward columnwise T recurrence (dlarft)
 *   orc_chase_bulges        Stages/BulgeChasing.swift:188-321 (chaseBulgeLower, sequential sweeps, scalar loops)
 *   orc_back_bulges         Stages/BulgeBackTransformation.swift:11-30 (one reflector at a time)
 *   orc_back_band           Stages/BandFormBackTransformation.swift:10-21 with WYTransform.swift:39-67 (gemm form)
 *   orc_sytrd_upper         Tridiagonalization.swift:510-618 (LAPACK dsytrd 'U', nb = 32, nx = 32: dlatrd panels,
 *                           syr2k, dsytd2 tail), scalar loops, no symmetric-kernel fusion
 *   orc_ormtr_upper         TridiagonalBackTransformation.swift:16-45 (32-reflector blocks, larftBackward, 3 gemm)
 * Indices 0-based, matrices column-major.
 *
 * Pinning (tests/test_oracle.py): every function against the host LAPACK (dsytrd / dormtr / dsyevd / dstedc through
 * scipy's OpenBLAS) and against invariants: spectra preserved, exact zeros outside the band, orthogonality of the
 * accumulated transformations, brute-force products on tiny n.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define A_(i, j) a[(size_t)(i) + (size_t)(j) * lda]

static double sgn(double a, double b) { return b >= 0 ? fabs(a) : -fabs(a); }

/* dlarfg on (alpha, x[0..n-2]) without the tiny-beta rescaling loop. Returns tau; alpha becomes beta. */
static double larfg(int n, double *alpha, double *x, int incx) {
  if (n <= 1) return 0.0;
  double xn = 0;
  for (int i = 0; i < n - 1; ++i) xn += x[(size_t)i * incx] * x[(size_t)i * incx];
  if (xn == 0) return 0.0;
  double beta = -sgn(sqrt(*alpha * *alpha + xn), *alpha);
  double tau = (beta - *alpha) / beta;
  double scal = 1.0 / (*alpha - beta);
  for (int i = 0; i < n - 1; ++i) x[(size_t)i * incx] *= scal;
  *alpha = beta;
  return tau;
}

/* C (m x n, ldc) = alpha op(A) op(B) + beta C */
static void gemm(char ta, char tb, int m, int n, int k, double alpha, const double *A, int lda, const double *B,
                 int ldb, double beta, double *C, int ldc) {
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      double s = 0;
      for (int l = 0; l < k; ++l) {
        double av = ta == 'N' ? A[i + (size_t)l * lda] : A[l + (size_t)i * lda];
        double bv = tb == 'N' ? B[l + (size_t)j * ldb] : B[j + (size_t)l * ldb];
        s += av * bv;
      }
      C[i + (size_t)j * ldc] = alpha * s + (beta == 0 ? 0 : beta * C[i + (size_t)j * ldc]);
    }
}

/* dgeqr2 on p (m x nb, ld), tau (min(m, nb)); V in geqrf layout below the diagonal. */
static void geqr2(int m, int nb, double *p, int ld, double *tau) {
  int k = m < nb ? m : nb;
  for (int c = 0; c < k; ++c) {
    double *col = p + c + (size_t)c * ld;
    double t = larfg(m - c, col, col + 1, 1);
    tau[c] = t;
    if (t != 0) {
      double keep = col[0];
      col[0] = 1;
      for (int cc = c + 1; cc < nb; ++cc) {
        double *o = p + c + (size_t)cc * ld;
        double s = 0;
        for (int i = 0; i < m - c; ++i) s += col[i] * o[i];
        for (int i = 0; i < m - c; ++i) o[i] -= t * s * col[i];
      }
      col[0] = keep;
    }
  }
}

/* dlarft('F', 'C'): T (k x k, ldt) of H_1 ... H_k = I - V T V^T, V unit lower trapezoidal (explicit, ld ldv). */
static void larft_forward(int m, int k, const double *v, int ldv, const double *tau, double *t, int ldt) {
  for (int i = 0; i < k; ++i) {
    for (int r = 0; r < k; ++r) t[r + (size_t)i * ldt] = 0;
    if (tau[i] == 0) continue;
    for (int r = 0; r < i; ++r) {
      double s = 0;
      for (int q = 0; q < m; ++q) s += v[q + (size_t)r * ldv] * v[q + (size_t)i * ldv];
      t[r + (size_t)i * ldt] = -tau[i] * s;
    }
    for (int r = 0; r < i; ++r) {
      double s = 0;
      for (int q = r; q < i; ++q) s += t[r + (size_t)q * ldt] * t[q + (size_t)i * ldt];
      t[r + (size_t)i * ldt] = s;
    }
    t[i + (size_t)i * ldt] = tau[i];
  }
}

/* Stage 1. a: full symmetric n x n (lda), updated in both triangles of the trailing block; on exit the lower band
 * (diagonals 0 .. b) holds the band matrix and everything below it is zero. v1 (n x n, ld n): per panel j the
 * explicit V (rows j + b .., columns j ..). t1 (b x n, ld b): per panel the T factor. */
void orc_reduce_to_band(int n, int b, double *a, int lda, double *v1, double *t1) {
  double *tau = malloc(sizeof(double) * (b + 1));
  double *vf = malloc(sizeof(double) * (size_t)n * b + 1);
  double *vt = malloc(sizeof(double) * (size_t)n * b + 1);
  double *x = malloc(sizeof(double) * (size_t)n * b + 1);
  double *y = malloc(sizeof(double) * (size_t)b * b + 1);
  double *pan = malloc(sizeof(double) * (size_t)n * b + 1);
  for (int j = 0; j < n - b - 1; j += b) {
    int r = j + b, m = n - r, k = m < b ? m : b;
    for (int c = 0; c < b; ++c)
      for (int i = 0; i < m; ++i) pan[i + (size_t)c * m] = A_(r + i, j + c);
    geqr2(m, b, pan, m, tau);
    for (int c = 0; c < b; ++c)
      for (int i = 0; i < m; ++i) {
        double val = i <= c ? pan[i + (size_t)c * m] : 0;
        A_(r + i, j + c) = val;
        A_(j + c, r + i) = val;
      }
    for (int c = 0; c < k; ++c)
      for (int i = 0; i < m; ++i) {
        double val = i < c ? 0 : (i == c ? 1 : pan[i + (size_t)c * m]);
        vf[i + (size_t)c * m] = val;
        v1[(r + i) + (size_t)(j + c) * n] = val;
      }
    double *t = t1 + (size_t)j * b;
    larft_forward(m, k, vf, m, tau, t, b);
    /* X = A2 (V T), W = X - V (T^T V^T A2 V T) / 2, A2 -= V W^T + W V^T */
    double *a2 = a + r + (size_t)r * lda;
    gemm('N', 'N', m, k, k, 1, vf, m, t, b, 0, vt, m);
    gemm('N', 'N', m, k, m, 1, a2, lda, vt, m, 0, x, m);
    gemm('T', 'N', k, k, m, 1, vt, m, x, m, 0, y, k);
    gemm('N', 'N', m, k, k, -0.5, vf, m, y, k, 1, x, m);
    for (int jj = 0; jj < m; ++jj)
      for (int ii = 0; ii < m; ++ii) {
        double s = 0;
        for (int l = 0; l < k; ++l)
          s += vf[ii + (size_t)l * m] * x[jj + (size_t)l * m] + x[ii + (size_t)l * m] * vf[jj + (size_t)l * m];
        a2[ii + (size_t)jj * lda] -= s;
      }
  }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i)
      if (abs(i - j) > b) A_(i, j) = 0;
  free(tau); free(vf); free(vt); free(x); free(y); free(pan);
}

/* One bulge (s, k) of the lower-triangle chase on a (full storage, lower triangle referenced). */
static int chase_bulge(int n, int b, double *a, int lda, double *v2, int s, int k, double *w, double *y, double *p) {
  int rs = s + 1 + k * b;
  if (rs >= n) return 0;
  int re = rs + b < n ? rs + b : n, len = re - rs;
  int c = k == 0 ? s : rs - b;
  double x0 = A_(rs, c);
  double xs[128];
  for (int i = 1; i < len; ++i) xs[i - 1] = A_(rs + i, c);
  double alpha = x0;
  double tau = larfg(len, &alpha, xs, 1);
  double beta = alpha;
  double sc = sqrt(tau);
  w[0] = sc;
  for (int i = 1; i < len; ++i) w[i] = sc * xs[i - 1];
  for (int i = 0; i < len; ++i) v2[(rs + i) + (size_t)s * n] = w[i];
  if (tau != 0) {
    int lo = rs - b > 0 ? rs - b : 0, hi = re + b < n ? re + b : n;
    for (int j = lo; j < rs; ++j) {
      double dot = 0;
      for (int i = 0; i < len; ++i) dot += w[i] * A_(rs + i, j);
      for (int i = 0; i < len; ++i) A_(rs + i, j) -= w[i] * dot;
    }
    int below = hi - re;
    for (int i = 0; i < below; ++i) y[i] = 0;
    for (int jj = 0; jj < len; ++jj)
      for (int i = 0; i < below; ++i) y[i] += A_(re + i, rs + jj) * w[jj];
    for (int jj = 0; jj < len; ++jj)
      for (int i = 0; i < below; ++i) A_(re + i, rs + jj) -= y[i] * w[jj];
    for (int i = 0; i < len; ++i) p[i] = 0;
    for (int jj = 0; jj < len; ++jj) {
      double acc = A_(rs + jj, rs + jj) * w[jj];
      for (int i = jj + 1; i < len; ++i) {
        p[i] += A_(rs + i, rs + jj) * w[jj];
        acc += A_(rs + i, rs + jj) * w[i];
      }
      p[jj] += acc;
    }
    double half = 0;
    for (int i = 0; i < len; ++i) half += w[i] * p[i];
    half /= 2;
    for (int i = 0; i < len; ++i) p[i] -= half * w[i];
    for (int jj = 0; jj < len; ++jj)
      for (int i = jj; i < len; ++i) A_(rs + i, rs + jj) -= w[i] * p[jj] + p[i] * w[jj];
  }
  A_(rs, c) = beta;
  for (int i = 1; i < len; ++i) A_(rs + i, c) = 0;
  return 1;
}

/* Stage 2 on a (n x n, lda; lower triangle: the band of width b, zero below). d (n), e (n - 1); v2 (n x n, ld n). */
void orc_chase_bulges(int n, int b, double *a, int lda, double *v2, double *d, double *e) {
  double w[128], y[128], p[128];
  if (b > 1 && b <= 64)
    for (int s = 0; s < n - 2; ++s)
      for (int k = 0; chase_bulge(n, b, a, lda, v2, s, k, w, y, p); ++k) {}
  for (int i = 0; i < n; ++i) d[i] = A_(i, i);
  for (int i = 0; i + 1 < n; ++i) e[i] = A_(i + 1, i);
}

/* z = H_1 ... H_N z, reverse chase order, one reflector at a time. */
void orc_back_bulges(int n, int b, const double *v2, double *z, int ldz, int m) {
  if (b <= 1) return;
  for (int s = n - 3; s >= 0; --s) {
    int kmax = 0;
    while (s + 1 + (kmax + 1) * b < n) ++kmax;
    for (int k = kmax; k >= 0; --k) {
      int rs = s + 1 + k * b, re = rs + b < n ? rs + b : n;
      const double *w = v2 + rs + (size_t)s * n;
      for (int col = 0; col < m; ++col) {
        double *zc = z + rs + (size_t)col * ldz;
        double dot = 0;
        for (int i = 0; i < re - rs; ++i) dot += w[i] * zc[i];
        for (int i = 0; i < re - rs; ++i) zc[i] -= w[i] * dot;
      }
    }
  }
}

/* z = Q1 z: per panel, last first, z(r:n, :) -= V (T (V^T z(r:n, :))). */
void orc_back_band(int n, int b, const double *v1, const double *t1, double *z, int ldz, int m) {
  int last = -1;
  for (int j = 0; j < n - b - 1; j += b) last = j;
  double *w1 = malloc(sizeof(double) * (size_t)b * (m > 0 ? m : 1));
  double *w2 = malloc(sizeof(double) * (size_t)b * (m > 0 ? m : 1));
  for (int j = last; j >= 0; j -= b) {
    int r = j + b, rows = n - r, k = rows < b ? rows : b;
    const double *v = v1 + r + (size_t)j * n;
    gemm('T', 'N', k, m, rows, 1, v, n, z + r, ldz, 0, w1, k);
    gemm('N', 'N', k, m, k, 1, t1 + (size_t)j * b, b, w1, k, 0, w2, k);
    gemm('N', 'N', rows, m, k, -1, v, n, w2, k, 1, z + r, ldz);
  }
  free(w1); free(w2);
}

/* ---- 1-stage reduction, LAPACK dsytrd('U') with nb = 32, nx = 32 (Tridiagonalization.swift:510-618) -------- */

#define AT(i, j) a[(size_t)((i) - 1) + (size_t)((j) - 1) * lda] /* 1-based */
#define WT(i, j) w[(size_t)((i) - 1) + (size_t)((j) - 1) * ldw]

/* y = alpha * A(1:n, 1:n) x from the upper triangle */
static void symv_upper(int n, double alpha, const double *a, int lda, const double *x, double *y) {
  for (int i = 0; i < n; ++i) y[i] = 0;
  for (int j = 1; j <= n; ++j) {
    double t1 = alpha * x[j - 1], t2 = 0;
    for (int i = 1; i < j; ++i) {
      y[i - 1] += t1 * AT(i, j);
      t2 += AT(i, j) * x[i - 1];
    }
    y[j - 1] += t1 * AT(j, j) + alpha * t2;
  }
}

/* dlatrd, uplo = 'U' */
static void latrd_upper(int n, int nb, double *a, int lda, double *e, double *tau, double *w, int ldw) {
  for (int i = n; i >= n - nb + 1; --i) {
    int iw = i - n + nb;
    if (i < n) {
      /* A(1:i, i) -= A(1:i, i+1:n) W(i, iw+1:nb)^T + W(1:i, iw+1:nb) A(i, i+1:n)^T */
      for (int r = 1; r <= i; ++r) {
        double s = 0;
        for (int c = 1; c <= n - i; ++c) s += AT(r, i + c) * WT(i, iw + c) + WT(r, iw + c) * AT(i, i + c);
        AT(r, i) -= s;
      }
    }
    if (i > 1) {
      int im1 = i - 1;
      tau[i - 2] = larfg(im1, &AT(i - 1, i), &AT(1, i), 1);
      e[i - 2] = AT(i - 1, i);
      AT(i - 1, i) = 1;
      symv_upper(im1, 1.0, a, lda, &AT(1, i), &WT(1, iw));
      if (i < n) {
        int m = n - i;
        double *u = malloc(sizeof(double) * m), *u2 = malloc(sizeof(double) * m);
        for (int c = 1; c <= m; ++c) {
          double s1 = 0, s2 = 0;
          for (int r = 1; r <= im1; ++r) {
            s1 += WT(r, iw + c) * AT(r, i);
            s2 += AT(r, i + c) * AT(r, i);
          }
          u[c - 1] = s1;
          u2[c - 1] = s2;
        }
        for (int c = 1; c <= m; ++c) WT(i + c, iw) = u2[c - 1];
        for (int r = 1; r <= im1; ++r) {
          double s = 0;
          for (int c = 1; c <= m; ++c) s += AT(r, i + c) * u[c - 1] + WT(r, iw + c) * u2[c - 1];
          WT(r, iw) -= s;
        }
        free(u); free(u2);
      }
      for (int r = 1; r <= im1; ++r) WT(r, iw) *= tau[i - 2];
      double dot = 0;
      for (int r = 1; r <= im1; ++r) dot += WT(r, iw) * AT(r, i);
      double alpha = -0.5 * tau[i - 2] * dot;
      for (int r = 1; r <= im1; ++r) WT(r, iw) += alpha * AT(r, i);
    }
  }
}

static void sytd2_upper(int n, double *a, int lda, double *d, double *e, double *tau) {
  double *t = malloc(sizeof(double) * (n > 0 ? n : 1));
  for (int i = n - 1; i >= 1; --i) {
    double taui = larfg(i, &AT(i, i + 1), &AT(1, i + 1), 1);
    e[i - 1] = AT(i, i + 1);
    if (taui != 0) {
      AT(i, i + 1) = 1;
      symv_upper(i, taui, a, lda, &AT(1, i + 1), t);
      double dot = 0;
      for (int r = 1; r <= i; ++r) dot += t[r - 1] * AT(r, i + 1);
      double alpha = -0.5 * taui * dot;
      for (int r = 1; r <= i; ++r) t[r - 1] += alpha * AT(r, i + 1);
      for (int c = 1; c <= i; ++c)
        for (int r = 1; r <= c; ++r) AT(r, c) -= AT(r, i + 1) * t[c - 1] + t[r - 1] * AT(c, i + 1);
      AT(i, i + 1) = e[i - 1];
    }
    d[i] = AT(i + 1, i + 1);
    tau[i - 1] = taui;
  }
  d[0] = AT(1, 1);
  free(t);
}

void orc_sytrd_upper(int n, double *a, int lda, double *d, double *e, double *tau) {
  const int nb = 32, nx = 32;
  int ldw = n;
  double *w = malloc(sizeof(double) * (size_t)n * nb + 1);
  int kk = n;
  if (n > nx) {
    kk = n - ((n - nx + nb - 1) / nb) * nb;
    for (int i = n - nb + 1; i >= kk + 1; i -= nb) {
      latrd_upper(i + nb - 1, nb, a, lda, e, tau, w, ldw);
      /* syr2k('U', 'N', i - 1, nb, -1, A(1, i), W) */
      for (int c = 1; c <= i - 1; ++c)
        for (int r = 1; r <= c; ++r) {
          double s = 0;
          for (int l = 0; l < nb; ++l)
            s += AT(r, i + l) * w[(c - 1) + (size_t)l * ldw] + w[(r - 1) + (size_t)l * ldw] * AT(c, i + l);
          AT(r, c) -= s;
        }
      for (int j = i; j <= i + nb - 1; ++j) {
        AT(j - 1, j) = e[j - 2];
        d[j - 1] = AT(j, j);
      }
    }
  }
  sytd2_upper(kk, a, lda, d, e, tau);
  free(w);
}

/* z = Q z (dormtr 'L', 'U', 'N') for m columns, Q from orc_sytrd_upper: 32-reflector blocks in increasing order. */
void orc_ormtr_upper(int n, int m, const double *a, int lda, const double *tau, double *z, int ldz) {
  const int nb = 32, k = n - 1;
  if (m <= 0 || k <= 0) return;
  double *v = malloc(sizeof(double) * (size_t)n * nb), *t = malloc(sizeof(double) * nb * nb);
  double *w1 = malloc(sizeof(double) * (size_t)nb * m), *w2 = malloc(sizeof(double) * (size_t)nb * m);
  for (int i = 0; i < k; i += nb) {
    int ib = k - i < nb ? k - i : nb, rows = i + ib;
    for (int c = 0; c < ib; ++c) {
      int r = i + c;
      for (int q = 0; q < r; ++q) v[q + (size_t)c * rows] = a[q + (size_t)(r + 1) * lda];
      v[r + (size_t)c * rows] = 1;
      for (int q = r + 1; q < rows; ++q) v[q + (size_t)c * rows] = 0;
    }
    /* larft('B', 'C'): H(ib) ... H(1) ordering for xormql blocks; Q_block = H_{i+ib-1} ... H_i applied as
       I - V T V^T with T lower triangular */
    for (int c = ib - 1; c >= 0; --c) {
      for (int r2 = 0; r2 < ib; ++r2) t[r2 + (size_t)c * ib] = 0;
      double tc = tau[i + c];
      if (tc == 0) continue;
      for (int r2 = c + 1; r2 < ib; ++r2) {
        double s = 0;
        for (int q = 0; q < rows; ++q) s += v[q + (size_t)r2 * rows] * v[q + (size_t)c * rows];
        t[r2 + (size_t)c * ib] = -tc * s;
      }
      for (int r2 = ib - 1; r2 > c; --r2) {
        double s = 0;
        for (int q = c + 1; q <= r2; ++q) s += t[r2 + (size_t)q * ib] * t[q + (size_t)c * ib];
        t[r2 + (size_t)c * ib] = s;
      }
      t[c + (size_t)c * ib] = tc;
    }
    gemm('T', 'N', ib, m, rows, 1, v, rows, z, ldz, 0, w1, ib);
    gemm('N', 'N', ib, m, ib, 1, t, ib, w1, ib, 0, w2, ib);
    gemm('N', 'N', rows, m, ib, -1, v, rows, w2, ib, 1, z, ldz);
  }
  free(v); free(t); free(w1); free(w2);
}
