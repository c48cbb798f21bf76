// Fermi-smeared density matrix P = C f(eps) C^T of a generalized eigensystem (H C = S C diag(eps), C^T S C = I) and
// its backward with respect to H and S. The reference consumes the eigenpairs this way in its SCC study
// (scripts/kernels/refine.f90:85-132: Fermi occupations with a bisected chemical potential, P = C f C^T); the
// backward has no reference implementation and follows from first-order perturbation of the generalized problem:
//   Gamma = C^T Gbar C (symmetrized),
//   Hbar = C (F o Gamma - kappa diag(f')) C^T,     F_ij = (f_i - f_j) / (eps_i - eps_j), F_ii = f'(eps_i),
//   Sbar = -C (G o Gamma - kappa diag(eps f')) C^T, G_ij = (eps_i f_i - eps_j f_j) / (eps_i - eps_j), G_ii = f_i +
//          eps_i f'_i,
// kappa = sum_j f'_j Gamma_jj / sum_j f'_j when mu is re-bisected with the electron count fixed (0 for a fixed mu).
// With x = (eps - mu) / kT and f = 1 / (1 + e^x): f_i - f_j = -sinh((x_i - x_j) / 2) / (2 cosh(x_i / 2) cosh(x_j / 2)),
// so F_ij = -(sinh(a / 2) / a) e^{-lc_i - lc_j} / (2 kT), a = |x_i - x_j|, lc = log cosh(x / 2): every factor is
// formed to full relative accuracy, for any gap (the degenerate limit a -> 0 gives f' exactly) and for states deep
// in the occupied or empty range. G_ij = (f_i + f_j) / 2 + (eps_i + eps_j) F_ij / 2.
// Spin channels: with nspin = 2 the occupations are those of the reference's open-shell case (refine.f90:85-93),
// f = f_alpha + f_beta, each Fermi-smeared with its own chemical potential (sum f_alpha = nel_alpha, sum f_beta =
// nel_beta); the backward is linear in the occupation function, so F, G and the kappa terms are sums over the
// channels (one kappa per channel, each mu re-bisected with its own count fixed). nspin = 1 is one channel with
// sum f = nel.
// All of it runs on batches of equal-n problems: the products are batched GEMMs on the matrix cores (FP64), the
// elementwise and reduction kernels take the problem index in blockIdx.z / blockIdx.x.
#include "common.hpp"
#include "kernels.hpp"

#include <cmath>

namespace xtbeig {

namespace {

__device__ inline double log_cosh_half(double x) {
  const double ax = fabs(x);
  return 0.5 * ax + log1p(exp(-ax)) - 0.6931471805599453;
}

// F_ij for one pair (x: scaled energies, lc: log cosh(x / 2)).
__device__ inline double fermi_divdiff(double xi, double xj, double lci, double lcj, double kT) {
  const double a = fabs(xi - xj);
  double r;
  if (a < 1e-4) {
    r = 0.5 + a * a / 48.0;  // sinh(a / 2) / a, next term a^4 / 3840 below 1e-17
    return -r * exp(-lci - lcj) / (2.0 * kT);
  }
  if (a < 30.0) {
    r = sinh(0.5 * a) / a;
    return -r * exp(-lci - lcj) / (2.0 * kT);
  }
  // log(sinh(a / 2)) = a / 2 + log1p(-e^{-a}) - log 2
  const double ls = 0.5 * a + log1p(-exp(-a)) - 0.6931471805599453;
  return -exp(ls - log(a) - lci - lcj) / (2.0 * kT);
}

// One workgroup per (problem, spin channel): mu by bisection so that sum_i f_i = the channel's count (fixed-mu
// callers pass the mu they want and bisect = 0), then f, f' and lc per state into occ (3 n per channel: f | f' | lc;
// channel q of problem p at (p nspin + q) 3 n; mu likewise at p nspin + q).
__global__ __launch_bounds__(256) void occupation_kernel(int n, int nspin, const double* __restrict__ w, int64_t sw,
                                                         double kT, double nel_a, double nel_b, int bisect,
                                                         double* __restrict__ mu, double* __restrict__ occ) {
  __shared__ double red[4];
  __shared__ double smn[256], smx[256];
  const int p = blockIdx.x, q = blockIdx.y, t = threadIdx.x;
  const int ch = p * nspin + q;
  const double nel = q == 0 ? nel_a : nel_b;
  const double* ev = w + (int64_t)p * sw;
  double* o = occ + (int64_t)ch * 3 * n;
  double m = mu[ch];
  if (bisect) {
    // bracket: every state fully occupied below lo, empty above hi
    double mn = 1e300, mx = -1e300;
    for (int i = t; i < n; i += blockDim.x) {
      mn = fmin(mn, ev[i]);
      mx = fmax(mx, ev[i]);
    }
    smn[t] = mn;
    smx[t] = mx;
    __syncthreads();
    for (int s2 = 128; s2 > 0; s2 >>= 1) {
      if (t < s2) {
        smn[t] = fmin(smn[t], smn[t + s2]);
        smx[t] = fmax(smx[t], smx[t + s2]);
      }
      __syncthreads();
    }
    double lo = smn[0] - 60.0 * kT - 1.0;
    double hi = smx[0] + 60.0 * kT + 1.0;
    for (int it = 0; it < 200; ++it) {
      const double mid = 0.5 * (lo + hi);
      double cnt = 0;
      for (int i = t; i < n; i += blockDim.x) cnt += 1.0 / (1.0 + exp((ev[i] - mid) / kT));
      cnt = block_sum(cnt, red);
      if (cnt < nel) lo = mid;
      else hi = mid;
      if (hi - lo <= 1e-15 * fmax(1.0, fabs(mid))) break;
    }
    m = 0.5 * (lo + hi);
    if (t == 0) mu[ch] = m;
  }
  for (int i = t; i < n; i += blockDim.x) {
    const double x = (ev[i] - m) / kT;
    const double lc = log_cosh_half(x);
    o[i] = 1.0 / (1.0 + exp(x));
    o[n + i] = -exp(-2.0 * lc) / (4.0 * kT);
    o[2 * n + i] = lc;
  }
}

// cf(:, j) = c(:, j) f_j, f_j summed over the spin channels (problem blockIdx.z)
__global__ void scale_columns_kernel(int n, int nspin, const double* __restrict__ c, int ldc, int64_t sc,
                                     const double* __restrict__ occ, double* __restrict__ cf, int64_t scf) {
  const int p = blockIdx.z, j = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double f = 0;
  for (int q = 0; q < nspin; ++q) f += occ[(int64_t)(p * nspin + q) * 3 * n + j];
  cf[(int64_t)p * scf + i + (int64_t)j * n] = c[(int64_t)p * sc + idx(i, j, ldc)] * f;
}

// kappa[ch] = sum_j f'_j Gamma_jj / sum_j f'_j per (problem, channel) (0 when mu is fixed or sum f' underflows)
__global__ __launch_bounds__(256) void kappa_kernel(int n, int nspin, const double* __restrict__ gam, int64_t sg,
                                                    const double* __restrict__ occ, int mu_fixed,
                                                    double* __restrict__ kappa) {
  __shared__ double red[4];
  const int p = blockIdx.x, q = blockIdx.y, t = threadIdx.x;
  const int ch = p * nspin + q;
  const double* g = gam + (int64_t)p * sg;
  const double* df = occ + (int64_t)ch * 3 * n + n;
  double num = 0, den = 0;
  for (int j = t; j < n; j += blockDim.x) {
    num += df[j] * g[j + (int64_t)j * n];
    den += df[j];
  }
  num = block_sum(num, red);
  den = block_sum(den, red);
  if (t == 0) kappa[ch] = (mu_fixed || den == 0.0) ? 0.0 : num / den;
}

// mh = F o sym(Gamma) - sum_q kappa_q diag(f'_q), ms = -(G o sym(Gamma) - sum_q kappa_q diag(eps f'_q)), F and G
// summed over the spin channels
__global__ void loewner_kernel(int n, int nspin, const double* __restrict__ w, int64_t sw,
                               const double* __restrict__ gam, int64_t sg, const double* __restrict__ occ,
                               const double* __restrict__ mu, const double* __restrict__ kappa, double kT,
                               double* __restrict__ mh, double* __restrict__ ms) {
  const int p = blockIdx.z, j = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* ev = w + (int64_t)p * sw;
  const double* g = gam + (int64_t)p * sg;
  const double ei = ev[i], ej = ev[j];
  const double gs = 0.5 * (g[i + (int64_t)j * n] + g[j + (int64_t)i * n]);
  double F = 0, G = 0, dh = 0, ds = 0;
  for (int q = 0; q < nspin; ++q) {
    const int ch = p * nspin + q;
    const double* o = occ + (int64_t)ch * 3 * n;
    const double m = mu[ch];
    const double xi = (ei - m) / kT, xj = (ej - m) / kT;
    const double Fq = fermi_divdiff(xi, xj, o[2 * n + i], o[2 * n + j], kT);
    F += Fq;
    G += 0.5 * (o[i] + o[j]) + 0.5 * (ei + ej) * Fq;
    if (i == j) {
      dh += kappa[ch] * o[n + i];
      ds += kappa[ch] * o[n + i] * ei;
    }
  }
  double vh = F * gs, vs = G * gs;
  if (i == j) {
    vh -= dh;
    vs -= ds;
  }
  mh[(int64_t)p * sg + i + (int64_t)j * n] = vh;
  ms[(int64_t)p * sg + i + (int64_t)j * n] = -vs;
}

}  // namespace

size_t density_workspace(int batch, int n) { return (size_t)batch * (4 * (size_t)n * n + 3 * (size_t)n + 8) + 64; }

void density_forward(hipStream_t s, int batch, int n, int nspin, const double* w, int64_t sw, const double* c,
                     int ldc, int64_t sc, double kT, double nel_a, double nel_b, int bisect, double* mu, double* p,
                     int ldp, int64_t sp, double* occ, double* work) {
  if (batch <= 0 || n <= 0) return;
  if (!(kT > 0)) throw std::runtime_error("density: kT must be positive");
  if (nspin != 1 && nspin != 2) throw std::runtime_error("density: nspin must be 1 or 2");
  hipLaunchKernelGGL(occupation_kernel, dim3(batch, nspin), dim3(256), 0, s, n, nspin, w, sw, kT, nel_a, nel_b, bisect,
                     mu, occ);
  XTB_LAUNCH_CHECK();
  const int64_t nn = (int64_t)n * n;
  double* cf = work;
  hipLaunchKernelGGL(scale_columns_kernel, dim3(cdiv(n, 256), n, batch), dim3(256), 0, s, n, nspin, c, ldc, sc, occ,
                     cf, nn);
  XTB_LAUNCH_CHECK();
  gemm<double>(s, 'N', 'T', n, n, n, 1.0, cf, n, c, ldc, 0.0, p, ldp, batch, nn, sc, sp);
}

void density_backward(hipStream_t s, int batch, int n, int nspin, const double* w, int64_t sw, const double* c,
                      int ldc, int64_t sc, double kT, const double* mu, int mu_fixed, const double* occ,
                      const double* gbar, int ldg, int64_t sg, double* hbar, int ldhb, int64_t shb, double* sbar,
                      int ldsb, int64_t ssb, double* work) {
  if (batch <= 0 || n <= 0) return;
  if (nspin != 1 && nspin != 2) throw std::runtime_error("density: nspin must be 1 or 2");
  const int64_t nn = (int64_t)n * n;
  double* t1 = work;
  double* gam = t1 + batch * nn;
  double* mh = gam + batch * nn;
  double* ms = mh + batch * nn;
  double* kappa = ms + batch * nn;
  // Gamma = C^T (Gbar C)
  gemm<double>(s, 'N', 'N', n, n, n, 1.0, gbar, ldg, c, ldc, 0.0, t1, n, batch, sg, sc, nn);
  gemm<double>(s, 'T', 'N', n, n, n, 1.0, c, ldc, t1, n, 0.0, gam, n, batch, sc, nn, nn);
  hipLaunchKernelGGL(kappa_kernel, dim3(batch, nspin), dim3(256), 0, s, n, nspin, gam, nn, occ, mu_fixed, kappa);
  XTB_LAUNCH_CHECK();
  hipLaunchKernelGGL(loewner_kernel, dim3(cdiv(n, 256), n, batch), dim3(256), 0, s, n, nspin, w, sw, gam, nn, occ, mu,
                     kappa, kT, mh, ms);
  XTB_LAUNCH_CHECK();
  // Hbar = C mh C^T, Sbar = C ms C^T
  gemm<double>(s, 'N', 'N', n, n, n, 1.0, c, ldc, mh, n, 0.0, t1, n, batch, sc, nn, nn);
  gemm<double>(s, 'N', 'T', n, n, n, 1.0, t1, n, c, ldc, 0.0, hbar, ldhb, batch, nn, sc, shb);
  gemm<double>(s, 'N', 'N', n, n, n, 1.0, c, ldc, ms, n, 0.0, t1, n, batch, sc, nn, nn);
  gemm<double>(s, 'N', 'T', n, n, n, 1.0, t1, n, c, ldc, 0.0, sbar, ldsb, batch, nn, sc, ssb);
}

}  // namespace xtbeig
