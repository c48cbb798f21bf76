// The secular equation of a rank-one modified diagonal eigenproblem, one root per GPU lane:
//     f(lambda) = 1/rho + sum_j z_j^2 / (d_j - lambda) = 0,   d_0 < d_1 < ... < d_{K-1}, rho > 0, ||z|| = 1,
// root i in (d_i, d_{i+1}) (the last in (d_{K-1}, d_{K-1} + rho)). This restates the method of LAPACK's dlaed4 /
// dlaed5 / dlaed6 (R.-C. Li, "Solving secular equations stably and efficiently", LAPACK Working Note 89; the
// "middle way" of rational interpolation with the origin shifted to the nearer pole, switching to three-pole
// interpolation when that converges faster; Gragg's cubically convergent scheme for the three-pole equation), which
// the reference calls per root in its divide-and-conquer merge (Stages/ParallelDivideAndConquer.swift:134-145).
//
// Differences from the Fortran routine: the differences delta_j = d_j - lambda are recomputed from the origin pole and
// the current shift each iteration, (d_j - d_origin) - tau, instead of being updated by repeated subtraction, and the
// caller receives (origin index, tau) so that every later kernel can form delta_j the same way.
//
// The secular solver and the deflation follow LAPACK's dlaed2 / dlaed4 / dlaed5 / dlaed6 closely (their variable
// names and branch structure), so LAPACK's license applies to those parts:
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
#pragma once
#include "common.hpp"

namespace xtbeig {
namespace secular {

constexpr double EPS = 1.1102230246251565e-16;  // 2^-53, LAPACK dlamch('E')
constexpr int MAXIT = 30;

// dlaed6: the root tau of  rho + z1/(d1 - tau) + z2/(d2 - tau) + z3/(d3 - tau) = 0 nearest 0 in the interval
// given by orgati (between d2 and d3, else d1 and d2); finit = value at tau = 0. Returns info (0: converged).
__device__ inline int laed6(int kniter, bool orgati, double rho, const double d[3], const double z[3], double finit,
                            double& tau) {
  const int maxit = 40;
  double lbd, ubd;
  if (orgati) { lbd = d[1]; ubd = d[2]; } else { lbd = d[0]; ubd = d[1]; }
  if (finit < 0) lbd = 0; else ubd = 0;
  int niter = 1;
  tau = 0;
  if (kniter == 2) {
    double a, b, c, temp;
    if (orgati) {
      temp = (d[2] - d[1]) / 2;
      c = rho + z[0] / ((d[0] - d[1]) - temp);
      a = c * (d[1] + d[2]) + z[1] + z[2];
      b = c * d[1] * d[2] + z[1] * d[2] + z[2] * d[1];
    } else {
      temp = (d[0] - d[1]) / 2;
      c = rho + z[2] / ((d[2] - d[1]) - temp);
      a = c * (d[0] + d[1]) + z[0] + z[1];
      b = c * d[0] * d[1] + z[0] * d[1] + z[1] * d[0];
    }
    temp = fmax(fabs(a), fmax(fabs(b), fabs(c)));
    a /= temp; b /= temp; c /= temp;
    if (c == 0) tau = b / a;
    else if (a <= 0) tau = (a - sqrt(fabs(a * a - 4 * b * c))) / (2 * c);
    else tau = 2 * b / (a + sqrt(fabs(a * a - 4 * b * c)));
    if (tau < lbd || tau > ubd) tau = (lbd + ubd) / 2;
    if (d[0] == tau || d[1] == tau || d[2] == tau) {
      tau = 0;
    } else {
      temp = finit + tau * z[0] / (d[0] * (d[0] - tau)) + tau * z[1] / (d[1] * (d[1] - tau)) +
             tau * z[2] / (d[2] * (d[2] - tau));
      if (temp <= 0) lbd = tau; else ubd = tau;
      if (fabs(finit) <= fabs(temp)) tau = 0;
    }
  }
  // Scaling against overflow of 1/temp^3 when tau is within ~safmin^(1/3) of a pole.
  const double small1 = 0x1p-340, sminv1 = 0x1p340, small2 = 0x1p-680, sminv2 = 0x1p680;  // radix powers
  double temp = orgati ? fmin(fabs(d[1] - tau), fabs(d[2] - tau)) : fmin(fabs(d[0] - tau), fabs(d[1] - tau));
  bool scale = false;
  double sclfac = 1, sclinv = 1;
  double ds[3], zs[3];
  if (temp <= small1) {
    scale = true;
    if (temp <= small2) { sclfac = sminv2; sclinv = small2; } else { sclfac = sminv1; sclinv = small1; }
    for (int i = 0; i < 3; ++i) { ds[i] = d[i] * sclfac; zs[i] = z[i] * sclfac; }
    tau *= sclfac; lbd *= sclfac; ubd *= sclfac;
  } else {
    for (int i = 0; i < 3; ++i) { ds[i] = d[i]; zs[i] = z[i]; }
  }
  double fc = 0, df = 0, ddf = 0;
  for (int i = 0; i < 3; ++i) {
    const double t = 1 / (ds[i] - tau), t1 = zs[i] * t, t2 = t1 * t, t3 = t2 * t;
    fc += t1 / ds[i];
    df += t2;
    ddf += t3;
  }
  double f = finit + tau * fc;
  int info = 0;
  if (fabs(f) <= 0) goto done;
  if (f <= 0) lbd = tau; else ubd = tau;
  for (niter = niter + 1; niter <= maxit; ++niter) {
    double t1, t2;
    if (orgati) { t1 = ds[1] - tau; t2 = ds[2] - tau; } else { t1 = ds[0] - tau; t2 = ds[1] - tau; }
    double a = (t1 + t2) * f - t1 * t2 * df;
    double b = t1 * t2 * f;
    double c = f - (t1 + t2) * df + t1 * t2 * ddf;
    const double tm = fmax(fabs(a), fmax(fabs(b), fabs(c)));
    a /= tm; b /= tm; c /= tm;
    double eta;
    if (c == 0) eta = b / a;
    else if (a <= 0) eta = (a - sqrt(fabs(a * a - 4 * b * c))) / (2 * c);
    else eta = 2 * b / (a + sqrt(fabs(a * a - 4 * b * c)));
    if (f * eta >= 0) eta = -f / df;
    tau += eta;
    if (tau < lbd || tau > ubd) tau = (lbd + ubd) / 2;
    fc = 0; df = 0; ddf = 0;
    double erretm = 0;
    for (int i = 0; i < 3; ++i) {
      if (ds[i] - tau != 0) {
        const double t = 1 / (ds[i] - tau), u1 = zs[i] * t, u2 = u1 * t, u3 = u2 * t, u4 = u1 / ds[i];
        fc += u4;
        erretm += fabs(u4);
        df += u2;
        ddf += u3;
      } else {
        goto done;
      }
    }
    f = finit + tau * fc;
    erretm = 8 * (fabs(finit) + fabs(tau) * erretm) + fabs(tau) * df;
    if (fabs(f) <= 4 * EPS * erretm || (ubd - lbd) <= 4 * EPS * fabs(tau)) goto done;
    if (f <= 0) lbd = tau; else ubd = tau;
  }
  info = 1;
done:
  if (scale) tau *= sclinv;
  return info;
}

// dlaed5: both roots of the 2 x 2 problem; writes the normalized eigenvector of root i (0 or 1) to v and the root to
// lam.
__device__ inline void laed5(int i, const double d[2], const double z[2], double rho, double v[2], double& lam) {
  const double del = d[1] - d[0];
  double tau, b, c;
  if (i == 0) {
    const double w = 1 + 2 * rho * (z[1] * z[1] - z[0] * z[0]) / del;
    if (w > 0) {
      b = del + rho * (z[0] * z[0] + z[1] * z[1]);
      c = rho * z[0] * z[0] * del;
      tau = 2 * c / (b + sqrt(fabs(b * b - 4 * c)));
      lam = d[0] + tau;
      v[0] = -z[0] / tau;
      v[1] = z[1] / (del - tau);
    } else {
      b = -del + rho * (z[0] * z[0] + z[1] * z[1]);
      c = rho * z[1] * z[1] * del;
      if (b > 0) tau = -2 * c / (b + sqrt(b * b + 4 * c));
      else tau = (b - sqrt(b * b + 4 * c)) / 2;
      lam = d[1] + tau;
      v[0] = -z[0] / (del + tau);
      v[1] = -z[1] / tau;
    }
  } else {
    b = -del + rho * (z[0] * z[0] + z[1] * z[1]);
    c = rho * z[1] * z[1] * del;
    if (b > 0) tau = (b + sqrt(b * b + 4 * c)) / 2;
    else tau = 2 * c / (-b + sqrt(b * b + 4 * c));
    lam = d[1] + tau;
    v[0] = -z[0] / (del + tau);
    v[1] = -z[1] / tau;
  }
  const double nrm = sqrt(v[0] * v[0] + v[1] * v[1]);
  v[0] /= nrm;
  v[1] /= nrm;
}

// Sums over the poles for the current shift: psi (poles 0 .. ilo - 1) and phi (poles ihi .. k - 1), their
// derivatives and the error bound's accumulation. G lanes share one root (lane gl of the group takes poles
// gl, gl + G, ...; the group's sums are combined so that every lane holds the same bits). The Fortran loop adds the
// running psi after each pole (and the running phi, from the top) into the error bound: that is the weighted sum
// sum_j term_j (ilo - j) (and sum_j term_j (j - ihi + 1)), formed here directly.
struct Sums { double psi, dpsi, phi, dphi, erretm; };
template <int G>
__device__ inline double gsum(double v) {
  if constexpr (G == 1) return v;
  else return group_sum_lanes<G>(v);
}
template <int G>
__device__ inline Sums pole_sums(int k, int ilo, int ihi, const double* d, const double* z, double dorg, double tau,
                                 int gl) {
  double psi = 0, dpsi = 0, phi = 0, dphi = 0, e1 = 0, e2 = 0;
  for (int j = gl; j < ilo; j += G) {
    const double dj = (d[j] - dorg) - tau;
    const double t = z[j] / dj;
    const double term = z[j] * t;
    psi += term;
    dpsi += t * t;
    e1 += term * (double)(ilo - j);
  }
  for (int j = ihi + gl; j < k; j += G) {
    const double dj = (d[j] - dorg) - tau;
    const double t = z[j] / dj;
    const double term = z[j] * t;
    phi += term;
    dphi += t * t;
    e2 += term * (double)(j - ihi + 1);
  }
  Sums r;
  r.psi = gsum<G>(psi);
  r.dpsi = gsum<G>(dpsi);
  r.phi = gsum<G>(phi);
  r.dphi = gsum<G>(dphi);
  r.erretm = fabs(gsum<G>(e1)) + gsum<G>(e2);
  return r;
}

// dlaed4 for root i (0-based) of the k >= 3 problem, on the G lanes of a group (gl: lane within the group; every
// lane returns the same result). Outputs the origin pole index and tau with lambda = d[orig] + tau. Returns info
// (0: converged).
template <int G>
__device__ inline int laed4(int k, int i, const double* d, const double* z, double rho, int& orig, double& tau,
                            int gl, int maxit = MAXIT) {
  const double rhoinv = 1 / rho;
  if (i == k - 1) {
    // the largest root, origin d[k - 1]
    const int n = k, ii = n - 2;  // 0-based index of pole N-1
    orig = n - 1;
    const double dorg = d[n - 1];
    const double midpt = rho / 2;
    double psi = 0;
    for (int j = gl; j < n - 2; j += G) psi += z[j] * z[j] / ((d[j] - dorg) - midpt);
    psi = gsum<G>(psi);
    double c = rhoinv + psi;
    double w = c + z[ii] * z[ii] / ((d[ii] - dorg) - midpt) + z[n - 1] * z[n - 1] / ((d[n - 1] - dorg) - midpt);
    double dltlb, dltub;
    if (w <= 0) {
      const double temp = z[n - 2] * z[n - 2] / (d[n - 1] - d[n - 2] + rho) + z[n - 1] * z[n - 1] / rho;
      if (c <= temp) {
        tau = rho;
      } else {
        const double del = d[n - 1] - d[n - 2];
        const double a = -c * del + z[n - 2] * z[n - 2] + z[n - 1] * z[n - 1];
        const double b = z[n - 1] * z[n - 1] * del;
        if (a < 0) tau = 2 * b / (sqrt(a * a + 4 * b * c) - a);
        else tau = (a + sqrt(a * a + 4 * b * c)) / (2 * c);
      }
      dltlb = midpt;
      dltub = rho;
    } else {
      const double del = d[n - 1] - d[n - 2];
      const double a = -c * del + z[n - 2] * z[n - 2] + z[n - 1] * z[n - 1];
      const double b = z[n - 1] * z[n - 1] * del;
      if (a < 0) tau = 2 * b / (sqrt(a * a + 4 * b * c) - a);
      else tau = (a + sqrt(a * a + 4 * b * c)) / (2 * c);
      dltlb = 0;
      dltub = midpt;
    }
    for (int niter = 1; niter <= maxit + 1; ++niter) {
      // psi over poles 0 .. n - 2, phi from pole n - 1 alone
      Sums sm = pole_sums<G>(n, ii + 1, n, d, z, dorg, tau, gl);
      const double dn = (d[n - 1] - dorg) - tau, dnm1 = (d[n - 2] - dorg) - tau;
      const double t = z[n - 1] / dn;
      const double phi = z[n - 1] * t, dphi = t * t;
      const double erretm = 8 * (-phi - sm.psi) + sm.erretm - phi + rhoinv + fabs(tau) * (sm.dpsi + dphi);
      w = rhoinv + phi + sm.psi;
      if (fabs(w) <= EPS * erretm) return 0;
      if (niter > maxit) break;
      if (w <= 0) dltlb = fmax(dltlb, tau); else dltub = fmin(dltub, tau);
      double cc = w - dnm1 * sm.dpsi - dn * dphi;
      const double a = (dnm1 + dn) * w - dnm1 * dn * (sm.dpsi + dphi);
      const double b = dnm1 * dn * w;
      if (cc < 0) cc = fabs(cc);
      double eta;
      if (cc == 0) eta = -w / (sm.dpsi + dphi);
      else if (a >= 0) eta = (a + sqrt(fabs(a * a - 4 * b * cc))) / (2 * cc);
      else eta = 2 * b / (a - sqrt(fabs(a * a - 4 * b * cc)));
      if (w * eta > 0) eta = -w / (sm.dpsi + dphi);
      const double temp = tau + eta;
      if (temp > dltub || temp < dltlb) eta = (w < 0) ? (dltub - tau) / 2 : (dltlb - tau) / 2;
      tau += eta;
    }
    return 1;
  }
  // 0 <= i < k - 1: root in (d[i], d[i + 1])
  const int ip1 = i + 1;
  const double del = d[ip1] - d[i];
  const double midpt = del / 2;
  double psi = 0, phi = 0;
  for (int j = gl; j < i; j += G) psi += z[j] * z[j] / ((d[j] - d[i]) - midpt);
  for (int j = i + 2 + gl; j < k; j += G) phi += z[j] * z[j] / ((d[j] - d[i]) - midpt);
  psi = gsum<G>(psi);
  phi = gsum<G>(phi);
  double c = rhoinv + psi + phi;
  double w = c + z[i] * z[i] / ((d[i] - d[i]) - midpt) + z[ip1] * z[ip1] / ((d[ip1] - d[i]) - midpt);
  bool orgati;
  double dltlb, dltub;
  if (w > 0) {
    orgati = true;
    const double a = c * del + z[i] * z[i] + z[ip1] * z[ip1];
    const double b = z[i] * z[i] * del;
    if (a > 0) tau = 2 * b / (a + sqrt(fabs(a * a - 4 * b * c)));
    else tau = (a - sqrt(fabs(a * a - 4 * b * c))) / (2 * c);
    dltlb = 0;
    dltub = midpt;
  } else {
    orgati = false;
    const double a = c * del - z[i] * z[i] - z[ip1] * z[ip1];
    const double b = z[ip1] * z[ip1] * del;
    if (a < 0) tau = 2 * b / (a - sqrt(fabs(a * a + 4 * b * c)));
    else tau = -(a + sqrt(fabs(a * a + 4 * b * c))) / (2 * c);
    dltlb = -midpt;
    dltub = 0;
  }
  orig = orgati ? i : ip1;
  const double dorg = d[orig];
  const int ii = orig, iim1 = ii - 1, iip1 = ii + 1;
  auto delta = [&](int j, double tt) { return (d[j] - dorg) - tt; };
  Sums sm = pole_sums<G>(k, ii, iip1, d, z, dorg, tau, gl);
  w = rhoinv + sm.phi + sm.psi;
  bool swtch3 = orgati ? (w < 0) : (w > 0);
  if (ii == 0 || ii == k - 1) swtch3 = false;
  double temp = z[ii] / delta(ii, tau);
  double dw = sm.dpsi + sm.dphi + temp * temp;
  temp = z[ii] * temp;
  w += temp;
  double erretm = 8 * (sm.phi - sm.psi) + sm.erretm + 2 * rhoinv + 3 * fabs(temp) + fabs(tau) * dw;
  if (fabs(w) <= EPS * erretm) return 0;
  if (w <= 0) dltlb = fmax(dltlb, tau); else dltub = fmin(dltub, tau);
  bool swtch = false;
  for (int niter = 2; niter <= maxit; ++niter) {
    double dpsi = sm.dpsi, dphi = sm.dphi;
    const double di = delta(i, tau), dip1 = delta(ip1, tau);
    double eta;
    if (!swtch3) {
      double cc;
      if (!swtch) {
        if (orgati) cc = w - dip1 * dw - (d[i] - d[ip1]) * (z[i] / di) * (z[i] / di);
        else cc = w - di * dw - (d[ip1] - d[i]) * (z[ip1] / dip1) * (z[ip1] / dip1);
      } else {
        const double t = z[ii] / delta(ii, tau);
        if (orgati) dpsi += t * t; else dphi += t * t;
        cc = w - di * dpsi - dip1 * dphi;
      }
      double a = (di + dip1) * w - di * dip1 * dw;
      const double b = di * dip1 * w;
      if (cc == 0) {
        if (a == 0) {
          if (!swtch) {
            if (orgati) a = z[i] * z[i] + dip1 * dip1 * (dpsi + dphi);
            else a = z[ip1] * z[ip1] + di * di * (dpsi + dphi);
          } else {
            a = di * di * dpsi + dip1 * dip1 * dphi;
          }
        }
        eta = b / a;
      } else if (a <= 0) {
        eta = (a - sqrt(fabs(a * a - 4 * b * cc))) / (2 * cc);
      } else {
        eta = 2 * b / (a + sqrt(fabs(a * a - 4 * b * cc)));
      }
    } else {
      // three poles iim1, ii, iip1
      const double t0 = rhoinv + sm.psi + sm.phi;
      const double dm = delta(iim1, tau), dp = delta(iip1, tau);
      double cc, zz[3];
      if (swtch) {
        cc = t0 - dm * dpsi - dp * dphi;
        zz[0] = dm * dm * dpsi;
        zz[2] = dp * dp * dphi;
      } else if (orgati) {
        double t1 = z[iim1] / dm;
        t1 *= t1;
        cc = t0 - dp * (dpsi + dphi) - (d[iim1] - d[iip1]) * t1;
        zz[0] = z[iim1] * z[iim1];
        zz[2] = dp * dp * ((dpsi - t1) + dphi);
      } else {
        double t1 = z[iip1] / dp;
        t1 *= t1;
        cc = t0 - dm * (dpsi + dphi) - (d[iip1] - d[iim1]) * t1;
        zz[0] = dm * dm * (dpsi + (dphi - t1));
        zz[2] = z[iip1] * z[iip1];
      }
      zz[1] = z[ii] * z[ii];
      const double dd[3] = {dm, delta(ii, tau), dp};
      if (laed6(niter, orgati, cc, dd, zz, w, eta) != 0) return 1;
    }
    if (w * eta >= 0) eta = -w / dw;
    const double tn = tau + eta;
    if (tn > dltub || tn < dltlb) eta = (w < 0) ? (dltub - tau) / 2 : (dltlb - tau) / 2;
    const double prew = w;
    tau += eta;
    sm = pole_sums<G>(k, ii, iip1, d, z, dorg, tau, gl);
    temp = z[ii] / delta(ii, tau);
    dw = sm.dpsi + sm.dphi + temp * temp;
    temp = z[ii] * temp;
    w = rhoinv + sm.phi + sm.psi + temp;
    erretm = 8 * (sm.phi - sm.psi) + sm.erretm + 2 * rhoinv + 3 * fabs(temp) + fabs(tau) * dw;
    if (niter == 2) {
      swtch = orgati ? (-w > fabs(prew) / 10) : (w > fabs(prew) / 10);
    } else if (w * prew > 0 && fabs(w) > fabs(prew) / 10) {
      swtch = !swtch;
    }
    if (fabs(w) <= EPS * erretm) return 0;
    if (w <= 0) dltlb = fmax(dltlb, tau); else dltub = fmin(dltub, tau);
  }
  return 1;
}

}  // namespace secular
}  // namespace xtbeig
