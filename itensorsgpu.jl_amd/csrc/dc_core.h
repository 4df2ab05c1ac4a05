// Numerical core of the symmetric tridiagonal divide-and-conquer eigensolver
// (Cuppen 1981; Gu & Eisenstat, SIAM J. Matrix Anal. Appl. 16 (1995)),
// shared by the GPU kernels (eigh.hip) and a CPU harness used in tests.
//
// One merge solves diag(dl) + rho * w w^T (rho > 0, |w| = 1) after
// deflation:
//  - dc_deflate: drop entries with negligible w and rotate away one of each
//    pair of (nearly) equal dl (the deflation rules of LAPACK dlaed2);
//  - dc_secular_root: root i of f(x) = 1 + rho sum_j w_j^2/(dl_j - x) as
//    (origin o, offset tau) with x = dl_o + tau, by a bracketed iteration on
//    a two-pole rational model of f (fixed-weight scheme of Bunch, Nielsen &
//    Sorensen 1978) with bisection as the safeguard;
//  - dc_zhat: w recomputed from the computed roots (Gu-Eisenstat), which
//    makes the eigenvectors zhat_i / (dl_i - x_j) orthogonal to working
//    precision.
#pragma once

#include <math.h>
#include <stdint.h>

#ifdef __HIPCC__
#define ITG_HD __host__ __device__
#else
#define ITG_HD
#endif

#define ITG_DC_EPS 1.1102230246251565e-16

// f(tau) with origin o; also the derivative split into the part from poles on
// the origin's side of the root (dpo) and from the far side (dpf).
ITG_HD static inline double dc_secular_eval(int k, const double* dl, const double* w, double rho,
                                            int o, double tau, int side, double* dpo,
                                            double* dpf, double* absum) {
  // side > 0: root right of the origin pole (origin side = j <= o)
  // side < 0: root left of the origin pole  (origin side = j >= o)
  double f = 1.0, po = 0.0, pf = 0.0, ab = 1.0;
  double dlo = dl[o];
  for (int j = 0; j < k; ++j) {
    double del = (dl[j] - dlo) - tau;
    double t = w[j] / del;
    double term = rho * w[j] * t;
    f += term;
    ab += fabs(term);
    double dt = rho * t * t;
    bool origin_side = side > 0 ? (j <= o) : (j >= o);
    if (origin_side) po += dt; else pf += dt;
  }
  *dpo = po;
  *dpf = pf;
  *absum = ab;
  return f;
}

// Root i of the secular equation, given `ev(o, tau, side, &dpo, &dpf, &ab)`
// that returns f at x = dl[o] + tau with the derivative parts and the
// absolute sum as in dc_secular_eval (serial on the host; the GPU kernel
// passes a wave-parallel evaluation of the same sums).
template <class Eval>
ITG_HD static inline void dc_secular_root_with(int i, int k, const double* dl, double rho,
                                               double w2sum, Eval&& ev, int* org_out,
                                               double* tau_out) {
  int o, side;
  double lo, hi, D;
  if (i == k - 1) {
    o = i;
    side = 1;
    lo = 0.0;
    hi = rho * w2sum;
    D = 0.0;   // no pole on the right
  } else {
    double gap = dl[i + 1] - dl[i];
    double mid = 0.5 * gap;
    double d0, d1, d2;
    double f = ev(i, mid, 1, &d0, &d1, &d2);
    if (f >= 0.0) {
      o = i; side = 1; lo = 0.0; hi = mid; D = gap;
    } else {
      o = i + 1; side = -1; lo = mid - gap; hi = 0.0; D = -gap;
    }
  }
  double tau = 0.5 * (lo + hi);
  for (int it = 0; it < 200; ++it) {
    double dpo, dpf, ab;
    double f = ev(o, tau, side, &dpo, &dpf, &ab);
    if (f > 0.0) hi = tau; else lo = tau;
    if (fabs(f) <= 4.0 * ITG_DC_EPS * ab) break;
    if (hi - lo <= 2.0 * ITG_DC_EPS * fmax(fabs(lo), fabs(hi))) break;
    // two-pole model c + a/(-x) + b/(D - x) matching f and f' at tau
    double a = tau * tau * dpo;
    double xn;
    if (D == 0.0 || dpf == 0.0) {
      double c = f + a / tau;            // f - a/(-tau); far side empty
      xn = (c > 0.0) ? a / c : 0.5 * (lo + hi);
    } else {
      double b = (D - tau) * (D - tau) * dpf;
      double c = f + a / tau - b / (D - tau);
      double A2 = c, B2 = -(c * D + a + b), C2 = a * D;
      double disc = B2 * B2 - 4.0 * A2 * C2;
      if (disc < 0.0) disc = 0.0;
      double sq = sqrt(disc);
      double q = -0.5 * (B2 + (B2 >= 0.0 ? sq : -sq));
      double x1 = (A2 != 0.0) ? q / A2 : (lo - 1.0);
      double x2 = (q != 0.0) ? C2 / q : (lo - 1.0);
      if (x1 > lo && x1 < hi) xn = x1;
      else if (x2 > lo && x2 < hi) xn = x2;
      else xn = 0.5 * (lo + hi);
    }
    if (!(xn > lo && xn < hi)) xn = 0.5 * (lo + hi);
    if (fabs(xn - tau) <= 2.0 * ITG_DC_EPS * fabs(xn)) { tau = xn; break; }
    tau = xn;
  }
  *org_out = o;
  *tau_out = tau;
}

ITG_HD static inline void dc_secular_root(int i, int k, const double* dl, const double* w,
                                          double rho, double w2sum, int* org_out,
                                          double* tau_out) {
  dc_secular_root_with(
      i, k, dl, rho, w2sum,
      [&](int o, double tau, int side, double* dpo, double* dpf, double* ab) {
        return dc_secular_eval(k, dl, w, rho, o, tau, side, dpo, dpf, ab);
      },
      org_out, tau_out);
}

// zhat_i from the roots (x_j = dl[org_j] + tau_j):
// zhat_i^2 = -prod_j (dl_i - x_j) / prod_{j != i} (dl_i - dl_j), up to 1/rho
ITG_HD static inline double dc_zhat(int i, int k, const double* dl, const int* org,
                                    const double* tau, const double* w) {
  double p = (dl[i] - dl[org[i]]) - tau[i];
  for (int j = 0; j < k; ++j) {
    if (j == i) continue;
    p *= ((dl[i] - dl[org[j]]) - tau[j]) / (dl[i] - dl[j]);
  }
  double z2 = -p;
  if (z2 < 0.0) z2 = 0.0;
  return copysign(sqrt(z2), w[i]);
}

// Deflation of one merge. Inputs sorted ascending by d (length s): d, z
// (|z| = 1), col (column of Q holding each entry's vector); rho > 0.
// Outputs: non-deflated (dl, w, gcol)[0..k); deflated (dval, dcol)[0..nd);
// rotations (rcol pairs, rcs = (c, s)) to apply to Q's columns in order:
// q_p <- c q_p + s q_n, q_n <- c q_n - s q_p. Returns k. d and z are read
// only.
//
// The rules are dlaed2's: an entry with rho |z| <= tol deflates; of two
// consecutive kept entries p, j with |t c s| <= tol (t = d_j - d_p,
// c = z_j / tau, s = -z_p / tau, tau^2 = z_p^2 + z_j^2) the rotation moves
// p's weight onto j, p deflates at d_p c^2 + d_j s^2 and j goes on at
// d_p s^2 + d_j c^2 = (d_p z_p^2 + d_j z_j^2) / tau^2 with z_j = tau. Along a
// run of such rotations the entry carried is thus the z^2-weighted mean of
// the run's d with weight sqrt(sum z^2): the pass keeps N = sum d z^2 and
// Z2 = sum z^2 and makes the test in the squared form
// (d_j Z2 - N)^2 z_j^2 <= tol^2 Z2 (Z2 + z_j^2)^2 (|t c s| <= tol squared
// and multiplied through by Z2 (Z2 + z_j^2)^2 > 0), so the step from one
// entry to the next is a few multiply-adds; square roots and divisions are
// only taken for the values written out. A run of one entry keeps its own d
// and signed z.
ITG_HD static inline int dc_deflate(int s, const double* d, const double* z, const int* col,
                                    double rho, double* dl, double* w, int* gcol, double* dval,
                                    int* dcol, int* nd_out, int* rcol, double* rcs,
                                    int* nrot_out) {
  double dmax = 0.0, zmax = 0.0;
  for (int j = 0; j < s; ++j) {
    dmax = fmax(dmax, fabs(d[j]));
    zmax = fmax(zmax, fabs(z[j]));
  }
  const double tol = 8.0 * ITG_DC_EPS * fmax(dmax, zmax);
  const double tol2 = tol * tol;
  int k = 0, nd = 0, nrot = 0;
  if (rho * zmax <= tol) {
    for (int j = 0; j < s; ++j) { dval[nd] = d[j]; dcol[nd] = col[j]; ++nd; }
    *nd_out = nd;
    *nrot_out = 0;
    return 0;
  }
  // the carried entry: column pc (-1: none yet), run sums N and Z2, and for
  // a run of one entry its own d0, z0
  int pc = -1;
  bool single = true;
  double N = 0.0, Z2 = 0.0, d0 = 0.0, z0 = 0.0;
  double dn = s > 0 ? d[0] : 0.0, zn = s > 0 ? z[0] : 0.0;
  for (int j = 0; j < s; ++j) {
    const double dj = dn, zj = zn;
    if (j + 1 < s) {   // the next entry's loads go out before this one's work
      dn = d[j + 1];
      zn = z[j + 1];
    }
    if (rho * fabs(zj) <= tol) {
      dval[nd] = dj; dcol[nd] = col[j]; ++nd;
      continue;
    }
    const double zj2 = zj * zj;
    if (pc < 0) {
      pc = col[j]; single = true; d0 = dj; z0 = zj; N = dj * zj2; Z2 = zj2;
      continue;
    }
    const double T2 = Z2 + zj2;
    const double x = dj * Z2 - N;
    if (x * x * zj2 <= tol2 * Z2 * T2 * T2) {
      // |z| <= 1 here and kept entries have rho |z| > tol, so the plain norms
      // neither overflow nor underflow
      const double tau = sqrt(T2), it = 1.0 / tau;
      const double c = zj * it, sn = -(single ? z0 : sqrt(Z2)) * it;
      const double dp = single ? d0 : N / Z2;
      rcol[2 * nrot] = pc;
      rcol[2 * nrot + 1] = col[j];
      rcs[2 * nrot] = c;
      rcs[2 * nrot + 1] = sn;
      ++nrot;
      dval[nd] = dp * c * c + dj * sn * sn; dcol[nd] = pc; ++nd;
      pc = col[j]; single = false; N += dj * zj2; Z2 = T2;
    } else {
      dl[k] = single ? d0 : N / Z2; w[k] = single ? z0 : sqrt(Z2); gcol[k] = pc; ++k;
      pc = col[j]; single = true; d0 = dj; z0 = zj; N = dj * zj2; Z2 = zj2;
    }
  }
  if (pc >= 0) {
    dl[k] = single ? d0 : N / Z2; w[k] = single ? z0 : sqrt(Z2); gcol[k] = pc; ++k;
  }
  *nd_out = nd;
  *nrot_out = nrot;
  return k;
}
