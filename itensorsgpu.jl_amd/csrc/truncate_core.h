// truncate! decision shared verbatim by the host entry point and the device
// kernel, so both produce identical (n, truncerr, docut).
//
// ITG_TRUNC_NDTENSORS: the CPU NDTensors truncate! (ITensors 0.2), which the
//   reference's CPU path runs: drop from the tail while n > maxdim, then while
//   the *cumulative* discarded weight stays <= cutoff*scale and n > mindim.
// ITG_TRUNC_REFGPU: src/tensor/cutruncate.jl:1-93 as written, including its
//   per-weight cutoff test and the n = length(P) - cut_ind rule.
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#define ITG_HD __host__ __device__
#else
#define ITG_HD
#endif

struct itg_trunc_result {
  int64_t n;
  double truncerr;
  double docut;
};

ITG_HD static inline double itg_trunc_docut(const double* P, int64_t n, int64_t origm) {
  // cutruncate.jl:81-87
  double docut = 0.0;
  if (n < origm) {
    docut = (P[n - 1] + P[n]) / 2.0;
    double diff = P[n - 1] - P[n];
    if (diff < 0) diff = -diff;
    if (diff < 1e-3 * P[n - 1]) docut += 1e-3 * P[n - 1];
  }
  return docut;
}

ITG_HD static inline double itg_nonneg(double x) {
  // map(x -> !signbit(x) ? x : 0.0, P): cutruncate.jl:23
  return (x > 0.0 || (x == 0.0 && !(1.0 / x < 0.0))) ? x : 0.0;
}

ITG_HD static inline itg_trunc_result itg_truncate_core(const double* P, int64_t origm,
                                                        int64_t maxdim, int64_t mindim,
                                                        double cutoff, int abs_cut,
                                                        int rel_cut, int semantics) {
  itg_trunc_result r = {1, 0.0, 0.0};
  if (origm <= 0) { r.n = 0; return r; }
  if (maxdim <= 0 || maxdim > origm) maxdim = origm;
  if (semantics == 1) {
    // ------------------------------------------------ reference GPU kernel
    if (mindim > maxdim) mindim = maxdim;                         // :4
    double maxP = P[0];
    for (int64_t i = 1; i < origm; ++i) maxP = P[i] > maxP ? P[i] : maxP;
    if (maxP == 0.0) return r;                                    // :11-14
    if (origm == 1) { r.docut = maxP / 2; return r; }             // :15-18
    int64_t n = origm;
    double truncerr = 0.0;
    if (n > maxdim) {                                             // :26-29
      for (int64_t i = 0; i < n - maxdim; ++i) truncerr += itg_nonneg(P[i]);
      n = maxdim;
    }
    if (abs_cut) {                                                // :32-41
      int64_t cut = 0;
      for (int64_t i = 0; i < origm; ++i)
        if (itg_nonneg(P[i]) - cutoff < 0.0) { cut = i; break; }
      n = origm - cut < maxdim ? origm - cut : maxdim;
      if (n < mindim) n = mindim;
      for (int64_t i = cut; i < origm; ++i) truncerr += itg_nonneg(P[i]);
    } else {
      double scale = 1.0;
      if (rel_cut) {                                              // :45-48
        scale = 0.0;
        for (int64_t i = 0; i < origm; ++i) scale += P[i];
        scale = scale > 0.0 ? scale : 1.0;
      }
      int64_t cut = 0;                                            // :53-56
      for (int64_t i = 0; i < origm; ++i)
        if (itg_nonneg(P[i]) + truncerr - cutoff * scale < 0.0) { cut = i; break; }
      if (cut > 0) {
        for (int64_t i = cut; i < origm; ++i) truncerr += itg_nonneg(P[i]);
      } else {
        for (int64_t i = 0; i < maxdim; ++i) truncerr += itg_nonneg(P[i]);
      }
      n = origm - cut < maxdim ? origm - cut : maxdim;
      if (n < mindim) n = mindim;
      truncerr = (scale == 0.0) ? 0.0 : truncerr / scale;
    }
    if (n < 1) n = 1;
    r.n = n;
    r.truncerr = truncerr;
    r.docut = itg_trunc_docut(P, n, origm);
    return r;
  }
  // -------------------------------------------------- NDTensors CPU semantics
  if (mindim < 1) mindim = 1;
  if (cutoff < 0.0) cutoff = 0.0;
  if (P[0] <= 0.0) return r;
  if (origm == 1) { r.docut = P[0] / 2; return r; }
  // negative weights at the tail count as zero
  int64_t lastpos = origm;
  while (lastpos > 0 && P[lastpos - 1] < 0.0) --lastpos;
#define ITG_W(i) ((i) < lastpos ? P[(i)] : 0.0)
  int64_t n = origm;
  double truncerr = 0.0;
  while (n > maxdim) { truncerr += ITG_W(n - 1); --n; }
  if (abs_cut) {
    while (n > mindim && ITG_W(n - 1) <= cutoff) { truncerr += ITG_W(n - 1); --n; }
  } else {
    double scale = 1.0;
    if (rel_cut) {
      scale = 0.0;
      for (int64_t i = 0; i < origm; ++i) scale += ITG_W(i);
      if (scale == 0.0) scale = 1.0;
    }
    while (n > mindim && truncerr + ITG_W(n - 1) <= cutoff * scale) {
      truncerr += ITG_W(n - 1);
      --n;
    }
    truncerr /= scale;
  }
  if (n < 1) n = 1;
  r.n = n;
  r.truncerr = truncerr;
  // docut uses the clamped weights
  if (n < origm) {
    double a = ITG_W(n - 1), b = ITG_W(n);
    r.docut = (a + b) / 2.0;
    double diff = a - b;
    if (diff < 0) diff = -diff;
    if (diff < 1e-3 * a) r.docut += 1e-3 * a;
  }
#undef ITG_W
  return r;
}
