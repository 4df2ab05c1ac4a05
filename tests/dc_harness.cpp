// CPU driver of the tridiagonal divide-and-conquer used by libitg's eigh,
// built from the same numerical core (itensorsgpu.jl_amd/csrc/dc_core.h) and
// the same bottom-up merge schedule as the GPU kernels, so the numerics can be
// checked against LAPACK without a GPU. Test infrastructure only.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "dc_core.h"

extern "C" int dc_eig(int n, const double* d_in, const double* e_in, double* evals,
                      double* Qout) {
  if (n <= 0) return 0;
  std::vector<double> D(n), Dn(n), Q((size_t)n * n, 0.0), Qn((size_t)n * n, 0.0);
  // tears at every split point: d_i -= |e_{i-1}| + |e_i|
  for (int i = 0; i < n; ++i) {
    double v = d_in[i];
    if (i > 0) v -= fabs(e_in[i - 1]);
    if (i < n - 1) v -= fabs(e_in[i]);
    D[i] = v;
    Q[(size_t)i * n + i] = 1.0;
  }
  std::vector<double> zs(n), ds(n), dl(n), w(n), dval(n), rcs(2 * n), tau(n), zh(n);
  std::vector<int> cs(n), gcol(n), dcol(n), rcol(2 * n), org(n);
  for (int half = 1; half < n; half *= 2) {
    int s_lvl = 2 * half;
    for (int r0 = 0; r0 < n; r0 += s_lvl) {
      int s = std::min(s_lvl, n - r0);
      int s1 = std::min(half, s);
      if (s1 == s) {  // no right sibling: carry over
        for (int c = r0; c < r0 + s; ++c) {
          Dn[c] = D[c];
          for (int r = r0; r < r0 + s; ++r) Qn[(size_t)c * n + r] = Q[(size_t)c * n + r];
        }
        continue;
      }
      double beta = e_in[r0 + s1 - 1];
      double rho = 2.0 * fabs(beta);
      double sg = beta >= 0 ? 1.0 : -1.0;
      std::vector<int> idx(s);
      std::iota(idx.begin(), idx.end(), r0);
      std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return D[a] < D[b]; });
      for (int i = 0; i < s; ++i) {
        int c = idx[i];
        double z = (c < r0 + s1) ? Q[(size_t)c * n + (r0 + s1 - 1)]
                                 : sg * Q[(size_t)c * n + (r0 + s1)];
        zs[i] = z * 0.7071067811865476;
        ds[i] = D[c];
        cs[i] = c;
      }
      int nd = 0, nrot = 0;
      int k = dc_deflate(s, ds.data(), zs.data(), cs.data(), rho, dl.data(), w.data(),
                         gcol.data(), dval.data(), dcol.data(), &nd, rcol.data(), rcs.data(),
                         &nrot);
      for (int t = 0; t < nrot; ++t) {
        int p = rcol[2 * t], q = rcol[2 * t + 1];
        double c = rcs[2 * t], sn = rcs[2 * t + 1];
        for (int r = r0; r < r0 + s; ++r) {
          double x = Q[(size_t)p * n + r], y = Q[(size_t)q * n + r];
          Q[(size_t)p * n + r] = c * x + sn * y;
          Q[(size_t)q * n + r] = c * y - sn * x;
        }
      }
      double w2 = 0.0;
      for (int j = 0; j < k; ++j) w2 += w[j] * w[j];
      for (int i = 0; i < k; ++i) dc_secular_root(i, k, dl.data(), w.data(), rho, w2, &org[i], &tau[i]);
      for (int i = 0; i < k; ++i) zh[i] = dc_zhat(i, k, dl.data(), org.data(), tau.data(), w.data());
      // eigenvectors of the rank-one update, then Q_new = Q_old[:, gcol] * V
      std::vector<double> V((size_t)k * k);
      for (int j = 0; j < k; ++j) {
        double nr = 0.0;
        for (int i = 0; i < k; ++i) {
          double v = zh[i] / ((dl[i] - dl[org[j]]) - tau[j]);
          V[(size_t)j * k + i] = v;
          nr += v * v;
        }
        nr = 1.0 / sqrt(nr);
        for (int i = 0; i < k; ++i) V[(size_t)j * k + i] *= nr;
      }
      for (int j = 0; j < k; ++j) {
        for (int r = r0; r < r0 + s; ++r) {
          double acc = 0.0;
          for (int t = 0; t < k; ++t) acc += Q[(size_t)gcol[t] * n + r] * V[(size_t)j * k + t];
          Qn[(size_t)(r0 + j) * n + r] = acc;
        }
        Dn[r0 + j] = dl[org[j]] + tau[j];
      }
      for (int t = 0; t < nd; ++t) {
        for (int r = r0; r < r0 + s; ++r) Qn[(size_t)(r0 + k + t) * n + r] = Q[(size_t)dcol[t] * n + r];
        Dn[r0 + k + t] = dval[t];
      }
    }
    std::swap(Q, Qn);
    std::swap(D, Dn);
  }
  std::vector<int> p(n);
  std::iota(p.begin(), p.end(), 0);
  std::stable_sort(p.begin(), p.end(), [&](int a, int b) { return D[a] < D[b]; });
  for (int i = 0; i < n; ++i) {
    evals[i] = D[p[i]];
    for (int r = 0; r < n; ++r) Qout[(size_t)i * n + r] = Q[(size_t)p[i] * n + r];
  }
  return 0;
}
