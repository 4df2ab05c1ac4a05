// Davidson eigensolver for one bond's two-site effective Hamiltonian, on the
// device: ITensors-style Krylov/Davidson (the eigsolve dmrg runs per bond,
// test/dmrg.jl:27) with the FP64 Krylov basis, batched FP64 dot products
// (one read-back per Gram-Schmidt pass), device-side Ritz combinations and
// the small projected eigenproblem on the host (cyclic Jacobi). In split-BF16
// mode the prepared Heff apply (five launches) is captured once into a HIP
// graph per (dims, basis slot) and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "itg_internal.h"

namespace {

constexpr int kMaxBasis = 17;   // maxiter + 1 <= 17

// part[b * cnt + k] = sum over this block's range of V_k . x (V_k = V + k n)
__global__ __launch_bounds__(256) void multidot_kernel(int64_t n, const double* __restrict__ V,
                                                       int cnt, const double* __restrict__ x,
                                                       double* __restrict__ part) {
  __shared__ double sh[4][kMaxBasis];
  double s[kMaxBasis];
#pragma unroll
  for (int k = 0; k < kMaxBasis; ++k) s[k] = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double xi = x[i];
#pragma unroll
    for (int k = 0; k < kMaxBasis; ++k)
      if (k < cnt) s[k] += V[(size_t)k * n + i] * xi;
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < kMaxBasis; ++k) {
    if (k >= cnt) break;
    double v = s[k];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) sh[w][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < cnt)
    part[(size_t)blockIdx.x * cnt + threadIdx.x] =
        (sh[0][threadIdx.x] + sh[1][threadIdx.x]) + (sh[2][threadIdx.x] + sh[3][threadIdx.x]);
}

// out[k] = sum_b part[b][k]: wave w sums the dots k = w (mod 4), its lanes
// striding over the block partials (a fixed order, so every call agrees)
__global__ __launch_bounds__(256) void multidot_sum_kernel(int nb, int cnt,
                                                           const double* __restrict__ part,
                                                           double* __restrict__ out) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int k = wv; k < cnt; k += 4) {
    double s = 0.0;
    for (int b = lane; b < nb; b += 64) s += part[(size_t)b * cnt + k];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) out[k] = s;
  }
}

struct Coefs {
  double c[kMaxBasis];
};

// y = beta * y + sum_k c_k V_k  (V_k = V + k n)
__global__ void lincomb_kernel(int64_t n, const double* __restrict__ V, int cnt, Coefs c,
                               double beta, double* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double s = beta == 0.0 ? 0.0 : beta * y[i];
#pragma unroll
    for (int k = 0; k < kMaxBasis; ++k)
      if (k < cnt) s += c.c[k] * V[(size_t)k * n + i];
    y[i] = s;
  }
}

// dots of x against V_0..V_{cnt-1}, read back to the host
int multidot(itg_ctx* ctx, int64_t n, const double* V, int cnt, const double* x, double* host) {
  unsigned nb = itg_blocks(n, 256, 512);
  double* part = ctx->red + 32;
  multidot_kernel<<<nb, 256, 0, ctx->stream>>>(n, V, cnt, x, part);
  multidot_sum_kernel<<<1, 256, 0, ctx->stream>>>((int)nb, cnt, part, ctx->red);
  ITG_LAUNCH_CHECK(ctx);
  ITG_HIP(ctx, hipMemcpyAsync(ctx->red_host, ctx->red, cnt * sizeof(double),
                              hipMemcpyDeviceToHost, ctx->stream));
  ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::memcpy(host, ctx->red_host, cnt * sizeof(double));
  return ITG_OK;
}

int lincomb(itg_ctx* ctx, int64_t n, const double* V, int cnt, const double* c, double beta,
            double* y) {
  Coefs cf;
  for (int k = 0; k < kMaxBasis; ++k) cf.c[k] = k < cnt ? c[k] : 0.0;
  lincomb_kernel<<<itg_blocks(n, 256, 8192), 256, 0, ctx->stream>>>(n, V, cnt, cf, beta, y);
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}

// lowest eigenpair of the symmetric m x m matrix M (row-major) by cyclic
// Jacobi; u normalised with u[0] >= 0
void small_lowest(int m, const double* M, double* lam, double* u) {
  std::vector<double> a(M, M + m * m), v(m * m, 0.0);
  for (int i = 0; i < m; ++i) v[i * m + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j) {
        tot += a[i * m + j] * a[i * m + j];
        if (i != j) off += a[i * m + j] * a[i * m + j];
      }
    if (off <= 1e-32 * tot || off == 0.0) break;
    for (int p = 0; p < m - 1; ++p)
      for (int q = p + 1; q < m; ++q) {
        double apq = a[p * m + q];
        if (apq == 0.0) continue;
        double theta = (a[q * m + q] - a[p * m + p]) / (2.0 * apq);
        double t = std::copysign(1.0, theta) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < m; ++k) {
          double akp = a[k * m + p], akq = a[k * m + q];
          a[k * m + p] = c * akp - s * akq;
          a[k * m + q] = s * akp + c * akq;
        }
        for (int k = 0; k < m; ++k) {
          double apk = a[p * m + k], aqk = a[q * m + k];
          a[p * m + k] = c * apk - s * aqk;
          a[q * m + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < m; ++k) {
          double vkp = v[k * m + p], vkq = v[k * m + q];
          v[k * m + p] = c * vkp - s * vkq;
          v[k * m + q] = s * vkp + c * vkq;
        }
      }
  }
  int best = 0;
  for (int i = 1; i < m; ++i)
    if (a[i * m + i] < a[best * m + best]) best = i;
  *lam = a[best * m + best];
  double nrm = 0.0;
  for (int k = 0; k < m; ++k) nrm += v[k * m + best] * v[k * m + best];
  nrm = std::sqrt(nrm);
  double sg = v[best] < 0.0 ? -1.0 : 1.0;
  for (int k = 0; k < m; ++k) u[k] = sg * v[k * m + best] / nrm;
}

// Heff apply x -> y, through a captured graph in split-BF16 mode
int heff_apply_graph(itg_ctx* ctx, int dtype, int mode, const int64_t* dims, const void* L,
                     const void* R, const void* prep, const void* x, void* y, void* work) {
  if (mode != ITG_MODE_BF16X3 || ctx->timing)
    return itg_heff2_apply_internal(ctx, dtype, mode, dims, L, R, prep, x, y, work);
  char key[256];
  std::snprintf(key, sizeof key, "%d/%d/%lld,%lld,%lld,%lld,%lld,%lld/%p/%p/%p/%p", dtype, mode,
                (long long)dims[0], (long long)dims[1], (long long)dims[2], (long long)dims[3],
                (long long)dims[4], (long long)dims[5], prep, x, y, work);
  auto it = ctx->graphs.find(key);
  if (it == ctx->graphs.end()) {
    // first use: eager (it also runs any one-time set-up outside a capture)
    ctx->graphs[key] = nullptr;
    return itg_heff2_apply_internal(ctx, dtype, mode, dims, L, R, prep, x, y, work);
  }
  if (!it->second) {
    if (ctx->graphs.size() > 256) {
      for (auto& kv : ctx->graphs)
        if (kv.second) hipGraphExecDestroy(kv.second);
      ctx->graphs.clear();
      return itg_heff2_apply_internal(ctx, dtype, mode, dims, L, R, prep, x, y, work);
    }
    hipGraph_t g = nullptr;
    ITG_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    int rc = itg_heff2_apply_internal(ctx, dtype, mode, dims, L, R, prep, x, y, work);
    hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
    if (rc || ce != hipSuccess || !g) {
      if (g) hipGraphDestroy(g);
      hipGetLastError();
      return rc ? rc : itg_fail(ctx, ITG_EHIP, "heff graph capture failed");
    }
    hipGraphExec_t ge = nullptr;
    hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ie != hipSuccess) {
      hipGetLastError();
      return itg_fail(ctx, ITG_EHIP, "heff graph instantiate failed: %s", hipGetErrorString(ie));
    }
    it->second = ge;
  }
  ITG_HIP(ctx, hipGraphLaunch(it->second, ctx->stream));
  return ITG_OK;
}

}  // namespace

void itg_graph_cache_clear(itg_ctx* ctx) {
  for (auto& kv : ctx->graphs)
    if (kv.second) hipGraphExecDestroy(kv.second);
  ctx->graphs.clear();
}

extern "C" int64_t itg_davidson2_work_elems(int64_t chiL, int64_t d, int64_t chiR, int maxiter) {
  int64_t n = chiL * d * d * chiR;
  if (maxiter < 1) maxiter = 1;
  return (int64_t)(2 * (maxiter + 1) + 3) * n + 64;
}

// The learned Davidson preconditioner run inside the device loop: feature
// rows of the residual, the fused network, scatter, plus the skip step
// (0 none, 1 the residual, 2 the diagonal step), each skip at rms 1.
struct LearnedStep {
  const float* params = nullptr;
  int in = 0, hidden = 0, out = 0;
  int skip = 2;
  int net_steps = 0;        // Davidson steps per call that try the network (0: all)
  double floor = 1e-8;
  float* X = nullptr;       // rows x in
  float* Y = nullptr;       // rows x out
  float* work = nullptr;    // activations when the fused kernel does not fit
  double* tskip = nullptr;  // n doubles
};

static int learned_correction(itg_ctx* ctx, const LearnedStep& ls, int64_t chiL, int64_t d,
                              int64_t chiR, const double* q, const double* diag, double lam,
                              double* t) {
  const int64_t rows = chiL * chiR, n = rows * d * d;
  int rc = itg_mlp_features(ctx, chiL, d, chiR, q, diag, lam, ls.floor, 0, 1, rows, ls.X);
  if (rc) return rc;
  rc = itg_mlp_infer(ctx, (int)rows, ls.in, ls.hidden, ls.out, ls.params, ls.X, ls.Y, ls.work);
  if (rc) return rc;
  rc = itg_mlp_scatter(ctx, chiL, d, chiR, ls.Y, 0, rows, t);
  if (rc) return rc;
  if (ls.skip == 0) return ITG_OK;
  const double* sv = q;
  if (ls.skip == 2) {
    rc = itg_precond_diag(ctx, ITG_F64, n, q, diag, lam, ls.floor, ls.tskip);
    if (rc) return rc;
    sv = ls.tskip;
  }
  double h[kMaxBasis];
  rc = multidot(ctx, n, sv, 1, sv, h);
  if (rc) return rc;
  const double sn = std::sqrt(h[0]);
  double sc = sn > 0.0 ? std::sqrt((double)n) / sn : 0.0;
  return lincomb(ctx, n, sv, 1, &sc, 1.0, t);
}

// ITensors' davidson on one bond with the basis on the device. With a
// learned step, each correction is first taken with the network and kept
// only if it lowers the residual norm; otherwise the basis vector is dropped,
// the previous Ritz pair restored and the diagonal step taken (`reverted`
// counts these).
static int davidson2_core(itg_ctx* ctx, int dtype, int mode, int64_t chiL, int64_t wl, int64_t wm,
                          int64_t wr, int64_t d, int64_t chiR, const void* L, const void* R,
                          const void* prep, void* psi, int maxiter, double errgoal,
                          const void* diag, const LearnedStep* ls, void* work, void* kwork,
                          double* energy, int* matvecs, int* reverted) {
  ITG_CHECK_CTX(ctx);
  if (dtype != ITG_F64) return itg_fail(ctx, ITG_EDTYPE, "davidson2: FP64 vectors only");
  if (maxiter < 1 || maxiter + 1 > kMaxBasis)
    return itg_fail(ctx, ITG_EARG, "davidson2: maxiter %d outside [1, %d]", maxiter,
                    kMaxBasis - 1);
  if (!psi || !prep || !work || !kwork || !energy)
    return itg_fail(ctx, ITG_EARG, "davidson2: null pointer");
  const int64_t dims[6] = {chiL, wl, wm, wr, d, chiR};
  const int64_t n = chiL * d * d * chiR;
  if (n <= 0) return itg_fail(ctx, ITG_EDIM, "davidson2: empty vector");
  if (maxiter > n - 1) maxiter = (int)std::max<int64_t>(n - 1, 0);
  const bool learned = ls != nullptr && ls->params != nullptr;
  double* V = (double*)kwork;                      // basis, maxiter + 1 vectors
  double* AV = V + (size_t)(maxiter + 1) * n;       // Heff applied to the basis
  double* q = AV + (size_t)(maxiter + 1) * n;       // residual
  double* t = q + n;                                // correction
  double* phi = t + n;                              // Ritz vector
  double h[kMaxBasis], u[kMaxBasis];
  int rc;
  // V0 = psi / |psi|
  rc = multidot(ctx, n, (const double*)psi, 1, (const double*)psi, h);
  if (rc) return rc;
  double nrm = std::sqrt(h[0]);
  if (!(nrm > 0.0)) return itg_fail(ctx, ITG_EARG, "davidson2: zero start vector");
  double one = 1.0 / nrm;
  rc = lincomb(ctx, n, (const double*)psi, 1, &one, 0.0, V);
  if (rc) return rc;
  rc = heff_apply_graph(ctx, dtype, mode, dims, L, R, prep, V, AV, work);
  if (rc) return rc;
  int nmv = 1, m = 1, nrev = 0;
  rc = multidot(ctx, n, V, 1, AV, h);
  if (rc) return rc;
  double lam = h[0], last = NAN;
  std::vector<double> M(kMaxBasis * kMaxBasis, 0.0);
  M[0] = lam;
  // q = AV0 - lam V0
  {
    double c1 = 1.0, cl = -lam;
    rc = lincomb(ctx, n, AV, 1, &c1, 0.0, q);
    if (rc) return rc;
    rc = lincomb(ctx, n, V, 1, &cl, 1.0, q);
    if (rc) return rc;
  }
  u[0] = 1.0;
  rc = multidot(ctx, n, q, 1, q, h);
  if (rc) return rc;
  double qn = std::sqrt(h[0]);
  for (int ni = 1; ni <= maxiter; ++ni) {
    bool conv = (qn < errgoal && std::fabs(lam - last) < errgoal) ||
                qn < std::max(1e-12, errgoal * 1e-3);
    if (qn < 1e-20 || (conv && ni > 1)) break;
    last = lam;
    bool stepped = false;
    const bool try_net = learned && (ls->net_steps <= 0 || ni <= ls->net_steps);
    const int attempts = try_net ? 2 : 1;
    for (int att = 0; att < attempts && !stepped; ++att) {
      const bool net = try_net && att == 0;
      const bool guarded = net;   // the diagonal step is the fallback
      // correction
      if (net) rc = learned_correction(ctx, *ls, chiL, d, chiR, q, (const double*)diag, lam, t);
      else if (diag) rc = itg_precond_diag(ctx, ITG_F64, n, q, diag, lam, 1e-8, t);
      else rc = itg_memcpy_d2d(ctx, t, q, (size_t)n * sizeof(double));
      if (rc) return rc;
      // two Gram-Schmidt passes against the basis
      for (int pass = 0; pass < 2; ++pass) {
        rc = multidot(ctx, n, V, m, t, h);
        if (rc) return rc;
        for (int k = 0; k < m; ++k) h[k] = -h[k];
        rc = lincomb(ctx, n, V, m, h, 1.0, t);
        if (rc) return rc;
      }
      rc = multidot(ctx, n, t, 1, t, h);
      if (rc) return rc;
      double tn = std::sqrt(h[0]);
      if (tn < 1e-14 * std::max(1.0, qn)) {
        if (guarded) {
          ++nrev;
          continue;
        }
        break;
      }
      // the Ritz pair before this step, restored if the step is reverted
      const std::vector<double> Mprev(M.begin(), M.begin() + m * m);
      const double lamprev = lam;
      double uprev[kMaxBasis];
      for (int k = 0; k < m; ++k) uprev[k] = u[k];
      double sc = 1.0 / tn;
      rc = lincomb(ctx, n, t, 1, &sc, 0.0, V + (size_t)m * n);
      if (rc) return rc;
      rc = heff_apply_graph(ctx, dtype, mode, dims, L, R, prep, V + (size_t)m * n,
                            AV + (size_t)m * n, work);
      if (rc) return rc;
      ++nmv;
      rc = multidot(ctx, n, V, m + 1, AV + (size_t)m * n, h);
      if (rc) return rc;
      ++m;
      std::vector<double> Mn(m * m);
      for (int a = 0; a < m - 1; ++a)
        for (int b = 0; b < m - 1; ++b) Mn[a * m + b] = M[a * (m - 1) + b];
      for (int k = 0; k < m; ++k) Mn[k * m + (m - 1)] = Mn[(m - 1) * m + k] = h[k];
      M.assign(Mn.begin(), Mn.end());
      small_lowest(m, M.data(), &lam, u);
      // phi = sum u_k V_k, q = sum u_k AV_k - lam phi
      rc = lincomb(ctx, n, V, m, u, 0.0, phi);
      if (rc) return rc;
      rc = lincomb(ctx, n, AV, m, u, 0.0, q);
      if (rc) return rc;
      double ml = -lam;
      rc = lincomb(ctx, n, phi, 1, &ml, 1.0, q);
      if (rc) return rc;
      rc = multidot(ctx, n, q, 1, q, h);
      if (rc) return rc;
      const double qn_new = std::sqrt(h[0]);
      if (guarded && !(qn_new < qn)) {
        // the learned step did not lower the residual: drop it
        --m;
        M.assign(Mprev.begin(), Mprev.end());
        lam = lamprev;
        for (int k = 0; k < m; ++k) u[k] = uprev[k];
        rc = lincomb(ctx, n, V, m, u, 0.0, phi);
        if (rc) return rc;
        rc = lincomb(ctx, n, AV, m, u, 0.0, q);
        if (rc) return rc;
        ml = -lam;
        rc = lincomb(ctx, n, phi, 1, &ml, 1.0, q);
        if (rc) return rc;
        ++nrev;
        continue;
      }
      qn = qn_new;
      stepped = true;
    }
    if (!stepped) break;
  }
  // psi = Ritz vector (V0 when no step was taken), normalised
  const double* src = (m == 1) ? V : phi;
  rc = multidot(ctx, n, src, 1, src, h);
  if (rc) return rc;
  double inv = 1.0 / std::sqrt(h[0]);
  rc = lincomb(ctx, n, src, 1, &inv, 0.0, (double*)psi);
  if (rc) return rc;
  *energy = lam;
  if (matvecs) *matvecs = nmv;
  if (reverted) *reverted = nrev;
  return ITG_OK;
}

extern "C" int itg_davidson2(itg_ctx* ctx, int dtype, int mode, int64_t chiL, int64_t wl,
                             int64_t wm, int64_t wr, int64_t d, int64_t chiR, const void* L,
                             const void* R, const void* prep, void* psi, int maxiter,
                             double errgoal, const void* diag, void* work, void* kwork,
                             double* energy, int* matvecs) {
  return davidson2_core(ctx, dtype, mode, chiL, wl, wm, wr, d, chiR, L, R, prep, psi, maxiter,
                        errgoal, diag, nullptr, work, kwork, energy, matvecs, nullptr);
}

static bool mlp_fused(int in, int hidden, int out) {
  return hidden == 256 && ((in <= 16 && out <= 4) || (in <= 32 && out <= 9));
}

extern "C" int64_t itg_davidson2_mlp_work_elems(int64_t chiL, int64_t d, int64_t chiR,
                                                int hidden) {
  const int64_t rows = chiL * chiR, in = 3 * d * d, out = d * d;
  int64_t f = rows * in + rows * out + 2 * rows * d * d + 64;
  if (!mlp_fused((int)in, hidden, (int)out)) f += 2 * rows * (int64_t)hidden;
  return f;
}

extern "C" int itg_davidson2_mlp(itg_ctx* ctx, int dtype, int mode, int64_t chiL, int64_t wl,
                                 int64_t wm, int64_t wr, int64_t d, int64_t chiR, const void* L,
                                 const void* R, const void* prep, void* psi, int maxiter,
                                 double errgoal, const void* diag, const float* params,
                                 int hidden, int skip, int net_steps, double floor, void* mwork,
                                 void* work, void* kwork, double* energy, int* matvecs,
                                 int* reverted) {
  ITG_CHECK_CTX(ctx);
  if (!diag || !params || !mwork)
    return itg_fail(ctx, ITG_EARG, "davidson2_mlp: diagonal, parameters and work are required");
  if (skip < 0 || skip > 2) return itg_fail(ctx, ITG_EARG, "davidson2_mlp: skip %d", skip);
  if (!(floor > 0.0)) return itg_fail(ctx, ITG_EARG, "davidson2_mlp: floor must be > 0");
  const int64_t rows = chiL * chiR;
  if (rows <= 0 || rows > 0x7fffffff) return itg_fail(ctx, ITG_EDIM, "davidson2_mlp: rows");
  LearnedStep ls;
  ls.params = params;
  ls.in = (int)(3 * d * d);
  ls.hidden = hidden;
  ls.out = (int)(d * d);
  ls.skip = skip;
  ls.net_steps = net_steps;
  ls.floor = floor;
  // work: X | Y | tskip (n doubles, 8-byte aligned) | activations
  float* w = (float*)mwork;
  ls.X = w;
  ls.Y = ls.X + rows * ls.in;
  size_t off = (size_t)(rows * ls.in + rows * ls.out);
  off = (off + 1) & ~(size_t)1;
  ls.tskip = (double*)(w + off);
  off += 2 * (size_t)rows * d * d;
  ls.work = mlp_fused(ls.in, hidden, ls.out) ? nullptr : w + off;
  return davidson2_core(ctx, dtype, mode, chiL, wl, wm, wr, d, chiR, L, R, prep, psi, maxiter,
                        errgoal, diag, &ls, work, kwork, energy, matvecs, reverted);
}
