// Two-site effective Hamiltonian product (ProjMPO product in ITensors:
// noprime(((psi*L)*W_b)*W_{b+1})*R, fixed order as in the reference's dmrg).
//
//   C1  T1(a',wl,s1,s2,b) = sum_a L(a',wl,a) psi(a,s1,s2,b)
//       GEMM M = chiL*wl, N = d*d*chiR, K = chiL               (MFMA, `mode`)
//   C23 T3(a',s1',s2',wr,b) = sum_{wl,s1,s2} T1(a',wl,s1,s2,b) X[(wl,s1,s2),(s1',s2',wr)]
//       X = W1.W2 contracted over wm; one HBM pass, FP64 arithmetic
//   C4  out(a',s1',s2',b') = sum_{wr,b} T3(a',s1',s2',wr,b) R(b',wr,b)
//       GEMM M = chiL*d*d, N = chiR, K = wr*chiR                (MFMA, `mode`)
//
// Algorithmic work at chiL = chiR = chi and wl = wr = w:
// 4*w*d^2*chi^3 (C1 + C4) + 2*w^2*d^4*chi^2 (C23).
#include <hip/hip_runtime.h>

#include "itg_internal.h"

// X[q][o], q = wl + WL*(s1 + d*s2), o = s1' + d*(s2' + d*wr)
template <typename T>
__global__ void heff_x_kernel(int64_t WL, int64_t WM, int64_t WR, int64_t d, const T* W1,
                              const T* W2, double* X) {
  int64_t Q = WL * d * d, O = d * d * WR;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < Q * O;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t q = e % Q, o = e / Q;
    int64_t wl = q % WL, s1 = (q / WL) % d, s2 = q / (WL * d);
    int64_t s1p = o % d, s2p = (o / d) % d, wr = o / (d * d);
    double acc = 0.0;
    for (int64_t wm = 0; wm < WM; ++wm) {
      double a = (double)W1[wl + WL * (wm + WM * (s1p + d * s1))];
      double b = (double)W2[wm + WM * (wr + WR * (s2p + d * s2))];
      acc += a * b;
    }
    X[q + Q * o] = acc;
  }
}

// one block: 64 consecutive a' for one b; tile T1(a', q) staged in LDS
template <typename T>
__global__ __launch_bounds__(256) void heff_wpass_kernel(int64_t chiL, int64_t Q, int64_t O,
                                                         const T* __restrict__ T1,
                                                         const double* __restrict__ X,
                                                         T* __restrict__ T3) {
  extern __shared__ double sh[];
  double* tile = sh;            // [Q][64]
  double* xs = sh + Q * 64;     // [Q][O]
  int64_t tiles_a = (chiL + 63) / 64;
  int64_t b = blockIdx.x / tiles_a;
  int64_t a0 = (blockIdx.x % tiles_a) * 64;
  int t = threadIdx.x;
  for (int64_t e = t; e < Q * O; e += 256) xs[e] = X[e];
  const T* src = T1 + Q * chiL * b;
  for (int64_t e = t; e < Q * 64; e += 256) {
    int64_t q = e / 64, al = e % 64;
    int64_t a = a0 + al;
    tile[e] = a < chiL ? (double)src[a + chiL * q] : 0.0;
  }
  __syncthreads();
  int al = t & 63, part = t >> 6;
  int64_t a = a0 + al;
  if (a < chiL) {
    T* dst = T3 + O * chiL * b;
    for (int64_t o = part; o < O; o += 4) {
      double acc = 0.0;
      for (int64_t q = 0; q < Q; ++q) acc += tile[q * 64 + al] * xs[q + Q * o];
      dst[a + chiL * o] = (T)acc;
    }
  }
}

template <typename T>
__global__ void heff_diag_kernel(int64_t chiL, int64_t WL, int64_t WM, int64_t WR, int64_t d,
                                 int64_t chiR, const T* L, const T* W1, const T* W2, const T* R,
                                 T* out) {
  int64_t n = chiL * d * d * chiR;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t a = e % chiL, s1 = (e / chiL) % d, s2 = (e / (chiL * d)) % d, b = e / (chiL * d * d);
    double acc = 0.0;
    for (int64_t wl = 0; wl < WL; ++wl) {
      double l = (double)L[a + chiL * (wl + WL * a)];
      if (l == 0.0) continue;
      for (int64_t wr = 0; wr < WR; ++wr) {
        double r = (double)R[b + chiR * (wr + WR * b)];
        if (r == 0.0) continue;
        double x = 0.0;
        for (int64_t wm = 0; wm < WM; ++wm)
          x += (double)W1[wl + WL * (wm + WM * (s1 + d * s1))] *
               (double)W2[wm + WM * (wr + WR * (s2 + d * s2))];
        acc += l * x * r;
      }
    }
    out[e] = (T)acc;
  }
}

extern "C" int64_t itg_heff2_work_elems(int64_t chiL, int64_t wl, int64_t wm, int64_t wr,
                                        int64_t d, int64_t chiR) {
  (void)wm;
  int64_t t1 = chiL * wl * d * d * chiR;
  int64_t t3 = chiL * d * d * wr * chiR;
  int64_t x = (wl * d * d) * (d * d * wr);
  // X is stored as doubles: reserve 2 elements per entry for FP32 storage
  return t1 + t3 + 2 * x + 64;
}

extern "C" int itg_heff2_apply(itg_ctx* ctx, int dtype, int mode, int64_t chiL, int64_t wl,
                               int64_t wm, int64_t wr, int64_t d, int64_t chiR, const void* L,
                               const void* W1, const void* W2, const void* R, const void* psi,
                               void* out, void* work) {
  ITG_CHECK_CTX(ctx);
  ITG_CHECK_DTYPE(ctx, dtype);
  if (mode < 0 || mode > 2) return itg_fail(ctx, ITG_EARG, "unknown mode %d", mode);
  if (chiL <= 0 || chiR <= 0 || wl <= 0 || wm <= 0 || wr <= 0 || d <= 0)
    return itg_fail(ctx, ITG_EDIM, "heff2: non-positive dimension");
  if (!L || !W1 || !W2 || !R || !psi || !out || !work)
    return itg_fail(ctx, ITG_EARG, "heff2: null pointer");
  size_t es = itg_dtype_size(dtype);
  int64_t Q = wl * d * d, O = d * d * wr;
  if (Q * 64 + Q * O > 8192)   // LDS budget of the W pass (64 KB)
    return itg_fail(ctx, ITG_EUNSUPPORTED, "heff2: MPO bond dimension too large for the W pass");
  char* wk = (char*)work;
  void* T1 = wk;
  void* T3 = wk + (size_t)chiL * Q * chiR * es;
  double* X = (double*)(wk + ((size_t)chiL * Q * chiR + (size_t)chiL * O * chiR) * es);
  X = (double*)(((uintptr_t)X + 7) & ~(uintptr_t)7);
  // X = W1.W2
  if (dtype == ITG_F64)
    heff_x_kernel<double><<<itg_blocks(Q * O, 256), 256, 0, ctx->stream>>>(
        wl, wm, wr, d, (const double*)W1, (const double*)W2, X);
  else
    heff_x_kernel<float><<<itg_blocks(Q * O, 256), 256, 0, ctx->stream>>>(
        wl, wm, wr, d, (const float*)W1, (const float*)W2, X);
  ITG_LAUNCH_CHECK(ctx);
  // C1: T1 = L(a'wl, a) . psi(a, s1 s2 b)
  int rc = itg_gemm_internal(ctx, dtype, mode, 'N', 'N', chiL * wl, d * d * chiR, chiL, 1.0, L,
                             chiL * wl, psi, chiL, 0.0, T1, chiL * wl);
  if (rc) return rc;
  // C23: W pass
  int64_t blocks = ((chiL + 63) / 64) * chiR;
  size_t lds = (size_t)(Q * 64 + Q * O) * sizeof(double);
  if (dtype == ITG_F64)
    heff_wpass_kernel<double><<<(unsigned)blocks, 256, lds, ctx->stream>>>(
        chiL, Q, O, (const double*)T1, X, (double*)T3);
  else
    heff_wpass_kernel<float><<<(unsigned)blocks, 256, lds, ctx->stream>>>(
        chiL, Q, O, (const float*)T1, X, (float*)T3);
  ITG_LAUNCH_CHECK(ctx);
  // C4: out(a's1's2', b') = T3(a's1's2', wr b) . R(b', wr b)^T
  return itg_gemm_internal(ctx, dtype, mode, 'N', 'T', chiL * d * d, chiR, wr * chiR, 1.0, T3,
                           chiL * d * d, R, chiR, 0.0, out, chiL * d * d);
}

extern "C" int itg_heff2_diag(itg_ctx* ctx, int dtype, int64_t chiL, int64_t wl, int64_t wm,
                              int64_t wr, int64_t d, int64_t chiR, const void* L, const void* W1,
                              const void* W2, const void* R, void* diag) {
  ITG_CHECK_CTX(ctx);
  ITG_CHECK_DTYPE(ctx, dtype);
  if (chiL <= 0 || chiR <= 0 || wl <= 0 || wm <= 0 || wr <= 0 || d <= 0)
    return itg_fail(ctx, ITG_EDIM, "heff2_diag: non-positive dimension");
  int64_t n = chiL * d * d * chiR;
  if (dtype == ITG_F64)
    heff_diag_kernel<double><<<itg_blocks(n, 256, 8192), 256, 0, ctx->stream>>>(
        chiL, wl, wm, wr, d, chiR, (const double*)L, (const double*)W1, (const double*)W2,
        (const double*)R, (double*)diag);
  else
    heff_diag_kernel<float><<<itg_blocks(n, 256, 8192), 256, 0, ctx->stream>>>(
        chiL, wl, wm, wr, d, chiR, (const float*)L, (const float*)W1, (const float*)W2,
        (const float*)R, (float*)diag);
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}
