// ComplexF64 storage: interleaved (re, im) pairs, Julia's ComplexF64 layout
// and NDTensors' Dense{ComplexF64} data. The data-movement and real-linear
// operations reuse the real kernels on a leading mode of dimension 2; the
// contraction is one real contraction over an expanded operand; the
// Hermitian eigendecomposition reduces to a real symmetric tridiagonal
// (zhetrd scheme), solved by the real divide and conquer, then transformed
// back with the complex reflectors (zunmtr scheme).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "itg_internal.h"

namespace {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// block sum of one complex value per thread (256 threads), result in every
// thread; fixed order, so bit-reproducible
__device__ __forceinline__ double2 block_csum_256(double2 v, double2* sh) {
  for (int o = 32; o > 0; o >>= 1) {
    v.x += __shfl_down(v.x, o, 64);
    v.y += __shfl_down(v.y, o, 64);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double2 r = make_double2(sh[0].x + sh[1].x + sh[2].x + sh[3].x,
                           sh[0].y + sh[1].y + sh[2].y + sh[3].y);
  return r;
}

__global__ void zfill_kernel(int64_t n, double re, double2* x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = make_double2(re, 0.0);
}

__global__ void zscal_kernel(int64_t n, double ar, double ai, double2* x) {
  const double2 a = make_double2(ar, ai);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = cmul(a, x[i]);
}

__global__ void zaxpby_kernel(int64_t n, double ar, double ai, const double2* x, double br,
                              double bi, double2* y) {
  const double2 a = make_double2(ar, ai), b = make_double2(br, bi);
  const bool useb = br != 0.0 || bi != 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double2 v = cmul(a, x[i]);
    if (useb) {
      double2 w = cmul(b, y[i]);
      v.x += w.x;
      v.y += w.y;
    }
    y[i] = v;
  }
}

__global__ void conj_kernel(int64_t n, double2* x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i].y = -x[i].y;
}

// real <-> complex storage
__global__ void to_complex_kernel(int64_t n, const double* s, double2* d) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = make_double2(s[i], 0.0);
}
__global__ void real_part_kernel(int64_t n, const double2* s, double* d) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = s[i].x;
}

__global__ __launch_bounds__(256) void zdot_partial_kernel(int64_t n, const double2* x,
                                                           const double2* y, double* part) {
  __shared__ double2 sh[4];
  double2 s = make_double2(0.0, 0.0);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double2 p = cmulc(x[i], y[i]);
    s.x += p.x;
    s.y += p.y;
  }
  double2 r = block_csum_256(s, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r.x;
    part[2 * blockIdx.x + 1] = r.y;
  }
}

__global__ __launch_bounds__(256) void zsum_parts_kernel(int nb, const double* part, double* out) {
  __shared__ double2 sh[4];
  double2 s = make_double2(0.0, 0.0);
  for (int i = threadIdx.x; i < nb; i += 256) {
    s.x += part[2 * i];
    s.y += part[2 * i + 1];
  }
  double2 r = block_csum_256(s, sh);
  if (threadIdx.x == 0) {
    out[0] = r.x;
    out[1] = r.y;
  }
}

// E(c, b, j) = sum_a M(a, b, c) X_a(j): the complex product table folded
// into one operand, E[4j + c + 2b] with c the product's and b the other
// operand's real/imaginary index: (re,re) = Xr, (im,re) = Xi, (re,im) = -Xi,
// (im,im) = Xr
__global__ void zexpand_kernel(int64_t n, const double2* __restrict__ X, double4* __restrict__ E) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double2 v = X[i];
    E[i] = make_double4(v.x, v.y, -v.y, v.x);
  }
}

// ------------------------------------------------ Hermitian tridiagonal
// Column j of the reduction (0-based, lower storage, A column-major ld n):
// reflector H = I - tau v v^H with v(0) = 1 and H^H x = beta e1, beta real,
// for x = A(j+1:n, j) (the zlarfg convention). v overwrites x.
__global__ __launch_bounds__(256) void zlarfg_kernel(double2* __restrict__ A, int n, int j,
                                                     double* d, double* e, double2* tau) {
  __shared__ double2 sh[4];
  double2* col = A + (size_t)j * n;
  double s = 0.0;
  for (int r = j + 2 + threadIdx.x; r < n; r += 256) {
    double2 v = col[r];
    s += v.x * v.x + v.y * v.y;
  }
  double xn2 = block_csum_256(make_double2(s, 0.0), sh).x;
  const double2 alpha = col[j + 1];
  __syncthreads();   // every thread has read alpha before thread 0 overwrites it
  double beta;
  double2 t;
  bool scale = false;
  double2 sc = make_double2(0.0, 0.0);
  if (xn2 == 0.0 && alpha.y == 0.0) {
    t = make_double2(0.0, 0.0);
    beta = alpha.x;
  } else {
    beta = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2), alpha.x);
    t = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
    // 1 / (alpha - beta)
    double2 dd = make_double2(alpha.x - beta, alpha.y);
    double den = dd.x * dd.x + dd.y * dd.y;
    sc = make_double2(dd.x / den, -dd.y / den);
    scale = true;
  }
  if (scale)
    for (int r = j + 2 + threadIdx.x; r < n; r += 256) col[r] = cmul(col[r], sc);
  if (threadIdx.x == 0) {
    d[j] = col[j].x;
    e[j] = beta;
    tau[j] = t;
    col[j + 1] = make_double2(1.0, 0.0);
  }
}

// partial y = A22 v over a 64-row x 128-column tile: part[cb * m + i]
__global__ __launch_bounds__(256) void zhemv_part_kernel(const double2* __restrict__ A, int n,
                                                         int j, double2* __restrict__ part) {
  __shared__ double2 sh[4][64];
  const int m = n - j - 1;
  const int tr = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + tr;
  const int c0 = blockIdx.y * 128, c1 = std::min(c0 + 128, m);
  const double2* v = A + (size_t)j * n + j + 1;
  double2 acc = make_double2(0.0, 0.0);
  if (i < m) {
    const double2* a = A + (size_t)(j + 1) * n + j + 1 + i;
    for (int c = c0 + g; c < c1; c += 4) {
      double2 p = cmul(a[(size_t)c * n], v[c]);
      acc.x += p.x;
      acc.y += p.y;
    }
  }
  sh[g][tr] = acc;
  __syncthreads();
  if (g == 0 && i < m) {
    double2 r = sh[0][tr];
    for (int q = 1; q < 4; ++q) { r.x += sh[q][tr].x; r.y += sh[q][tr].y; }
    part[(size_t)blockIdx.y * m + i] = r;
  }
}

// x = tau * sum(part); block partials of x^H v
__global__ __launch_bounds__(256) void zhemv_fin_kernel(const double2* __restrict__ A, int n, int j,
                                                        const double2* __restrict__ part, int ncb,
                                                        const double2* tau,
                                                        double2* __restrict__ x,
                                                        double2* __restrict__ dpart) {
  __shared__ double2 sh[4];
  const int m = n - j - 1;
  const int i = blockIdx.x * 256 + threadIdx.x;
  double2 c = make_double2(0.0, 0.0);
  if (i < m) {
    double2 y = make_double2(0.0, 0.0);
    for (int q = 0; q < ncb; ++q) {
      double2 p = part[(size_t)q * m + i];
      y.x += p.x;
      y.y += p.y;
    }
    double2 xi = cmul(tau[j], y);
    x[i] = xi;
    c = cmulc(xi, A[(size_t)j * n + j + 1 + i]);
  }
  double2 r = block_csum_256(c, sh);
  if (threadIdx.x == 0) dpart[blockIdx.x] = r;
}

// A22 -= v w^H + w v^H, w = x + alpha v, alpha = -tau (x^H v) / 2
__global__ __launch_bounds__(256) void zher2_kernel(double2* __restrict__ A, int n, int j,
                                                    const double2* __restrict__ x,
                                                    const double2* __restrict__ dpart, int ndp,
                                                    const double2* tau) {
  const int m = n - j - 1;
  double2 s = make_double2(0.0, 0.0);
  for (int q = 0; q < ndp; ++q) { s.x += dpart[q].x; s.y += dpart[q].y; }
  double2 al = cmul(tau[j], s);
  al.x *= -0.5;
  al.y *= -0.5;
  const double2* v = A + (size_t)j * n + j + 1;
  const int tr = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int r = blockIdx.x * 64 + tr;
  if (r >= m) return;
  const double2 vr = v[r];
  double2 wr = cmul(al, vr);
  wr.x += x[r].x;
  wr.y += x[r].y;
  double2* a = A + (size_t)(j + 1) * n + j + 1 + r;
  const int c0 = blockIdx.y * 64, c1 = std::min(c0 + 64, m);
  for (int c = c0 + g; c < c1; c += 4) {
    const double2 vc = v[c];
    double2 wc = cmul(al, vc);
    wc.x += x[c].x;
    wc.y += x[c].y;
    // v_r conj(w_c) + w_r conj(v_c)
    double2 p = cmulc(wc, vr), q = cmulc(vc, wr);
    double2 cur = a[(size_t)c * n];
    cur.x -= p.x + q.x;
    cur.y -= p.y + q.y;
    a[(size_t)c * n] = cur;
  }
}

__global__ void zhetrd_last_kernel(const double2* A, int n, double* d, double* e) {
  d[n - 1] = A[(size_t)(n - 1) * n + n - 1].x;
  e[n - 1] = 0.0;
}

// U(:, c) <- H_{j0} ... H_{j1-1} U(:, c) for one column c per block, the
// column's rows j0+1..n-1 held in LDS while the group's reflectors (columns
// j of A, v(j+1) = 1) are applied last to first
__global__ __launch_bounds__(256) void zunm_group_kernel(const double2* __restrict__ A, int n,
                                                         int j0, int j1, const double2* tau,
                                                         double2* __restrict__ U) {
  extern __shared__ double2 u[];
  __shared__ double2 sh[4];
  const int rlo = j0 + 1, mr = n - rlo;
  double2* col = U + (size_t)blockIdx.x * n + rlo;
  for (int r = threadIdx.x; r < mr; r += 256) u[r] = col[r];
  __syncthreads();
  for (int j = j1 - 1; j >= j0; --j) {
    const double2* v = A + (size_t)j * n;
    double2 s = make_double2(0.0, 0.0);
    for (int r = j + 1 + threadIdx.x; r < n; r += 256) {
      double2 p = cmulc(v[r], u[r - rlo]);
      s.x += p.x;
      s.y += p.y;
    }
    double2 t = cmul(tau[j], block_csum_256(s, sh));
    for (int r = j + 1 + threadIdx.x; r < n; r += 256) {
      double2 p = cmul(v[r], t);
      u[r - rlo].x -= p.x;
      u[r - rlo].y -= p.y;
    }
    __syncthreads();
  }
  for (int r = threadIdx.x; r < mr; r += 256) col[r] = u[r];
}

}  // namespace

// ---------------------------------------------------------- BLAS-1 / fills
int itg_c64_fill(itg_ctx* ctx, int64_t n, double value, void* x) {
  if (n <= 0) return ITG_OK;
  zfill_kernel<<<itg_blocks(n, 256), 256, 0, ctx->stream>>>(n, value, (double2*)x);
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}

int itg_c64_convert(itg_ctx* ctx, int64_t n, int sdt, const void* src, int ddt, void* dst) {
  if (n <= 0) return ITG_OK;
  unsigned b = itg_blocks(n, 256);
  if (sdt == ITG_F64 && ddt == ITG_C64)
    to_complex_kernel<<<b, 256, 0, ctx->stream>>>(n, (const double*)src, (double2*)dst);
  else if (sdt == ITG_C64 && ddt == ITG_F64)
    real_part_kernel<<<b, 256, 0, ctx->stream>>>(n, (const double2*)src, (double*)dst);
  else
    return itg_fail(ctx, ITG_EDTYPE, "convert: dtype %d -> %d", sdt, ddt);
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}

extern "C" int itg_zscal(itg_ctx* ctx, int64_t n, double re, double im, void* x) {
  ITG_CHECK_CTX(ctx);
  if (n <= 0 || (re == 1.0 && im == 0.0)) return ITG_OK;
  zscal_kernel<<<itg_blocks(n, 256, 4096), 256, 0, ctx->stream>>>(n, re, im, (double2*)x);
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}

extern "C" int itg_zaxpby(itg_ctx* ctx, int64_t n, double are, double aim, const void* x,
                          double bre, double bim, void* y) {
  ITG_CHECK_CTX(ctx);
  if (n <= 0) return ITG_OK;
  zaxpby_kernel<<<itg_blocks(n, 256, 4096), 256, 0, ctx->stream>>>(
      n, are, aim, (const double2*)x, bre, bim, (double2*)y);
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}

extern "C" int itg_conj(itg_ctx* ctx, int dtype, int64_t n, void* x) {
  ITG_CHECK_CTX(ctx);
  if (dtype == ITG_F64 || dtype == ITG_F32) return ITG_OK;
  if (dtype != ITG_C64) return itg_fail(ctx, ITG_EDTYPE, "unsupported dtype %d", dtype);
  if (n <= 0) return ITG_OK;
  conj_kernel<<<itg_blocks(n, 256, 4096), 256, 0, ctx->stream>>>(n, (double2*)x);
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}

int itg_c64_dot(itg_ctx* ctx, int64_t n, const void* x, const void* y, double* out) {
  if (n <= 0) {
    out[0] = out[1] = 0.0;
    return ITG_OK;
  }
  int nb = (int)std::min<int64_t>((n + 255) / 256, 2048);
  double* part = ctx->red + 16;
  zdot_partial_kernel<<<nb, 256, 0, ctx->stream>>>(n, (const double2*)x, (const double2*)y, part);
  zsum_parts_kernel<<<1, 256, 0, ctx->stream>>>(nb, part, ctx->red);
  ITG_LAUNCH_CHECK(ctx);
  ITG_HIP(ctx, hipMemcpyAsync(ctx->red_host, ctx->red, 2 * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
  ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  out[0] = ctx->red_host[0];
  out[1] = ctx->red_host[1];
  return ITG_OK;
}

// ------------------------------------------------------------ contraction
static int cplx_workspace(itg_ctx* ctx, size_t bytes, void** out) {
  if (bytes > ctx->cws_bytes) {
    if (ctx->cws) {
      ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
      ITG_HIP(ctx, hipFree(ctx->cws));
      ctx->cws = nullptr;
      ctx->cws_bytes = 0;
    }
    if (hipMalloc(&ctx->cws, bytes) != hipSuccess) {
      hipGetLastError();
      return itg_fail(ctx, ITG_EOOM, "complex workspace of %zu bytes failed", bytes);
    }
    ctx->cws_bytes = bytes;
  }
  *out = ctx->cws;
  return ITG_OK;
}

int itg_c64_contract(itg_ctx* ctx, int mode, int rA, const int64_t* dA, const int32_t* lA,
                     const void* A, int rB, const int64_t* dB, const int32_t* lB, const void* B,
                     int rC, const int64_t* dC, const int32_t* lC, void* C, double alpha,
                     double beta) {
  if (mode != ITG_MODE_NATIVE)
    return itg_fail(ctx, ITG_EUNSUPPORTED, "complex contraction runs in FP64 (native) mode only");
  if (rA < 0 || rB < 0 || rC < 0 || rA > ITG_MAX_RANK || rB > ITG_MAX_RANK || rC > ITG_MAX_RANK)
    return itg_fail(ctx, ITG_EDIM, "rank out of range");
  int64_t nA = 1, nB = 1;
  int32_t lmax = 0;
  for (int i = 0; i < rA; ++i) { nA *= dA[i]; lmax = std::max(lmax, lA[i]); }
  for (int i = 0; i < rB; ++i) { nB *= dB[i]; lmax = std::max(lmax, lB[i]); }
  for (int i = 0; i < rC; ++i) lmax = std::max(lmax, lC[i]);
  if (lmax > 0x7ffffff0) return itg_fail(ctx, ITG_EARG, "label out of range");
  // expand the smaller operand (it grows by 2x), if its rank leaves room
  bool expA = nA < nB;
  if (expA && rA + 2 > ITG_MAX_RANK) expA = false;
  if (!expA && rB + 2 > ITG_MAX_RANK) expA = true;
  if ((expA ? rA + 2 : rA + 1) > ITG_MAX_RANK || (expA ? rB + 1 : rB + 2) > ITG_MAX_RANK ||
      rC + 1 > ITG_MAX_RANK)
    return itg_fail(ctx, ITG_EUNSUPPORTED, "complex contraction: rank above %d", ITG_MAX_RANK - 2);
  const int32_t la = lmax + 1, lc = lmax + 2;   // the factors' and the product's re/im modes
  int64_t dA2[ITG_MAX_RANK], dB2[ITG_MAX_RANK], dC2[ITG_MAX_RANK];
  int32_t lA2[ITG_MAX_RANK], lB2[ITG_MAX_RANK], lC2[ITG_MAX_RANK];
  auto build = [&](bool expand, int r, const int64_t* d, const int32_t* l, int64_t* d2,
                   int32_t* l2) {
    int o = 0;
    if (expand) { d2[o] = 2; l2[o++] = lc; }
    d2[o] = 2;
    l2[o++] = la;
    for (int i = 0; i < r; ++i) { d2[o] = d[i]; l2[o++] = l[i]; }
    return o;
  };
  int rA2 = build(expA, rA, dA, lA, dA2, lA2);
  int rB2 = build(!expA, rB, dB, lB, dB2, lB2);
  dC2[0] = 2;
  lC2[0] = lc;
  for (int i = 0; i < rC; ++i) { dC2[i + 1] = dC[i]; lC2[i + 1] = lC[i]; }
  const int64_t ne = expA ? nA : nB;
  void* E = nullptr;
  if (ne > 0) {
    int rc = cplx_workspace(ctx, (size_t)ne * 32, &E);
    if (rc) return rc;
    zexpand_kernel<<<itg_blocks(ne, 256, 16384), 256, 0, ctx->stream>>>(
        ne, (const double2*)(expA ? A : B), (double4*)E);
    ITG_LAUNCH_CHECK(ctx);
  }
  return itg_contract(ctx, ITG_F64, mode, rA2, dA2, lA2, expA ? E : A, rB2, dB2, lB2,
                      expA ? B : E, rC + 1, dC2, lC2, C, alpha, beta);
}

// ---------------------------------------------------- Hermitian eigen parts
size_t itg_zhetrd_scratch_doubles(int n) {
  const size_t ncb = (size_t)(n + 127) / 128;
  return 2 * (ncb * n + n + 64 + (size_t)n);
}

int itg_zhetrd(itg_ctx* ctx, int n, void* Av, double* d, double* e, void* tauv, double* scratch) {
  double2* A = (double2*)Av;
  double2* tau = (double2*)tauv;
  const size_t ncbmax = (size_t)(n + 127) / 128;
  double2* part = (double2*)scratch;
  double2* x = part + ncbmax * n;
  double2* dpart = x + n;
  for (int j = 0; j + 1 < n; ++j) {
    const int m = n - j - 1;
    zlarfg_kernel<<<1, 256, 0, ctx->stream>>>(A, n, j, d, e, tau);
    const int ncb = (m + 127) / 128;
    zhemv_part_kernel<<<dim3((m + 63) / 64, ncb), 256, 0, ctx->stream>>>(A, n, j, part);
    const int ndp = (m + 255) / 256;
    zhemv_fin_kernel<<<ndp, 256, 0, ctx->stream>>>(A, n, j, part, ncb, tau, x, dpart);
    zher2_kernel<<<dim3((m + 63) / 64, (m + 63) / 64), 256, 0, ctx->stream>>>(A, n, j, x, dpart,
                                                                             ndp, tau);
    ITG_LAUNCH_CHECK(ctx);
  }
  zhetrd_last_kernel<<<1, 1, 0, ctx->stream>>>(A, n, d, e);
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}

int itg_zunmtr(itg_ctx* ctx, int n, int k, const void* Av, const void* tauv, const double* Z,
               void* Uv) {
  if (k <= 0) return ITG_OK;
  if (n > 8192) return itg_fail(ctx, ITG_EUNSUPPORTED, "complex eigh: n = %d > 8192", n);
  const double2* A = (const double2*)Av;
  const double2* tau = (const double2*)tauv;
  double2* U = (double2*)Uv;
  to_complex_kernel<<<itg_blocks((int64_t)n * k, 256, 16384), 256, 0, ctx->stream>>>(
      (int64_t)n * k, Z, U);
  ITG_LAUNCH_CHECK(ctx);
  static bool attr = false;
  if (!attr) {
    ITG_HIP(ctx, hipFuncSetAttribute((const void*)zunm_group_kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     8192 * (int)sizeof(double2)));
    attr = true;
  }
  constexpr int G = 32;
  const int nrefl = n - 1;
  for (int j1 = nrefl; j1 > 0; j1 -= G) {
    const int j0 = std::max(0, j1 - G);
    const size_t lds = (size_t)(n - j0 - 1) * sizeof(double2);
    zunm_group_kernel<<<k, 256, lds, ctx->stream>>>(A, n, j0, j1, tau, U);
    ITG_LAUNCH_CHECK(ctx);
  }
  return ITG_OK;
}

// ==================================================================== QR
// Householder QR of a complex m x n matrix (zgeqr2): reflector j from
// column j, rows j..m-1 (v(j) = 1 implicit, beta real on the diagonal), and
// H_j^H applied to the later columns; Q = H_0 .. H_{r-1} [I; 0] formed by
// one workgroup per column of Q with the column held in LDS.
namespace {

__global__ __launch_bounds__(256) void zqr_reflect_kernel(double2* __restrict__ W, int m, int j,
                                                          double2* tau) {
  __shared__ double2 sh[4];
  double2* col = W + (size_t)j * m;
  double s = 0.0;
  for (int r = j + 1 + threadIdx.x; r < m; r += 256) {
    double2 v = col[r];
    s += v.x * v.x + v.y * v.y;
  }
  const double xn2 = block_csum_256(make_double2(s, 0.0), sh).x;
  const double2 alpha = col[j];
  __syncthreads();   // alpha read by every thread before it is overwritten
  double beta = alpha.x;
  double2 t = make_double2(0.0, 0.0), sc = make_double2(1.0, 0.0);
  if (!(xn2 == 0.0 && alpha.y == 0.0)) {
    beta = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2), alpha.x);
    t = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
    double2 dd = make_double2(alpha.x - beta, alpha.y);
    double den = dd.x * dd.x + dd.y * dd.y;
    sc = make_double2(dd.x / den, -dd.y / den);
  }
  for (int r = j + 1 + threadIdx.x; r < m; r += 256) col[r] = cmul(col[r], sc);
  if (threadIdx.x == 0) {
    col[j] = make_double2(beta, 0.0);
    tau[j] = t;
  }
}

// W(j:m, c) -= conj(tau_j) v (v^H W(j:m, c)) for c > j, one block per column
__global__ __launch_bounds__(256) void zqr_apply_kernel(double2* __restrict__ W, int m, int n,
                                                        int j, const double2* tau) {
  __shared__ double2 sh[4];
  const int c = j + 1 + blockIdx.x;
  if (c >= n) return;
  const double2 tj = tau[j];
  if (tj.x == 0.0 && tj.y == 0.0) return;
  const double2* v = W + (size_t)j * m;
  double2* col = W + (size_t)c * m;
  double2 s = make_double2(0.0, 0.0);
  for (int r = j + threadIdx.x; r < m; r += 256) {
    double2 p = r == j ? col[r] : cmulc(v[r], col[r]);
    s.x += p.x;
    s.y += p.y;
  }
  s = block_csum_256(s, sh);
  const double2 f = cmulc(tj, s);   // conj(tau) (v^H col)
  for (int r = j + threadIdx.x; r < m; r += 256) {
    double2 p = r == j ? f : cmul(v[r], f);
    col[r].x -= p.x;
    col[r].y -= p.y;
  }
}

// column c of Q = H_0 .. H_{r-1} e_c, the column in LDS
__global__ __launch_bounds__(256) void zqr_form_q_kernel(const double2* __restrict__ W, int m,
                                                         int r, const double2* tau,
                                                         double2* __restrict__ Q) {
  extern __shared__ double2 u[];
  __shared__ double2 sh[4];
  const int c = blockIdx.x;
  for (int i = threadIdx.x; i < m; i += 256) u[i] = make_double2(i == c ? 1.0 : 0.0, 0.0);
  __syncthreads();
  for (int j = std::min(c, r - 1); j >= 0; --j) {
    // H_j e_c vanishes above row j: reflectors past c leave the column as is
    const double2* v = W + (size_t)j * m;
    double2 s = make_double2(0.0, 0.0);
    for (int i = j + threadIdx.x; i < m; i += 256) {
      double2 p = i == j ? u[i] : cmulc(v[i], u[i]);
      s.x += p.x;
      s.y += p.y;
    }
    const double2 f = cmul(tau[j], block_csum_256(s, sh));
    for (int i = j + threadIdx.x; i < m; i += 256) {
      double2 p = i == j ? f : cmul(v[i], f);
      u[i].x -= p.x;
      u[i].y -= p.y;
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < m; i += 256) Q[(size_t)c * m + i] = u[i];
}

__global__ void zqr_extract_r_kernel(int m, int n, int r, const double2* __restrict__ W,
                                     double2* __restrict__ R) {
  int64_t tot = (int64_t)r * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    int i = (int)(e % r), c = (int)(e / r);
    R[e] = i <= c ? W[i + (size_t)c * m] : make_double2(0.0, 0.0);
  }
}

// ------------------------------------------------------- complex Jacobi
// One round of the one-sided Jacobi SVD on the columns of W (m rows): pair
// (p, q) of the round-robin schedule is rotated so that W(:,p)^H W(:,q) = 0,
// the same rotation applied to V's columns: with g = a^H b = |g| e^{i phi},
// a' = c a - s e^{-i phi} b, b' = s e^{i phi} a + c b (the real rotation of
// (a, e^{-i phi} b)).
__global__ __launch_bounds__(256) void zjac_round_kernel(double2* __restrict__ W, int m,
                                                         double2* __restrict__ V, int nv,
                                                         int ncols, int N, int rnd, double tol,
                                                         double small2, double* __restrict__ offs) {
  __shared__ double2 sh[4];
  const int i = blockIdx.x;
  int p, q;
  if (i == 0) { p = rnd; q = N - 1; }
  else { p = (rnd + i) % (N - 1); q = (rnd - i + (N - 1)) % (N - 1); }
  if (p >= ncols || q >= ncols) return;
  double2* a = W + (size_t)p * m;
  double2* b = W + (size_t)q * m;
  double sa = 0.0, sb = 0.0;
  double2 g = make_double2(0.0, 0.0);
  for (int r = threadIdx.x; r < m; r += 256) {
    double2 x = a[r], y = b[r];
    sa += x.x * x.x + x.y * x.y;
    sb += y.x * y.x + y.y * y.y;
    double2 pr = cmulc(x, y);
    g.x += pr.x;
    g.y += pr.y;
  }
  const double2 ab = block_csum_256(make_double2(sa, sb), sh);
  sa = ab.x;
  sb = ab.y;
  g = block_csum_256(g, sh);
  const double ga = sqrt(g.x * g.x + g.y * g.y);
  const double rel = (sa > small2 && sb > small2) ? ga / sqrt(sa * sb) : 0.0;
  if (threadIdx.x == 0) offs[i] = fmax(offs[i], rel);
  if (rel <= tol) return;
  const double2 ph = make_double2(g.x / ga, g.y / ga);    // e^{i phi}
  const double zeta = (sb - sa) / (2.0 * ga);
  const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
  const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
  const double2 sp = make_double2(s * ph.x, s * ph.y);      // s e^{i phi}
  for (int r = threadIdx.x; r < m; r += 256) {
    double2 x = a[r], y = b[r];
    double2 u1 = cmulc(sp, y);   // s e^{-i phi} y
    double2 u2 = cmul(sp, x);    // s e^{i phi} x
    a[r] = make_double2(c * x.x - u1.x, c * x.y - u1.y);
    b[r] = make_double2(u2.x + c * y.x, u2.y + c * y.y);
  }
  double2* va = V + (size_t)p * nv;
  double2* vb = V + (size_t)q * nv;
  for (int r = threadIdx.x; r < nv; r += 256) {
    double2 x = va[r], y = vb[r];
    double2 u1 = cmulc(sp, y);
    double2 u2 = cmul(sp, x);
    va[r] = make_double2(c * x.x - u1.x, c * x.y - u1.y);
    vb[r] = make_double2(u2.x + c * y.x, u2.y + c * y.y);
  }
}

__global__ __launch_bounds__(256) void zcol_norm_kernel(const double2* __restrict__ W, int m,
                                                        int ncols, double* __restrict__ sig) {
  __shared__ double2 sh[4];
  const int c = blockIdx.x;
  if (c >= ncols) return;
  double s = 0.0;
  for (int r = threadIdx.x; r < m; r += 256) {
    double2 x = W[(size_t)c * m + r];
    s += x.x * x.x + x.y * x.y;
  }
  s = block_csum_256(make_double2(s, 0.0), sh).x;
  if (threadIdx.x == 0) sig[c] = sqrt(s);
}

// kept column j: U(:, j) = W(:, perm[j]) / sigma_j (zero for sigma_j = 0),
// V(:, j) = V0(:, perm[j]); both optionally conjugated (the transposed case)
__global__ void zsvd_gather_kernel(int rows, int cols, int k, const double2* __restrict__ W,
                                   const double2* __restrict__ V, const int* __restrict__ perm,
                                   const double* __restrict__ sig, double2* __restrict__ Uo,
                                   double2* __restrict__ Vo, int conj_out) {
  const int64_t tot = (int64_t)std::max(rows, cols) * k;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)(e / std::max(rows, cols)), r = (int)(e % std::max(rows, cols));
    const int c = perm[j];
    const double inv = sig[j] > 0.0 ? 1.0 / sig[j] : 0.0;
    if (r < rows) {
      double2 x = W[(size_t)c * rows + r];
      x = make_double2(x.x * inv, x.y * inv);
      if (conj_out) x.y = -x.y;
      Uo[(size_t)j * rows + r] = x;
    }
    if (r < cols) {
      double2 x = V[(size_t)c * cols + r];
      if (conj_out) x.y = -x.y;
      Vo[(size_t)j * cols + r] = x;
    }
  }
}

__global__ void zconj_transpose_kernel(int m, int n, const double2* __restrict__ A,
                                       double2* __restrict__ B) {
  const int64_t tot = (int64_t)m * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % m), c = (int)(e / m);
    double2 x = A[e];
    B[c + (size_t)i * n] = make_double2(x.x, -x.y);
  }
}

}  // namespace

static int zws(itg_ctx* ctx, size_t bytes, void** out) {
  if (bytes > ctx->zla_ws_bytes) {
    if (ctx->zla_ws) {
      ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
      ITG_HIP(ctx, hipFree(ctx->zla_ws));
      ctx->zla_ws = nullptr;
      ctx->zla_ws_bytes = 0;
    }
    if (hipMalloc(&ctx->zla_ws, bytes) != hipSuccess) {
      hipGetLastError();
      return itg_fail(ctx, ITG_EOOM, "complex factorisation workspace of %zu bytes failed", bytes);
    }
    ctx->zla_ws_bytes = bytes;
  }
  *out = ctx->zla_ws;
  return ITG_OK;
}

int itg_c64_qr(itg_ctx* ctx, int64_t m64, int64_t n64, const void* A, void* Qo, void* Ro) {
  if (m64 > 8192) return itg_fail(ctx, ITG_EUNSUPPORTED, "complex qr: m = %lld > 8192",
                                  (long long)m64);
  const int m = (int)m64, n = (int)n64, r = std::min(m, n);
  void* ws;
  int rc = zws(ctx, ((size_t)m * n + r) * sizeof(double2) + 512, &ws);
  if (rc) return rc;
  double2* W = (double2*)ws;
  double2* tau = W + (size_t)m * n;
  ITG_HIP(ctx, hipMemcpyAsync(W, A, (size_t)m * n * sizeof(double2), hipMemcpyDeviceToDevice,
                              ctx->stream));
  for (int j = 0; j < r; ++j) {
    zqr_reflect_kernel<<<1, 256, 0, ctx->stream>>>(W, m, j, tau);
    if (j + 1 < n) zqr_apply_kernel<<<n - j - 1, 256, 0, ctx->stream>>>(W, m, n, j, tau);
    ITG_LAUNCH_CHECK(ctx);
  }
  static bool attr = false;
  if (!attr) {
    ITG_HIP(ctx, hipFuncSetAttribute((const void*)zqr_form_q_kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     8192 * (int)sizeof(double2)));
    attr = true;
  }
  zqr_form_q_kernel<<<r, 256, (size_t)m * sizeof(double2), ctx->stream>>>(W, m, r, tau,
                                                                          (double2*)Qo);
  zqr_extract_r_kernel<<<itg_blocks((int64_t)r * n, 256, 8192), 256, 0, ctx->stream>>>(
      m, n, r, W, (double2*)Ro);
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}

int itg_c64_svd_trunc(itg_ctx* ctx, int64_t m64, int64_t n64, void* Ain, void* Uo, double* So,
                      void* Vo, int64_t maxdim, int64_t mindim, double cutoff, int abs_cut,
                      int rel_cut, int64_t* kout, double* truncerr, double* docut) {
  if (std::min(m64, n64) > 8192 || m64 * n64 > ((int64_t)1 << 30))
    return itg_fail(ctx, ITG_EUNSUPPORTED, "complex svd: matrix too large");
  const int m = (int)m64, n = (int)n64;
  const bool tr = m < n;   // work on A^H so that rows >= cols
  const int rows = tr ? n : m, cols = tr ? m : n;
  const int N = cols + (cols & 1);
  const size_t szW = (size_t)rows * cols, szV = (size_t)cols * cols;
  void* ws;
  int rc = zws(ctx, (3 * szW + 3 * szV) * sizeof(double2) + (size_t)N * 24 + 4096, &ws);
  if (rc) return rc;
  double2* W = (double2*)ws;
  double2* W2 = W + szW;
  double2* Wc = W2 + szW;             // conj(W), for the Gram product
  double2* G = Wc + szW;
  double2* V = G + szV;
  double* wG = (double*)(V + szV);
  double* offs = wG + N;
  double* sig = offs + N;
  int* perm = (int*)(sig + N);
  // W = A (or A^H)
  if (tr)
    zconj_transpose_kernel<<<itg_blocks((int64_t)m * n, 256, 8192), 256, 0, ctx->stream>>>(
        m, n, (const double2*)Ain, W);
  else
    ITG_HIP(ctx, hipMemcpyAsync(W, Ain, szW * sizeof(double2), hipMemcpyDeviceToDevice,
                                ctx->stream));
  ITG_LAUNCH_CHECK(ctx);
  // preconditioning: G = W^H W = V0 diag V0^H, W <- W V0, V <- V0
  ITG_HIP(ctx, hipMemcpyAsync(Wc, W, szW * sizeof(double2), hipMemcpyDeviceToDevice,
                              ctx->stream));
  rc = itg_conj(ctx, ITG_C64, (int64_t)szW, Wc);
  if (rc) return rc;
  {
    const int64_t dW[2] = {rows, cols}, dG[2] = {cols, cols};
    const int32_t lWc[2] = {1, 2}, lW[2] = {1, 3}, lG[2] = {2, 3};
    rc = itg_contract(ctx, ITG_C64, ITG_MODE_NATIVE, 2, dW, lWc, Wc, 2, dW, lW, W, 2, dG, lG, G,
                      1.0, 0.0);
    if (rc) return rc;
    int64_t kg = 0;
    double teg = 0.0, dcg = 0.0;
    rc = itg_eigh_trunc(ctx, ITG_C64, cols, G, wG, V, 0, -1.0, 0, 1, &kg, &teg, &dcg);
    if (rc) return rc;
    const int32_t lA[2] = {1, 2}, lV[2] = {2, 3}, lO[2] = {1, 3};
    rc = itg_contract(ctx, ITG_C64, ITG_MODE_NATIVE, 2, dW, lA, W, 2, dG, lV, V, 2, dW, lO, W2,
                      1.0, 0.0);
    if (rc) return rc;
  }
  W = W2;
  const double tol = 4.0 * 1.1102230246251565e-16 * std::sqrt((double)rows);
  double fro = 0.0;
  rc = itg_nrm2(ctx, ITG_C64, (int64_t)szW, W, &fro);
  if (rc) return rc;
  const double small2 = (1.1102230246251565e-16 * fro) * (1.1102230246251565e-16 * fro);
  std::vector<double> hoffs(N / 2);
  bool conv = cols < 2;
  for (int sweep = 0; sweep < 60 && !conv; ++sweep) {
    ITG_HIP(ctx, hipMemsetAsync(offs, 0, (N / 2) * sizeof(double), ctx->stream));
    for (int rnd = 0; rnd < N - 1; ++rnd)
      zjac_round_kernel<<<N / 2, 256, 0, ctx->stream>>>(W, rows, V, cols, cols, N, rnd, tol,
                                                        small2, offs);
    ITG_LAUNCH_CHECK(ctx);
    ITG_HIP(ctx, hipMemcpyAsync(hoffs.data(), offs, (N / 2) * sizeof(double),
                                hipMemcpyDeviceToHost, ctx->stream));
    ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double mx = 0.0;
    for (double x : hoffs) mx = std::max(mx, x);
    conv = mx <= tol;
  }
  if (!conv) return itg_fail(ctx, ITG_ECONV, "complex svd: Jacobi sweeps did not converge");
  zcol_norm_kernel<<<cols, 256, 0, ctx->stream>>>(W, rows, cols, wG);
  ITG_LAUNCH_CHECK(ctx);
  // descending order (on the host: one read-back of cols values)
  std::vector<double> hs(cols);
  ITG_HIP(ctx, hipMemcpyAsync(hs.data(), wG, cols * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
  ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<int> hp(cols);
  for (int i = 0; i < cols; ++i) hp[i] = i;
  std::stable_sort(hp.begin(), hp.end(), [&](int a, int b) { return hs[a] > hs[b]; });
  std::vector<double> sorted(cols), P(cols);
  for (int i = 0; i < cols; ++i) {
    sorted[i] = hs[hp[i]];
    P[i] = sorted[i] * sorted[i];
  }
  int64_t k = 0;
  double te = 0.0, dc = 0.0;
  rc = itg_truncate(ctx, P.data(), cols, maxdim, mindim, cutoff, abs_cut, rel_cut,
                    ITG_TRUNC_NDTENSORS, &k, &te, &dc);
  if (rc) return rc;
  ITG_HIP(ctx, hipMemcpyAsync(sig, sorted.data(), cols * sizeof(double), hipMemcpyHostToDevice,
                              ctx->stream));
  ITG_HIP(ctx, hipMemcpyAsync(perm, hp.data(), cols * sizeof(int), hipMemcpyHostToDevice,
                              ctx->stream));
  // A = U S V^H; for A^H = U' S V'^H, A = V' S U'^H: U = V', V = U'
  double2* Ua = (double2*)(tr ? Vo : Uo);
  double2* Va = (double2*)(tr ? Uo : Vo);
  zsvd_gather_kernel<<<itg_blocks((int64_t)std::max(rows, cols) * k, 256, 8192), 256, 0,
                       ctx->stream>>>(rows, cols, (int)k, W, V, perm, sig, Ua, Va, 0);
  ITG_LAUNCH_CHECK(ctx);
  ITG_HIP(ctx, hipMemcpyAsync(So, sig, std::min(m, n) * sizeof(double), hipMemcpyDeviceToDevice,
                              ctx->stream));
  ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *kout = k;
  if (truncerr) *truncerr = te;
  if (docut) *docut = dc;
  return ITG_OK;
}
