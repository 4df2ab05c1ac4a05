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
// In bf16x3 mode L and R are split once per bond (itg_heff2_prepare) and psi
// and T3 once per apply into k-contiguous bf16 planes, and C1/C4 run on the
// split-operand GEMM (bf16x3.hip).
#include <hip/hip_runtime.h>

#include <mutex>

#include "bf16x3.h"
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

// The same pass for fixed Q x O (the chain's w = 5, d = 2: 20 x 20): one
// thread per row a' for WPASS_B consecutive b, T1(a', q, b) loaded and
// T3(a', o, b) stored along a' (256 consecutive rows per block), X read from
// LDS by broadcast. The sum over q runs in the order of heff_wpass_kernel.
constexpr int WPASS_B = 4;

template <typename T, int QC, int OC>
__global__ __launch_bounds__(256) void heff_wpass_fixed_kernel(int64_t chiL, int64_t chiR,
                                                               const T* __restrict__ T1,
                                                               const double* __restrict__ X,
                                                               T* __restrict__ T3) {
  __shared__ double xs[QC * OC];
  for (int e = threadIdx.x; e < QC * OC; e += 256) xs[e] = X[e];
  __syncthreads();
  const int64_t tiles_a = (chiL + 255) / 256;
  const int64_t a = (blockIdx.x % tiles_a) * 256 + threadIdx.x;
  const int64_t b0 = (blockIdx.x / tiles_a) * WPASS_B;
  if (a >= chiL) return;
  const int64_t b1 = b0 + WPASS_B < chiR ? b0 + WPASS_B : chiR;
  for (int64_t b = b0; b < b1; ++b) {
    const T* src = T1 + (int64_t)QC * chiL * b + a;
    double t[QC];
#pragma unroll
    for (int q = 0; q < QC; ++q) t[q] = (double)src[chiL * q];
    T* dst = T3 + (int64_t)OC * chiL * b + a;
    // a few columns of X at a time: unrolled whole, the 400 loads would be
    // hoisted out of the b loop and spill
#pragma unroll 2
    for (int o = 0; o < OC; ++o) {
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < QC; ++q) acc += t[q] * xs[q + QC * o];
      dst[chiL * o] = (T)acc;
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

namespace {

struct HeffDims {
  int64_t chiL, wl, wm, wr, d, chiR, Q, O, M1, N1, K1, Kp1, M4, N4, K4, Kp4;
  HeffDims(int64_t chiL_, int64_t wl_, int64_t wm_, int64_t wr_, int64_t d_, int64_t chiR_)
      : chiL(chiL_), wl(wl_), wm(wm_), wr(wr_), d(d_), chiR(chiR_) {
    Q = wl * d * d;
    O = d * d * wr;
    M1 = chiL * wl;
    N1 = d * d * chiR;
    K1 = chiL;
    Kp1 = itg_split_ldk(K1);
    M4 = chiL * d * d;
    N4 = chiR;
    K4 = wr * chiR;
    Kp4 = itg_split_ldk(K4);
  }
};

inline size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// prep: X (Q x O doubles) | L planes (M1 x Kp1) | R planes (chiR x Kp4)
struct Prep {
  double* X;
  uint16_t *Lh, *Ll, *Rh, *Rl;
};
size_t prep_bytes(const HeffDims& h) {
  return al256(h.Q * h.O * sizeof(double)) + 2 * al256(h.M1 * h.Kp1 * 2) +
         2 * al256(h.N4 * h.Kp4 * 2) + 256;
}
Prep carve_prep(const HeffDims& h, void* base) {
  char* p = (char*)(((uintptr_t)base + 255) & ~(uintptr_t)255);
  Prep r;
  r.X = (double*)p;
  p += al256(h.Q * h.O * sizeof(double));
  r.Lh = (uint16_t*)p;
  p += al256(h.M1 * h.Kp1 * 2);
  r.Ll = (uint16_t*)p;
  p += al256(h.M1 * h.Kp1 * 2);
  r.Rh = (uint16_t*)p;
  p += al256(h.N4 * h.Kp4 * 2);
  r.Rl = (uint16_t*)p;
  return r;
}

// work: T1 | T3 | psi planes (N1 x Kp1) | T3 planes (M4 x Kp4) | prep
struct Work {
  void *T1, *T3;
  uint16_t *Ph, *Pl, *Th, *Tl;
  void* prep;
};
size_t work_bytes(const HeffDims& h, size_t es) {
  return al256(h.M1 * h.N1 * es) + al256(h.M4 * h.K4 * es) + 2 * al256(h.N1 * h.Kp1 * 2) +
         2 * al256(h.M4 * h.Kp4 * 2) + prep_bytes(h) + 256;
}
Work carve_work(const HeffDims& h, size_t es, void* base) {
  char* p = (char*)(((uintptr_t)base + 255) & ~(uintptr_t)255);
  Work w;
  w.T1 = p;
  p += al256(h.M1 * h.N1 * es);
  w.T3 = p;
  p += al256(h.M4 * h.K4 * es);
  w.Ph = (uint16_t*)p;
  p += al256(h.N1 * h.Kp1 * 2);
  w.Pl = (uint16_t*)p;
  p += al256(h.N1 * h.Kp1 * 2);
  w.Th = (uint16_t*)p;
  p += al256(h.M4 * h.Kp4 * 2);
  w.Tl = (uint16_t*)p;
  p += al256(h.M4 * h.Kp4 * 2);
  w.prep = p;
  return w;
}

int check_heff(itg_ctx* ctx, int dtype, int mode, const HeffDims& h) {
  ITG_CHECK_DTYPE(ctx, dtype);
  if (mode < 0 || mode > 2) return itg_fail(ctx, ITG_EARG, "unknown mode %d", mode);
  if (h.chiL <= 0 || h.chiR <= 0 || h.wl <= 0 || h.wm <= 0 || h.wr <= 0 || h.d <= 0)
    return itg_fail(ctx, ITG_EDIM, "heff2: non-positive dimension");
  if (h.Q * 64 + h.Q * h.O > 8192)   // LDS budget of the W pass (64 KB)
    return itg_fail(ctx, ITG_EUNSUPPORTED, "heff2: MPO bond dimension too large for the W pass");
  return ITG_OK;
}

int prepare(itg_ctx* ctx, int dtype, int mode, const HeffDims& h, const void* L, const void* W1,
            const void* W2, const void* R, const Prep& P) {
  if (dtype == ITG_F64)
    heff_x_kernel<double><<<itg_blocks(h.Q * h.O, 256), 256, 0, ctx->stream>>>(
        h.wl, h.wm, h.wr, h.d, (const double*)W1, (const double*)W2, P.X);
  else
    heff_x_kernel<float><<<itg_blocks(h.Q * h.O, 256), 256, 0, ctx->stream>>>(
        h.wl, h.wm, h.wr, h.d, (const float*)W1, (const float*)W2, P.X);
  ITG_LAUNCH_CHECK(ctx);
  if (mode == ITG_MODE_BF16X3) {
    int rc = itg_split_bf16(ctx, dtype, L, h.M1, h.K1, 1, h.M1, h.Kp1, P.Lh, P.Ll);
    if (rc) return rc;
    rc = itg_split_bf16(ctx, dtype, R, h.N4, h.K4, 1, h.chiR, h.Kp4, P.Rh, P.Rl);
    if (rc) return rc;
  }
  return ITG_OK;
}

int apply(itg_ctx* ctx, int dtype, int mode, const HeffDims& h, const void* L, const void* R,
          const Prep& P, const void* psi, void* out, const Work& w) {
  int rc;
  const bool bf = mode == ITG_MODE_BF16X3;
  // bf16x3 keeps T1 and T3 in FP32: the split operands carry 16 mantissa bits
  const int idt = bf ? ITG_F32 : dtype;
  // C1: T1 = L(a'wl, a) . psi(a, s1 s2 b)
  if (bf) {
    rc = itg_split_bf16(ctx, dtype, psi, h.N1, h.K1, h.K1, 1, h.Kp1, w.Ph, w.Pl);
    if (rc) return rc;
    rc = itg_gemm_split(ctx, idt, h.M1, h.N1, h.K1, h.Kp1, 1.0, P.Lh, P.Ll, h.Kp1, w.Ph, w.Pl,
                        h.Kp1, 0.0, w.T1, h.M1);
  } else {
    rc = itg_gemm_internal(ctx, dtype, mode, 'N', 'N', h.M1, h.N1, h.K1, 1.0, L, h.M1, psi, h.K1,
                           0.0, w.T1, h.M1);
  }
  if (rc) return rc;
  // C23: W pass
  int64_t blocks = ((h.chiL + 63) / 64) * h.chiR;
  size_t lds = (size_t)(h.Q * 64 + h.Q * h.O) * sizeof(double);
  int tk = itg_tstart(ctx, ITG_K_WPASS);
  if (h.Q == 20 && h.O == 20) {
    const int64_t fb = ((h.chiL + 255) / 256) * ((h.chiR + WPASS_B - 1) / WPASS_B);
    if (idt == ITG_F64)
      heff_wpass_fixed_kernel<double, 20, 20><<<(unsigned)fb, 256, 0, ctx->stream>>>(
          h.chiL, h.chiR, (const double*)w.T1, P.X, (double*)w.T3);
    else
      heff_wpass_fixed_kernel<float, 20, 20><<<(unsigned)fb, 256, 0, ctx->stream>>>(
          h.chiL, h.chiR, (const float*)w.T1, P.X, (float*)w.T3);
  } else if (idt == ITG_F64) {
    heff_wpass_kernel<double><<<(unsigned)blocks, 256, lds, ctx->stream>>>(
        h.chiL, h.Q, h.O, (const double*)w.T1, P.X, (double*)w.T3);
  } else {
    heff_wpass_kernel<float><<<(unsigned)blocks, 256, lds, ctx->stream>>>(
        h.chiL, h.Q, h.O, (const float*)w.T1, P.X, (float*)w.T3);
  }
  ITG_LAUNCH_CHECK(ctx);
  itg_tstop(ctx, tk, ITG_K_WPASS, 2.0 * h.Q * h.O * h.chiL * h.chiR,
            (double)(h.Q + h.O) * h.chiL * h.chiR * itg_dtype_size(idt));
  // C4: out(a's1's2', b') = T3(a's1's2', wr b) . R(b', wr b)^T
  if (bf) {
    rc = itg_split_bf16(ctx, idt, w.T3, h.M4, h.K4, 1, h.M4, h.Kp4, w.Th, w.Tl);
    if (rc) return rc;
    return itg_gemm_split(ctx, dtype, h.M4, h.N4, h.K4, h.Kp4, 1.0, w.Th, w.Tl, h.Kp4, P.Rh,
                          P.Rl, h.Kp4, 0.0, out, h.M4);
  }
  return itg_gemm_internal(ctx, dtype, mode, 'N', 'T', h.M4, h.N4, h.K4, 1.0, w.T3, h.M4, R,
                           h.chiR, 0.0, out, h.M4);
}

}  // namespace

extern "C" int64_t itg_heff2_work_elems(int64_t chiL, int64_t wl, int64_t wm, int64_t wr,
                                        int64_t d, int64_t chiR) {
  // counted in 4-byte units: enough for FP32 and FP64 buffers alike
  HeffDims h(chiL, wl, wm, wr, d, chiR);
  return (int64_t)((work_bytes(h, 8) + 3) / 4);
}

extern "C" int64_t itg_heff2_prep_elems(int64_t chiL, int64_t wl, int64_t wm, int64_t wr,
                                        int64_t d, int64_t chiR) {
  HeffDims h(chiL, wl, wm, wr, d, chiR);
  return (int64_t)((prep_bytes(h) + 3) / 4);
}

extern "C" int itg_heff2_prepare(itg_ctx* ctx, int dtype, int mode, int64_t chiL, int64_t wl,
                                 int64_t wm, int64_t wr, int64_t d, int64_t chiR, const void* L,
                                 const void* W1, const void* W2, const void* R, void* prep) {
  ITG_CHECK_CTX(ctx);
  HeffDims h(chiL, wl, wm, wr, d, chiR);
  int rc = check_heff(ctx, dtype, mode, h);
  if (rc) return rc;
  if (!L || !W1 || !W2 || !R || !prep) return itg_fail(ctx, ITG_EARG, "heff2: null pointer");
  return prepare(ctx, dtype, mode, h, L, W1, W2, R, carve_prep(h, prep));
}

extern "C" int itg_heff2_apply_prepared(itg_ctx* ctx, int dtype, int mode, int64_t chiL,
                                        int64_t wl, int64_t wm, int64_t wr, int64_t d,
                                        int64_t chiR, const void* L, const void* R,
                                        const void* prep, const void* psi, void* out,
                                        void* work) {
  ITG_CHECK_CTX(ctx);
  HeffDims h(chiL, wl, wm, wr, d, chiR);
  int rc = check_heff(ctx, dtype, mode, h);
  if (rc) return rc;
  if (!L || !R || !prep || !psi || !out || !work)
    return itg_fail(ctx, ITG_EARG, "heff2: null pointer");
  size_t es = itg_dtype_size(dtype);
  return apply(ctx, dtype, mode, h, L, R, carve_prep(h, (void*)prep), psi, out,
               carve_work(h, es, work));
}

int itg_heff2_apply_internal(itg_ctx* ctx, int dtype, int mode, const int64_t* dims,
                             const void* L, const void* R, const void* prep, const void* psi,
                             void* out, void* work) {
  HeffDims h(dims[0], dims[1], dims[2], dims[3], dims[4], dims[5]);
  int rc = check_heff(ctx, dtype, mode, h);
  if (rc) return rc;
  size_t es = itg_dtype_size(dtype);
  return apply(ctx, dtype, mode, h, L, R, carve_prep(h, (void*)prep), psi, out,
               carve_work(h, es, work));
}

extern "C" int itg_heff2_apply(itg_ctx* ctx, int dtype, int mode, int64_t chiL, int64_t wl,
                               int64_t wm, int64_t wr, int64_t d, int64_t chiR, const void* L,
                               const void* W1, const void* W2, const void* R, const void* psi,
                               void* out, void* work) {
  ITG_CHECK_CTX(ctx);
  HeffDims h(chiL, wl, wm, wr, d, chiR);
  int rc = check_heff(ctx, dtype, mode, h);
  if (rc) return rc;
  if (!L || !W1 || !W2 || !R || !psi || !out || !work)
    return itg_fail(ctx, ITG_EARG, "heff2: null pointer");
  size_t es = itg_dtype_size(dtype);
  Work w = carve_work(h, es, work);
  Prep P = carve_prep(h, w.prep);
  rc = prepare(ctx, dtype, mode, h, L, W1, W2, R, P);
  if (rc) return rc;
  return apply(ctx, dtype, mode, h, L, R, P, psi, out, w);
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

// ------------------------------------------------------ no-alloc variant
// itg_heff2_apply_noalloc keeps its work buffer (T1, T3, the split planes and
// the prepared operator) and any GEMM split-K partials in `scratch`, a
// file-scope device array: the HIP runtime gives every GPU its own copy when
// the library's code object is loaded there, and it is held until the
// process ends. In host code `scratch` names a host-side stand-in, not device
// memory, so host code never passes it to a kernel or a copy:
// itg_noalloc_setup resolves each device's address once, with that device
// current, and every later use goes through that address.
constexpr size_t kNoallocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {
constexpr int kNoallocMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kNoallocMaxDevices] = {};

// fixed split-K scratch and no event timing for the duration of one call
struct NoallocScope {
  itg_ctx* ctx;
  int timing;
  NoallocScope(itg_ctx* c, void* ws, size_t bytes) : ctx(c), timing(c->timing) {
    ctx->fixed_ws = ws;
    ctx->fixed_ws_bytes = bytes;
    ctx->timing = 0;
  }
  ~NoallocScope() {
    ctx->fixed_ws = nullptr;
    ctx->fixed_ws_bytes = 0;
    ctx->timing = timing;
  }
};
}  // namespace

int itg_noalloc_setup(itg_ctx* ctx) {
  if (ctx->device < 0 || ctx->device >= kNoallocMaxDevices)
    return itg_fail(ctx, ITG_EUNSUPPORTED, "no-alloc scratch: device %d out of range", ctx->device);
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  void* p = g_noalloc_ptr[ctx->device];
  if (!p) {
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return itg_fail(ctx, ITG_EHIP, "no-alloc scratch address on device %d: %s", ctx->device,
                      hipGetErrorString(e));
    g_noalloc_ptr[ctx->device] = p;
  }
  ctx->noalloc_scratch = p;
  return ITG_OK;
}

extern "C" int64_t itg_heff2_noalloc_max_bytes(void) { return (int64_t)kNoallocScratchBytes; }

extern "C" int itg_heff2_apply_noalloc(itg_ctx* ctx, int dtype, int mode, int64_t chiL,
                                       int64_t wl, int64_t wm, int64_t wr, int64_t d,
                                       int64_t chiR, const void* L, const void* W1,
                                       const void* W2, const void* R, const void* psi,
                                       void* out) {
  ITG_CHECK_CTX(ctx);
  HeffDims h(chiL, wl, wm, wr, d, chiR);
  int rc = check_heff(ctx, dtype, mode, h);
  if (rc) return rc;
  if (!L || !W1 || !W2 || !R || !psi || !out)
    return itg_fail(ctx, ITG_EARG, "heff2_apply_noalloc: null pointer");
  if (!ctx->noalloc_scratch)
    return itg_fail(ctx, ITG_EHIP, "heff2_apply_noalloc: no scratch on device %d", ctx->device);
  size_t es = itg_dtype_size(dtype);
  size_t need = al256(work_bytes(h, es));
  if (need > kNoallocScratchBytes)
    return itg_fail(ctx, ITG_EUNSUPPORTED,
                    "heff2_apply_noalloc: shape (%lld,%lld,%lld,%lld,%lld,%lld) needs %zu bytes "
                    "of work, more than the %zu-byte no-alloc scratch; use itg_heff2_apply "
                    "with a work buffer",
                    (long long)chiL, (long long)wl, (long long)wm, (long long)wr, (long long)d,
                    (long long)chiR, need, kNoallocScratchBytes);
  char* base = (char*)ctx->noalloc_scratch;
  // the rest of the array holds the split-K partials of the two GEMMs
  NoallocScope scope(ctx, base + need, kNoallocScratchBytes - need);
  Work w = carve_work(h, es, base);
  Prep P = carve_prep(h, w.prep);
  rc = prepare(ctx, dtype, mode, h, L, W1, W2, R, P);
  if (rc) return rc;
  return apply(ctx, dtype, mode, h, L, R, P, psi, out, w);
}
