// libitg core: context, errors, memory, Philox Gaussian fill, BLAS-1,
// permutation / permuted add, truncation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>

#include "itg_internal.h"
#include "truncate_core.h"

// ================================================================ errors
int itg_fail(itg_ctx* ctx, int code, const char* fmt, ...) {
  if (ctx) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    ctx->err = buf;
  }
  return code;
}

extern "C" const char* itg_last_error(itg_ctx* ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

// =============================================================== context
extern "C" int itg_init(int device, itg_ctx** out) {
  if (!out) return ITG_EARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return ITG_EHIP;
  if (device < 0 || device >= ndev) return ITG_EARG;
  if (hipSetDevice(device) != hipSuccess) return ITG_EHIP;
  itg_ctx* ctx = new itg_ctx();
  ctx->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete ctx; return ITG_EHIP; }
  ctx->num_cu = prop.multiProcessorCount;
  ctx->gfx = 0;
  if (strncmp(prop.gcnArchName, "gfx", 3) == 0) ctx->gfx = (int)strtol(prop.gcnArchName + 3, nullptr, 16);
  ctx->hbm_gb = (double)prop.totalGlobalMem / 1e9;
  if (hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return ITG_EHIP;
  }
  ctx->stream = ctx->own_stream;
  const char* lib = getenv("ITG_GEMM_LIB");
  ctx->libgemm = (lib && lib[0] == '0') ? 0 : 1;
  if (hipMalloc(&ctx->red, kRedSlots * sizeof(double)) != hipSuccess ||
      hipHostMalloc(&ctx->red_host, kRedSlots * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return ITG_EOOM;
  }
  if (itg_noalloc_setup(ctx) != ITG_OK) {
    hipFree(ctx->red);
    hipHostFree(ctx->red_host);
    hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return ITG_EHIP;
  }
  *out = ctx;
  return ITG_OK;
}

extern "C" int itg_destroy(itg_ctx* ctx) {
  ITG_CHECK_CTX(ctx);
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  itg_plan_cache_clear(ctx);
  itg_graph_cache_clear(ctx);
  itg_timing_free(ctx);
  itg_libgemm_free(ctx);
  if (ctx->ws) hipFree(ctx->ws);
  if (ctx->eig_ws) hipFree(ctx->eig_ws);
  if (ctx->mlp_ws) hipFree(ctx->mlp_ws);
  if (ctx->cws) hipFree(ctx->cws);
  if (ctx->svd_ws) hipFree(ctx->svd_ws);
  if (ctx->zla_ws) hipFree(ctx->zla_ws);
  if (ctx->red) hipFree(ctx->red);
  if (ctx->red_host) hipHostFree(ctx->red_host);
  hipStreamDestroy(ctx->own_stream);
  delete ctx;
  return ITG_OK;
}

extern "C" int itg_sync(itg_ctx* ctx) {
  ITG_CHECK_CTX(ctx);
  ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ITG_OK;
}

extern "C" void* itg_stream(itg_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

extern "C" int itg_set_stream(itg_ctx* ctx, void* stream) {
  ITG_CHECK_CTX(ctx);
  ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->stream = stream ? (hipStream_t)stream : ctx->own_stream;
  return ITG_OK;
}

extern "C" int itg_device_info(itg_ctx* ctx, int* num_cu, int* gfx, double* hbm_gb) {
  ITG_CHECK_CTX(ctx);
  if (num_cu) *num_cu = ctx->num_cu;
  if (gfx) *gfx = ctx->gfx;
  if (hbm_gb) *hbm_gb = ctx->hbm_gb;
  return ITG_OK;
}

// ------------------------------------------------------------- timing
static hipEvent_t timing_event(itg_ctx* ctx) {
  if (ctx->event_next == ctx->event_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    ctx->event_pool.push_back(e);
  }
  return ctx->event_pool[ctx->event_next++];
}

int itg_tstart(itg_ctx* ctx, int kind) {
  if (!(ctx->timing & (1 << kind))) return -1;
  hipEvent_t a = timing_event(ctx), b = timing_event(ctx);
  if (!a || !b) return -1;
  hipEventRecord(a, ctx->stream);
  ctx->launches.push_back({-1, a, b, 0.0, 0.0});
  return (int)ctx->launches.size() - 1;
}

itg_tev itg_tevents(itg_ctx* ctx, int kind) {
  itg_tev ev = {-1, nullptr, nullptr};
  if (!(ctx->timing & (1 << kind))) return ev;
  hipEvent_t a = timing_event(ctx), b = timing_event(ctx);
  if (!a || !b) return ev;
  ctx->launches.push_back({-1, a, b, 0.0, 0.0});
  ev.t = (int)ctx->launches.size() - 1;
  ev.a = a;
  ev.b = b;
  return ev;
}

void itg_tcommit(itg_ctx* ctx, const itg_tev& ev, int kind, double flops, double bytes) {
  if (ev.t < 0) return;
  auto& L = ctx->launches[ev.t];
  L.kind = kind;
  L.flops = flops;
  L.bytes = bytes;
}

void itg_tstop(itg_ctx* ctx, int t, int kind, double flops, double bytes) {
  if (t < 0) return;
  auto& L = ctx->launches[t];
  hipEventRecord(L.b, ctx->stream);
  L.kind = kind;
  L.flops = flops;
  L.bytes = bytes;
}

void itg_timing_free(itg_ctx* ctx) {
  for (hipEvent_t e : ctx->event_pool) hipEventDestroy(e);
  ctx->event_pool.clear();
  ctx->launches.clear();
  ctx->event_next = 0;
}

extern "C" int itg_timing(itg_ctx* ctx, int mask) {
  ITG_CHECK_CTX(ctx);
  if (mask) {
    ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->launches.clear();
    ctx->event_next = 0;
  }
  ctx->timing = mask & ((1 << ITG_K_COUNT) - 1);
  return ITG_OK;
}

extern "C" int itg_timing_set(itg_ctx* ctx, int mask) {
  ITG_CHECK_CTX(ctx);
  ctx->timing = mask & ((1 << ITG_K_COUNT) - 1);
  return ITG_OK;
}

extern "C" int itg_timing_report(itg_ctx* ctx, int kind, int64_t* launches, double* ms,
                                 double* flops, double* bytes) {
  ITG_CHECK_CTX(ctx);
  if (kind < 0 || kind >= ITG_K_COUNT) return itg_fail(ctx, ITG_EARG, "timing: kind %d", kind);
  ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  int64_t n = 0;
  double t = 0.0, f = 0.0, b = 0.0;
  for (auto& L : ctx->launches) {
    if (L.kind != kind) continue;
    float e = 0.0f;
    ITG_HIP(ctx, hipEventElapsedTime(&e, L.a, L.b));
    ++n;
    t += e;
    f += L.flops;
    b += L.bytes;
  }
  if (launches) *launches = n;
  if (ms) *ms = t;
  if (flops) *flops = f;
  if (bytes) *bytes = b;
  return ITG_OK;
}

int itg_workspace(itg_ctx* ctx, size_t bytes, void** out) {
  if (bytes == 0) bytes = 256;
  if (bytes > ctx->ws_bytes) {
    if (ctx->ws) {
      ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
      ITG_HIP(ctx, hipFree(ctx->ws));
      ctx->ws = nullptr;
      ctx->ws_bytes = 0;
    }
    size_t want = std::max(bytes, (size_t)1 << 20);
    if (hipMalloc(&ctx->ws, want) != hipSuccess)
      return itg_fail(ctx, ITG_EOOM, "workspace allocation of %zu bytes failed", want);
    ctx->ws_bytes = want;
  }
  *out = ctx->ws;
  return ITG_OK;
}

// ================================================================ memory
extern "C" int itg_malloc(itg_ctx* ctx, size_t bytes, void** ptr) {
  ITG_CHECK_CTX(ctx);
  if (!ptr) return itg_fail(ctx, ITG_EARG, "null output pointer");
  *ptr = nullptr;
  if (bytes == 0) bytes = 8;
  if (hipMalloc(ptr, bytes) != hipSuccess) {
    hipGetLastError();
    return itg_fail(ctx, ITG_EOOM, "hipMalloc(%zu) failed", bytes);
  }
  return ITG_OK;
}

extern "C" int itg_free(itg_ctx* ctx, void* ptr) {
  ITG_CHECK_CTX(ctx);
  if (!ptr) return ITG_OK;
  ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ITG_HIP(ctx, hipFree(ptr));
  return ITG_OK;
}

extern "C" int itg_memcpy_h2d(itg_ctx* ctx, void* dst, const void* src, size_t bytes) {
  ITG_CHECK_CTX(ctx);
  if (!bytes) return ITG_OK;
  ITG_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ITG_OK;
}

extern "C" int itg_memcpy_d2h(itg_ctx* ctx, void* dst, const void* src, size_t bytes) {
  ITG_CHECK_CTX(ctx);
  if (!bytes) return ITG_OK;
  ITG_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ITG_OK;
}

extern "C" int itg_memcpy_d2d(itg_ctx* ctx, void* dst, const void* src, size_t bytes) {
  ITG_CHECK_CTX(ctx);
  if (!bytes) return ITG_OK;
  ITG_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return ITG_OK;
}

template <typename T>
__global__ void fill_kernel(int64_t n, T v, T* x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = v;
}

extern "C" int itg_fill(itg_ctx* ctx, int dtype, int64_t n, double value, void* x) {
  ITG_CHECK_CTX(ctx);
  if (dtype == ITG_C64) return itg_c64_fill(ctx, n, value, x);
  ITG_CHECK_DTYPE(ctx, dtype);
  if (n <= 0) return ITG_OK;
  if (dtype == ITG_F64)
    fill_kernel<double><<<itg_blocks(n, 256), 256, 0, ctx->stream>>>(n, value, (double*)x);
  else
    fill_kernel<float><<<itg_blocks(n, 256), 256, 0, ctx->stream>>>(n, (float)value, (float*)x);
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}

template <typename S, typename D>
__global__ void convert_kernel(int64_t n, const S* s, D* d) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = (D)s[i];
}

extern "C" int itg_convert(itg_ctx* ctx, int64_t n, int sdt, const void* src, int ddt, void* dst) {
  ITG_CHECK_CTX(ctx);
  if (sdt == ITG_C64 && ddt == ITG_C64) return itg_memcpy_d2d(ctx, dst, src, (size_t)n * 16);
  if (sdt == ITG_C64 || ddt == ITG_C64) return itg_c64_convert(ctx, n, sdt, src, ddt, dst);
  ITG_CHECK_DTYPE(ctx, sdt);
  ITG_CHECK_DTYPE(ctx, ddt);
  if (n <= 0) return ITG_OK;
  unsigned b = itg_blocks(n, 256);
  if (sdt == ddt) return itg_memcpy_d2d(ctx, dst, src, n * itg_dtype_size(sdt));
  if (sdt == ITG_F64)
    convert_kernel<double, float><<<b, 256, 0, ctx->stream>>>(n, (const double*)src, (float*)dst);
  else
    convert_kernel<float, double><<<b, 256, 0, ctx->stream>>>(n, (const float*)src, (double*)dst);
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}

// ================================================================== rng
// Philox4x32-10, Salmon et al. SC'11; constants from the paper.
__device__ __forceinline__ void philox_round(uint32_t& c0, uint32_t& c1, uint32_t& c2,
                                             uint32_t& c3, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  uint32_t hi0 = __umulhi(M0, c0), lo0 = M0 * c0;
  uint32_t hi1 = __umulhi(M1, c2), lo1 = M1 * c2;
  uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
  c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
}

template <typename T>
__global__ void randn_kernel(int64_t n, uint64_t seed, uint64_t offset, T* x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t g = offset + (uint64_t)i;
    uint64_t ctr = g >> 1;
    uint32_t c0 = (uint32_t)ctr, c1 = (uint32_t)(ctr >> 32), c2 = 0, c3 = 0;
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    for (int r = 0; r < 10; ++r) {
      if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
      philox_round(c0, c1, c2, c3, k0, k1);
    }
    uint64_t a = ((((uint64_t)c1) << 32) | c0) >> 11;
    uint64_t b = ((((uint64_t)c3) << 32) | c2) >> 11;
    double u1 = ((double)a + 0.5) * 0x1p-53;
    double u2 = ((double)b + 0.5) * 0x1p-53;
    double r = sqrt(-2.0 * log(u1));
    double th = 2.0 * M_PI * u2;
    double z = (g & 1) ? r * sin(th) : r * cos(th);
    x[i] = (T)z;
  }
}

extern "C" int itg_randn(itg_ctx* ctx, int dtype, int64_t n, uint64_t seed, uint64_t offset,
                         void* x) {
  ITG_CHECK_CTX(ctx);
  if (dtype == ITG_C64) {
    // complex element i = (g[2(offset+i)] + i g[2(offset+i)+1]) / sqrt(2), g
    // the real stream: unit variance, as randn(ComplexF64)
    int rc = itg_randn(ctx, ITG_F64, 2 * n, seed, 2 * offset, x);
    if (rc) return rc;
    return itg_scal(ctx, ITG_F64, 2 * n, 0.70710678118654752440, x);
  }
  ITG_CHECK_DTYPE(ctx, dtype);
  if (n <= 0) return ITG_OK;
  if (dtype == ITG_F64)
    randn_kernel<double><<<itg_blocks(n, 256), 256, 0, ctx->stream>>>(n, seed, offset, (double*)x);
  else
    randn_kernel<float><<<itg_blocks(n, 256), 256, 0, ctx->stream>>>(n, seed, offset, (float*)x);
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}

// ================================================================ BLAS-1
template <typename T>
__global__ void scal_kernel(int64_t n, double a, T* x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = (T)(a * (double)x[i]);
}

template <typename T>
__global__ void axpby_kernel(int64_t n, double a, const T* x, double b, T* y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double v = a * (double)x[i];
    if (b != 0.0) v += b * (double)y[i];
    y[i] = (T)v;
  }
}

extern "C" int itg_scal(itg_ctx* ctx, int dtype, int64_t n, double alpha, void* x) {
  ITG_CHECK_CTX(ctx);
  if (dtype == ITG_C64) return itg_scal(ctx, ITG_F64, 2 * n, alpha, x);
  ITG_CHECK_DTYPE(ctx, dtype);
  if (n <= 0 || alpha == 1.0) return ITG_OK;
  if (dtype == ITG_F64)
    scal_kernel<double><<<itg_blocks(n, 256, 4096), 256, 0, ctx->stream>>>(n, alpha, (double*)x);
  else
    scal_kernel<float><<<itg_blocks(n, 256, 4096), 256, 0, ctx->stream>>>(n, alpha, (float*)x);
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}

extern "C" int itg_axpby(itg_ctx* ctx, int dtype, int64_t n, double alpha, const void* x,
                         double beta, void* y) {
  ITG_CHECK_CTX(ctx);
  if (dtype == ITG_C64) return itg_axpby(ctx, ITG_F64, 2 * n, alpha, x, beta, y);
  ITG_CHECK_DTYPE(ctx, dtype);
  if (n <= 0) return ITG_OK;
  if (dtype == ITG_F64)
    axpby_kernel<double><<<itg_blocks(n, 256, 4096), 256, 0, ctx->stream>>>(
        n, alpha, (const double*)x, beta, (double*)y);
  else
    axpby_kernel<float><<<itg_blocks(n, 256, 4096), 256, 0, ctx->stream>>>(
        n, alpha, (const float*)x, beta, (float*)y);
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}

template <typename T>
__global__ void precond_diag_kernel(int64_t n, const T* q, const T* dg, double lam, double fl,
                                    T* t) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double den = (double)dg[i] - lam;
    if (fabs(den) < fl) den = den < 0.0 ? -fl : fl;
    t[i] = (T)((double)q[i] / den);
  }
}

extern "C" int itg_precond_diag(itg_ctx* ctx, int dtype, int64_t n, const void* q,
                                const void* diag, double lambda, double floor, void* t) {
  ITG_CHECK_CTX(ctx);
  ITG_CHECK_DTYPE(ctx, dtype);
  if (n <= 0) return ITG_OK;
  if (!(floor > 0.0)) return itg_fail(ctx, ITG_EARG, "precond_diag: floor must be > 0");
  if (dtype == ITG_F64)
    precond_diag_kernel<double><<<itg_blocks(n, 256, 4096), 256, 0, ctx->stream>>>(
        n, (const double*)q, (const double*)diag, lambda, floor, (double*)t);
  else
    precond_diag_kernel<float><<<itg_blocks(n, 256, 4096), 256, 0, ctx->stream>>>(
        n, (const float*)q, (const float*)diag, lambda, floor, (float*)t);
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}

// block-level FP64 sum of one value per thread (256 threads)
__device__ __forceinline__ double block_sum_256(double v) {
  __shared__ double sh[4];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = sh[0] + sh[1] + sh[2] + sh[3];
  __syncthreads();
  return r;
}

template <typename T, bool SQ>
__global__ __launch_bounds__(256) void dot_partial_kernel(int64_t n, const T* x, const T* y,
                                                          double* part) {
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double a = (double)x[i];
    s += SQ ? a * a : a * (double)y[i];
  }
  double r = block_sum_256(s);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

__global__ __launch_bounds__(256) void sum_parts_kernel(int n, const double* part, double* out) {
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += part[i];
  double r = block_sum_256(s);
  if (threadIdx.x == 0) out[0] = r;
}

static int dot_common(itg_ctx* ctx, int dtype, int64_t n, const void* x, const void* y,
                      bool sq, double* out) {
  ITG_CHECK_CTX(ctx);
  if (dtype == ITG_C64) {
    if (!out) return itg_fail(ctx, ITG_EARG, "null output");
    if (sq) return dot_common(ctx, ITG_F64, 2 * n, x, nullptr, true, out);
    return itg_c64_dot(ctx, n, x, y, out);
  }
  ITG_CHECK_DTYPE(ctx, dtype);
  if (!out) return itg_fail(ctx, ITG_EARG, "null output");
  if (n <= 0) { *out = 0.0; return ITG_OK; }
  int nb = (int)std::min<int64_t>((n + 255) / 256, 2048);
  double* part = ctx->red + 16;
  if (dtype == ITG_F64) {
    if (sq) dot_partial_kernel<double, true><<<nb, 256, 0, ctx->stream>>>(n, (const double*)x, nullptr, part);
    else dot_partial_kernel<double, false><<<nb, 256, 0, ctx->stream>>>(n, (const double*)x, (const double*)y, part);
  } else {
    if (sq) dot_partial_kernel<float, true><<<nb, 256, 0, ctx->stream>>>(n, (const float*)x, nullptr, part);
    else dot_partial_kernel<float, false><<<nb, 256, 0, ctx->stream>>>(n, (const float*)x, (const float*)y, part);
  }
  sum_parts_kernel<<<1, 256, 0, ctx->stream>>>(nb, part, ctx->red);
  ITG_LAUNCH_CHECK(ctx);
  ITG_HIP(ctx, hipMemcpyAsync(ctx->red_host, ctx->red, sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
  ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *out = ctx->red_host[0];
  return ITG_OK;
}

extern "C" int itg_dot(itg_ctx* ctx, int dtype, int64_t n, const void* x, const void* y,
                       double* out) {
  return dot_common(ctx, dtype, n, x, y, false, out);
}

extern "C" int itg_nrm2(itg_ctx* ctx, int dtype, int64_t n, const void* x, double* out) {
  int rc = dot_common(ctx, dtype, n, x, nullptr, true, out);
  if (rc == ITG_OK) *out = std::sqrt(*out);
  return rc;
}

// ============================================================ permutation
// B = alpha*permutedims(A, perm) + beta*B. After merging modes that stay
// adjacent and in order, two kernels cover every case:
//  - tile_transpose: A's fastest mode and B's fastest mode differ; a 32x32
//    LDS tile is read coalesced along A's mode 0 and written coalesced along
//    B's mode 0, remaining modes are batch.
//  - elementwise: A's fastest mode stays B's fastest; reads and writes are
//    both coalesced along it.
struct PermDesc {
  int rank;
  int64_t dB[ITG_MAX_RANK];   // B's dims (output order)
  int64_t sA[ITG_MAX_RANK];   // A stride of B's mode q
};

template <typename T, bool BETA>
__global__ __launch_bounds__(256) void permute_elem_kernel(int64_t total, PermDesc d, double alpha,
                                                           const T* __restrict__ A, double beta,
                                                           T* __restrict__ B) {
  for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < total;
       o += (int64_t)gridDim.x * blockDim.x) {
    int64_t rem = o, offA = 0;
    for (int q = 0; q < d.rank; ++q) {
      int64_t j = rem % d.dB[q];
      rem /= d.dB[q];
      offA += j * d.sA[q];
    }
    T a = A[offA];
    if (BETA) {
      B[o] = (T)(alpha * (double)a + beta * (double)B[o]);
    } else {
      B[o] = alpha == 1.0 ? a : (T)(alpha * (double)a);
    }
  }
}

struct TransDesc {
  int64_t n0, np;           // A's mode-0 dim (i) and B's mode-0 dim (j)
  int64_t sAj;              // A stride of j
  int64_t sBi;              // B stride of i
  int nb;                   // batch modes
  int64_t bd[ITG_MAX_RANK], bsA[ITG_MAX_RANK], bsB[ITG_MAX_RANK];
  int64_t nbatch;
};

template <typename T, bool BETA>
__global__ __launch_bounds__(256) void permute_tile_kernel(TransDesc d, double alpha,
                                                           const T* __restrict__ A, double beta,
                                                           T* __restrict__ B) {
  __shared__ T tile[32][33];
  int64_t ti = (d.n0 + 31) / 32;
  int64_t t = blockIdx.x;
  int64_t i0 = (t % ti) * 32, j0 = (t / ti) * 32;
  int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int64_t bt = blockIdx.y; bt < d.nbatch; bt += gridDim.y) {
    int64_t rem = bt, oA = 0, oB = 0;
    for (int q = 0; q < d.nb; ++q) {
      int64_t j = rem % d.bd[q];
      rem /= d.bd[q];
      oA += j * d.bsA[q];
      oB += j * d.bsB[q];
    }
    for (int r = ty; r < 32; r += 8) {
      int64_t i = i0 + tx, j = j0 + r;
      if (i < d.n0 && j < d.np) tile[r][tx] = A[oA + i + j * d.sAj];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
      int64_t j = j0 + tx, i = i0 + r;
      if (i < d.n0 && j < d.np) {
        int64_t ob = oB + j + i * d.sBi;
        T a = tile[tx][r];
        if (BETA) B[ob] = (T)(alpha * (double)a + beta * (double)B[ob]);
        else B[ob] = alpha == 1.0 ? a : (T)(alpha * (double)a);
      }
    }
    __syncthreads();
  }
}

template <typename T>
static int permute_launch(itg_ctx* ctx, int rank, const int64_t* dimsA, const int32_t* perm,
                          double alpha, const T* A, double beta, T* B, bool use_beta) {
  // validate
  if (rank < 0 || rank > ITG_MAX_RANK) return itg_fail(ctx, ITG_EDIM, "rank %d out of range", rank);
  bool seen[ITG_MAX_RANK] = {false};
  int64_t total = 1;
  for (int q = 0; q < rank; ++q) {
    if (perm[q] < 0 || perm[q] >= rank || seen[perm[q]])
      return itg_fail(ctx, ITG_EDIM, "invalid permutation");
    seen[perm[q]] = true;
    if (dimsA[q] < 0) return itg_fail(ctx, ITG_EDIM, "negative dimension");
    total *= dimsA[q];
  }
  if (total == 0) return ITG_OK;
  // A strides
  int64_t sA[ITG_MAX_RANK];
  int64_t s = 1;
  for (int q = 0; q < rank; ++q) { sA[q] = s; s *= dimsA[q]; }
  // B modes in order: A-mode perm[q]; drop dims of 1 and merge runs
  std::vector<int> bm;            // A-mode indices in B order (non-unit)
  for (int q = 0; q < rank; ++q)
    if (dimsA[perm[q]] != 1) bm.push_back(perm[q]);
  std::vector<int64_t> md, msA, msB;  // merged: dim, A stride, B stride
  int64_t sb = 1;
  for (size_t q = 0; q < bm.size(); ++q) {
    int a = bm[q];
    if (!md.empty() && msA.back() * md.back() == sA[a]) {
      md.back() *= dimsA[a];
    } else {
      md.push_back(dimsA[a]);
      msA.push_back(sA[a]);
      msB.push_back(sb);
    }
    sb *= dimsA[a];
  }
  int r = (int)md.size();
  unsigned nthreads = 256;
  if (r == 0) { md.push_back(1); msA.push_back(1); msB.push_back(1); r = 1; }
  // fastest B mode is merged mode 0; find merged mode holding A stride 1
  int ia = -1;
  for (int q = 0; q < r; ++q) if (msA[q] == 1) ia = q;
  bool tile = (ia > 0) && md[0] >= 16 && md[ia] >= 16;
  if (tile) {
    TransDesc d;
    d.n0 = md[ia]; d.np = md[0]; d.sAj = msA[0]; d.sBi = msB[ia];
    d.nb = 0; d.nbatch = 1;
    for (int q = 1; q < r; ++q) {
      if (q == ia) continue;
      d.bd[d.nb] = md[q]; d.bsA[d.nb] = msA[q]; d.bsB[d.nb] = msB[q];
      d.nbatch *= md[q];
      ++d.nb;
    }
    int64_t tiles = ((d.n0 + 31) / 32) * ((d.np + 31) / 32);
    if (tiles > 0x7fffffff) return itg_fail(ctx, ITG_EUNSUPPORTED, "permute too large");
    dim3 grid((unsigned)tiles, (unsigned)std::min<int64_t>(d.nbatch, 65535));
    if (use_beta)
      permute_tile_kernel<T, true><<<grid, nthreads, 0, ctx->stream>>>(d, alpha, A, beta, B);
    else
      permute_tile_kernel<T, false><<<grid, nthreads, 0, ctx->stream>>>(d, alpha, A, beta, B);
  } else {
    PermDesc d;
    d.rank = r;
    for (int q = 0; q < r; ++q) { d.dB[q] = md[q]; d.sA[q] = msA[q]; }
    unsigned nb = itg_blocks(total, 256, 8192);
    if (use_beta)
      permute_elem_kernel<T, true><<<nb, nthreads, 0, ctx->stream>>>(total, d, alpha, A, beta, B);
    else
      permute_elem_kernel<T, false><<<nb, nthreads, 0, ctx->stream>>>(total, d, alpha, A, beta, B);
  }
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}

// complex storage as real storage with a leading re/im mode of dimension 2
// that the permutation keeps in front
static int c64_as_real(itg_ctx* ctx, int rank, const int64_t* dimsA, const int32_t* perm,
                       int64_t* d2, int32_t* p2) {
  if (rank < 0 || rank + 1 > ITG_MAX_RANK)
    return itg_fail(ctx, ITG_EDIM, "complex permute: rank %d out of range", rank);
  d2[0] = 2;
  p2[0] = 0;
  for (int q = 0; q < rank; ++q) {
    if (perm[q] < 0 || perm[q] >= rank) return itg_fail(ctx, ITG_EDIM, "invalid permutation");
    d2[q + 1] = dimsA[q];
    p2[q + 1] = perm[q] + 1;
  }
  return ITG_OK;
}

extern "C" int itg_permute(itg_ctx* ctx, int dtype, int rank, const int64_t* dimsA,
                           const int32_t* perm, double alpha, const void* A, void* B) {
  ITG_CHECK_CTX(ctx);
  if (dtype == ITG_C64) {
    if (A == B) return itg_fail(ctx, ITG_EARG, "permute cannot run in place");
    int64_t d2[ITG_MAX_RANK];
    int32_t p2[ITG_MAX_RANK];
    int rc = c64_as_real(ctx, rank, dimsA, perm, d2, p2);
    if (rc) return rc;
    return permute_launch<double>(ctx, rank + 1, d2, p2, alpha, (const double*)A, 0.0,
                                  (double*)B, false);
  }
  ITG_CHECK_DTYPE(ctx, dtype);
  if (A == B) return itg_fail(ctx, ITG_EARG, "permute cannot run in place");
  if (dtype == ITG_F64)
    return permute_launch<double>(ctx, rank, dimsA, perm, alpha, (const double*)A, 0.0,
                                  (double*)B, false);
  return permute_launch<float>(ctx, rank, dimsA, perm, alpha, (const float*)A, 0.0, (float*)B,
                               false);
}

extern "C" int itg_add_permuted(itg_ctx* ctx, int dtype, int rank, const int64_t* dimsA,
                                const int32_t* perm, double alpha, const void* A, double beta,
                                void* B) {
  ITG_CHECK_CTX(ctx);
  if (dtype == ITG_C64) {
    int64_t d2[ITG_MAX_RANK];
    int32_t p2[ITG_MAX_RANK];
    int rc = c64_as_real(ctx, rank, dimsA, perm, d2, p2);
    if (rc) return rc;
    return itg_add_permuted(ctx, ITG_F64, rank + 1, d2, p2, alpha, A, beta, B);
  }
  ITG_CHECK_DTYPE(ctx, dtype);
  if (A == B) {
    for (int q = 0; q < rank; ++q)
      if (perm[q] != q) return itg_fail(ctx, ITG_EARG, "in-place add needs identity permutation");
    int64_t n = 1;
    for (int q = 0; q < rank; ++q) n *= dimsA[q];
    return itg_scal(ctx, dtype, n, alpha + beta, B);
  }
  if (dtype == ITG_F64)
    return permute_launch<double>(ctx, rank, dimsA, perm, alpha, (const double*)A, beta,
                                  (double*)B, true);
  return permute_launch<float>(ctx, rank, dimsA, perm, alpha, (const float*)A, beta, (float*)B,
                               true);
}

// ============================================================ truncation
extern "C" int itg_truncate(itg_ctx* ctx, const double* P, int64_t n, int64_t maxdim,
                            int64_t mindim, double cutoff, int abs_cut, int rel_cut,
                            int semantics, int64_t* k, double* truncerr, double* docut) {
  ITG_CHECK_CTX(ctx);
  if (n < 0 || (n > 0 && !P)) return itg_fail(ctx, ITG_EARG, "bad weight array");
  if (semantics != ITG_TRUNC_NDTENSORS && semantics != ITG_TRUNC_REFGPU)
    return itg_fail(ctx, ITG_EARG, "unknown truncate semantics %d", semantics);
  itg_trunc_result r = itg_truncate_core(P, n, maxdim, mindim, cutoff, abs_cut, rel_cut,
                                         semantics);
  if (k) *k = r.n;
  if (truncerr) *truncerr = r.truncerr;
  if (docut) *docut = r.docut;
  return ITG_OK;
}

// The decision is the same serial rule as on the host; the weights are first
// staged in LDS (up to kTruncLds of them) so its scans do not wait on HBM.
constexpr int64_t kTruncLds = 8192;

// lane j's double, broadcast to the wave (j uniform)
__device__ __forceinline__ double wave_lane(double v, int j) {
  int lo = __builtin_amdgcn_readlane(__double2loint(v), j);
  int hi = __builtin_amdgcn_readlane(__double2hiint(v), j);
  return __hiloint2double(hi, lo);
}

// The NDTensors decision of itg_truncate_core for one wave: the weights are
// read 64 at a time into registers (the next 64 in flight while the current
// ones are used) and every sum is taken element by element in the core's
// order through lane broadcasts, so (n, truncerr, docut) are bit-identical
// to the host's while the serial chain runs on register operands.
__device__ void truncate_ndtensors_wave(const double* P, int64_t m, int64_t maxdim,
                                        int64_t mindim, double cutoff, int abs_cut, int rel_cut,
                                        double* out) {
  const int lane = threadIdx.x & 63;
  if (maxdim <= 0 || maxdim > m) maxdim = m;
  if (mindim < 1) mindim = 1;
  if (cutoff < 0.0) cutoff = 0.0;
  double rn = 1.0, rte = 0.0, rdc = 0.0;
  const double p0 = P[0];
  if (!(p0 <= 0.0) && m == 1) rdc = p0 / 2;
  if (!(p0 <= 0.0) && m > 1) {
    // lastpos = 1 + the last index whose weight is not negative
    int64_t lp = 0;
    for (int64_t i = lane; i < m; i += 64)
      if (!(P[i] < 0.0)) lp = i + 1;
    for (int o = 32; o > 0; o >>= 1) {
      int64_t t = __shfl_xor(lp, o, 64);
      lp = t > lp ? t : lp;
    }
    auto W = [&](int64_t i) { return (i >= 0 && i < lp && i < m) ? P[i] : 0.0; };
    int64_t n = m;
    double te = 0.0;
    // from the tail: chunk c holds W(hi - 63 .. hi), lane l = W(hi - l)
    int64_t hi = m - 1;
    double cur = W(hi - lane), nxt = W(hi - 64 - lane);
    int j = 0;   // next lane of `cur` to use
    auto advance = [&]() {
      if (++j == 64) {
        j = 0;
        hi -= 64;
        cur = nxt;
        nxt = W(hi - 64 - lane);
      }
    };
    // whole chunks of the maxdim tail in one unrolled chain (constant lanes),
    // the rest element by element
    while (n - maxdim >= 64) {
#pragma unroll
      for (int l = 0; l < 64; ++l) te += wave_lane(cur, l);
      n -= 64;
      hi -= 64;
      cur = nxt;
      nxt = W(hi - 64 - lane);
    }
    while (n > maxdim) {
      te += wave_lane(cur, j);
      --n;
      advance();
    }
    if (abs_cut) {
      while (n > mindim) {
        double w = wave_lane(cur, j);
        if (!(w <= cutoff)) break;
        te += w;
        --n;
        advance();
      }
    } else {
      double scale = 1.0;
      if (rel_cut) {
        scale = 0.0;
        double a = W(lane), b = W(64 + lane);
        int64_t base = 0;
        for (; base + 64 <= m; base += 64) {
#pragma unroll
          for (int l = 0; l < 64; ++l) scale += wave_lane(a, l);
          a = b;
          b = W(base + 128 + lane);
        }
        for (int l = 0; l < (int)(m - base); ++l) scale += wave_lane(a, l);
        if (scale == 0.0) scale = 1.0;
      }
      const double lim = cutoff * scale;
      while (n > mindim) {
        double w = wave_lane(cur, j);
        if (!(te + w <= lim)) break;
        te += w;
        --n;
        advance();
      }
      te /= scale;
    }
    if (n < 1) n = 1;
    rn = (double)n;
    rte = te;
    if (n < m) {
      double a = W(n - 1), b = W(n);
      rdc = (a + b) / 2.0;
      double diff = a - b;
      if (diff < 0) diff = -diff;
      if (diff < 1e-3 * a) rdc += 1e-3 * a;
    }
  }
  if (lane == 0) {
    out[0] = rn;
    out[1] = rte;
    out[2] = rdc;
  }
}

__global__ void truncate_kernel(const double* P, int64_t n, int64_t maxdim, int64_t mindim,
                                double cutoff, int abs_cut, int rel_cut, int semantics,
                                double* out) {
  extern __shared__ double Ps[];
  const bool staged = n <= kTruncLds;
  if (staged) {
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) Ps[i] = P[i];
    __syncthreads();
  }
  if (semantics == ITG_TRUNC_NDTENSORS) {
    if (threadIdx.x < 64)
      truncate_ndtensors_wave(staged ? Ps : P, n, maxdim, mindim, cutoff, abs_cut, rel_cut, out);
    return;
  }
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    itg_trunc_result r = itg_truncate_core(staged ? Ps : P, n, maxdim, mindim, cutoff, abs_cut,
                                           rel_cut, semantics);
    out[0] = (double)r.n;
    out[1] = r.truncerr;
    out[2] = r.docut;
  }
}

int itg_truncate_device_async(itg_ctx* ctx, const double* P, int64_t n, int64_t maxdim,
                              int64_t mindim, double cutoff, int abs_cut, int rel_cut,
                              int semantics, double* out3) {
  static bool lds_set = false;
  if (!lds_set) {
    ITG_HIP(ctx, hipFuncSetAttribute((const void*)truncate_kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)(kTruncLds * sizeof(double))));
    lds_set = true;
  }
  size_t lds = n <= kTruncLds ? (size_t)std::max<int64_t>(n, 1) * sizeof(double) : 0;
  truncate_kernel<<<1, 256, lds, ctx->stream>>>(P, n, maxdim, mindim, cutoff, abs_cut, rel_cut,
                                                semantics, out3);
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}

extern "C" int itg_truncate_device(itg_ctx* ctx, const double* P, int64_t n, int64_t maxdim,
                                   int64_t mindim, double cutoff, int abs_cut, int rel_cut,
                                   int semantics, int64_t* k, double* truncerr, double* docut) {
  ITG_CHECK_CTX(ctx);
  if (n < 0 || (n > 0 && !P)) return itg_fail(ctx, ITG_EARG, "bad weight array");
  if (semantics != ITG_TRUNC_NDTENSORS && semantics != ITG_TRUNC_REFGPU)
    return itg_fail(ctx, ITG_EARG, "unknown truncate semantics %d", semantics);
  int rc = itg_truncate_device_async(ctx, P, n, maxdim, mindim, cutoff, abs_cut, rel_cut,
                                     semantics, ctx->red);
  if (rc) return rc;
  ITG_HIP(ctx, hipMemcpyAsync(ctx->red_host, ctx->red, 3 * sizeof(double),
                              hipMemcpyDeviceToHost, ctx->stream));
  ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (k) *k = (int64_t)ctx->red_host[0];
  if (truncerr) *truncerr = ctx->red_host[1];
  if (docut) *docut = ctx->red_host[2];
  return ITG_OK;
}
