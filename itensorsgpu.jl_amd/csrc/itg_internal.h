// Internal declarations shared by the libitg translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/itg.h"

struct itg_plan;  // contraction plan (gett.hip)

struct itg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;       // stream all work is queued on
  hipStream_t own_stream = nullptr;   // the context's own stream
  std::string err;
  int num_cu = 0;
  int gfx = 0;
  double hbm_gb = 0.0;
  // grow-only device workspace for internal temporaries
  void* ws = nullptr;
  size_t ws_bytes = 0;
  // grow-only device workspace of the factorisations (eigh/svd/qr)
  void* eig_ws = nullptr;
  size_t eig_ws_bytes = 0;
  // grow-only device workspace of the MLP backward pass
  void* mlp_ws = nullptr;
  size_t mlp_ws_bytes = 0;
  // small device scratch for reductions + pinned host mirror
  double* red = nullptr;        // device, kRedSlots doubles
  double* red_host = nullptr;   // pinned host, kRedSlots doubles
  // contraction plan cache (keyed by the canonical contraction signature)
  std::unordered_map<std::string, itg_plan*> plans;
  size_t plan_bytes = 0;
  // per-kernel event timing (itg_timing)
  int timing = 0;   // bit mask of timed kernel classes
  struct TimedLaunch {
    int kind;
    hipEvent_t a, b;
    double flops, bytes;
  };
  std::vector<TimedLaunch> launches;
  std::vector<hipEvent_t> event_pool;
  size_t event_next = 0;
  // captured Heff-apply graphs of itg_davidson2 (null: seen once, eager)
  std::unordered_map<std::string, hipGraphExec_t> graphs;
  // this device's copy of the no-alloc scratch array (heff.hip), resolved
  // once per device by itg_noalloc_setup
  void* noalloc_scratch = nullptr;
  // while set, GEMM split-K partials go here (at most fixed_ws_bytes)
  // instead of the grow-only workspace: the no-alloc entry points
  void* fixed_ws = nullptr;
  size_t fixed_ws_bytes = 0;
  // grow-only device workspace of the SVD (apart from eig_ws: the SVD calls
  // the eigensolver)
  void* svd_ws = nullptr;
  size_t svd_ws_bytes = 0;
  // grow-only device workspace of the complex QR and SVD
  void* zla_ws = nullptr;
  size_t zla_ws_bytes = 0;
  // grow-only device workspace of the complex contraction's expanded operand
  void* cws = nullptr;
  size_t cws_bytes = 0;
  // vendor BLAS handle for plain FP64 GEMMs (created on first use) and
  // whether they go there (ITG_GEMM_LIB=0 at init keeps them on GETT)
  void* blas = nullptr;
  int libgemm = 1;
};

constexpr int kRedSlots = 16384;

// ---------------------------------------------------------------- errors
int itg_fail(itg_ctx* ctx, int code, const char* fmt, ...);

#define ITG_HIP(ctx, call)                                                    \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      return itg_fail((ctx), ITG_EHIP, "%s failed: %s (%s:%d)", #call,        \
                      hipGetErrorString(e_), __FILE__, __LINE__);             \
  } while (0)

#define ITG_LAUNCH_CHECK(ctx)                                                 \
  do {                                                                        \
    hipError_t e_ = hipGetLastError();                                        \
    if (e_ != hipSuccess)                                                     \
      return itg_fail((ctx), ITG_EHIP, "kernel launch failed: %s (%s:%d)",    \
                      hipGetErrorString(e_), __FILE__, __LINE__);             \
  } while (0)

#define ITG_CHECK_CTX(ctx)                                                    \
  do {                                                                        \
    if (!(ctx)) return ITG_EARG;                                              \
  } while (0)

#define ITG_CHECK_DTYPE(ctx, dt)                                              \
  do {                                                                        \
    if ((dt) != ITG_F64 && (dt) != ITG_F32)                                   \
      return itg_fail((ctx), ITG_EDTYPE, "unsupported dtype %d", (int)(dt));  \
  } while (0)

static inline size_t itg_dtype_size(int dt) { return dt == ITG_C64 ? 16 : dt == ITG_F64 ? 8 : 4; }

// Grow-only workspace; the returned pointer is valid until the next call.
int itg_workspace(itg_ctx* ctx, size_t bytes, void** out);

// Reduction helper: sums `nparts` doubles already in ctx->red[0..nparts) on
// device and copies the result (and extra slots) to the host synchronously.
int itg_read_reduction(itg_ctx* ctx, int nslots, double* host_out);

static inline unsigned itg_blocks(int64_t n, int threads, int64_t cap = 65535 * 8) {
  int64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned)b;
}

// Internal (non-exported) entry points used across translation units.
// lower_band: compute only the 128 x 128 output tiles on or below the
// diagonal and the first tile band above it (square C whose upper part far
// from the diagonal is never read)
int itg_gemm_internal(itg_ctx* ctx, int dtype, int mode, char ta, char tb,
                      int64_t m, int64_t n, int64_t k, double alpha,
                      const void* A, int64_t lda, const void* B, int64_t ldb,
                      double beta, void* C, int64_t ldc, bool lower_band = false);
void itg_plan_cache_clear(itg_ctx* ctx);
// Plain FP64 GEMMs on the vendor library (libgemm.hip). itg_libgemm_ok says
// whether a native FP64 product of this size may go there (not while the
// stream is being captured, not inside a no-alloc call, dimensions in 32
// bits); itg_libgemm is column-major C = alpha op(A) op(B) + beta C and
// itg_libsyrk the lower triangle of C = alpha op(A) op(A)^T.
bool itg_libgemm_ok(itg_ctx* ctx, int dtype, int mode, int64_t m, int64_t n, int64_t k);
int itg_libgemm(itg_ctx* ctx, char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha,
                const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                double* C, int64_t ldc);
int itg_libsyrk_lower(itg_ctx* ctx, char trans, int64_t n, int64_t k, double alpha,
                      const double* A, int64_t lda, double* C, int64_t ldc);
void itg_libgemm_free(itg_ctx* ctx);
// Event bracket of one launch for itg_timing: t = itg_tstart(ctx, kind)
// before the launch, itg_tstop(ctx, t, kind, flops, bytes) after it (no-ops
// unless the kind is being timed).
int itg_tstart(itg_ctx* ctx, int kind);
void itg_tstop(itg_ctx* ctx, int t, int kind, double flops, double bytes);
void itg_timing_free(itg_ctx* ctx);
// For a launch made with hipExtLaunchKernelGGL(..., ev.a, ev.b, 0, ...): the
// runtime stamps the two events at the kernel's own start and end, so short
// kernels are timed without the dispatch gap a separate event record adds.
// ev.a/ev.b are null when the kind is not timed (a plain launch).
struct itg_tev {
  int t;
  hipEvent_t a, b;
};
itg_tev itg_tevents(itg_ctx* ctx, int kind);
void itg_tcommit(itg_ctx* ctx, const itg_tev& ev, int kind, double flops, double bytes);
void itg_graph_cache_clear(itg_ctx* ctx);
// Resolves (once per device, with that device current) the device address of
// the no-alloc scratch array into ctx->noalloc_scratch.
int itg_noalloc_setup(itg_ctx* ctx);
// Prepared two-site Heff apply (heff.hip) with dims = {chiL, wl, wm, wr, d,
// chiR}; validates the shape, launches only.
int itg_heff2_apply_internal(itg_ctx* ctx, int dtype, int mode, const int64_t* dims,
                             const void* L, const void* R, const void* prep, const void* psi,
                             void* out, void* work);

// ComplexF64 (complex.hip): pieces the generic entry points dispatch to
int itg_c64_fill(itg_ctx* ctx, int64_t n, double value, void* x);
int itg_c64_convert(itg_ctx* ctx, int64_t n, int sdt, const void* src, int ddt, void* dst);
// out[0], out[1] = re, im of sum conj(x) y
int itg_c64_dot(itg_ctx* ctx, int64_t n, const void* x, const void* y, double* out);
int itg_c64_contract(itg_ctx* ctx, int mode, int rA, const int64_t* dA, const int32_t* lA,
                     const void* A, int rB, const int64_t* dB, const int32_t* lB, const void* B,
                     int rC, const int64_t* dC, const int32_t* lC, void* C, double alpha,
                     double beta);
// Hermitian A (n x n complex, full storage) -> real tridiagonal (d, e); the
// reflectors stay in A's strict lower part, their factors in tau (n complex);
// scratch: itg_zhetrd_scratch_doubles(n) doubles
size_t itg_zhetrd_scratch_doubles(int n);
int itg_zhetrd(itg_ctx* ctx, int n, void* A, double* d, double* e, void* tau, double* scratch);
// U (n x k complex) = Q Z, Q the product of itg_zhetrd's reflectors, Z real
// (n x k, leading dimension n)
int itg_zunmtr(itg_ctx* ctx, int n, int k, const void* A, const void* tau, const double* Z,
               void* U);
int itg_c64_qr(itg_ctx* ctx, int64_t m, int64_t n, const void* A, void* Q, void* R);
int itg_c64_svd_trunc(itg_ctx* ctx, int64_t m, int64_t n, void* A, void* U, double* S, void* V,
                      int64_t maxdim, int64_t mindim, double cutoff, int abs_cut, int rel_cut,
                      int64_t* k, double* truncerr, double* docut);
