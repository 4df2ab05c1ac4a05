// Plain FP64 GEMMs on the vendor library (rocBLAS). On the products of this
// path the library's FP64 MFMA kernels run at 61-76 TFLOP/s (4096^3,
// 8192x2048x10240, 10240x8192x2048) against 34 TFLOP/s for the GETT kernel,
// whose point is the permutation fused into its loads: contractions whose
// modes group into plain column-major matrices, and itg_gemm itself, come
// here; the rest stay on GETT. Products inside a stream capture or a
// no-alloc call stay on GETT too (the library may allocate its workspace).
#include <rocblas/rocblas.h>

#include <climits>
#include <cstdlib>

#include "itg_internal.h"

namespace {

// products with a dimension below this stay on GETT: on the eigensolver's
// skinny back-transform products (32 to 128 rows or a K of 32 to 128) the
// library was slower (73.1 against 69.4 ms per n = 4096 eigendecomposition
// with them on it), and small products launch faster there
constexpr int64_t kLibMinDim = 256;

rocblas_handle handle_of(itg_ctx* ctx) {
  if (!ctx->blas) {
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) return nullptr;
    // no atomics: every call sums in a fixed order (bit-reproducible)
    rocblas_set_atomics_mode(h, rocblas_atomics_not_allowed);
    rocblas_set_pointer_mode(h, rocblas_pointer_mode_host);
    ctx->blas = h;
  }
  return (rocblas_handle)ctx->blas;
}

bool fits32(int64_t x) { return x >= 0 && x <= INT_MAX; }

rocblas_operation op_of(char t) {
  return (t == 'T' || t == 't') ? rocblas_operation_transpose : rocblas_operation_none;
}

}  // namespace

bool itg_libgemm_ok(itg_ctx* ctx, int dtype, int mode, int64_t m, int64_t n, int64_t k) {
  if (!ctx->libgemm || dtype != ITG_F64 || mode != ITG_MODE_NATIVE) return false;
  if (ctx->fixed_ws) return false;
  if (!fits32(m) || !fits32(n) || !fits32(k)) return false;
  if (m < kLibMinDim || n < kLibMinDim || k < kLibMinDim) return false;
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(ctx->stream, &st) != hipSuccess) return false;
  return st == hipStreamCaptureStatusNone;
}

int itg_libgemm(itg_ctx* ctx, char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha,
                const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                double* C, int64_t ldc) {
  if (!fits32(lda) || !fits32(ldb) || !fits32(ldc))
    return itg_fail(ctx, ITG_EUNSUPPORTED, "leading dimension beyond 32 bits");
  rocblas_handle h = handle_of(ctx);
  if (!h) return itg_fail(ctx, ITG_EHIP, "rocblas_create_handle failed");
  rocblas_set_stream(h, ctx->stream);
  const int t = itg_tstart(ctx, ITG_K_GETT);
  rocblas_status s = rocblas_dgemm(h, op_of(ta), op_of(tb), (rocblas_int)m, (rocblas_int)n,
                                   (rocblas_int)k, &alpha, A, (rocblas_int)lda, B,
                                   (rocblas_int)ldb, &beta, C, (rocblas_int)ldc);
  itg_tstop(ctx, t, ITG_K_GETT, 2.0 * m * n * k,
            8.0 * (m * k + k * n + m * n * (beta != 0.0 ? 2 : 1)));
  if (s != rocblas_status_success)
    return itg_fail(ctx, ITG_EHIP, "rocblas_dgemm: %s", rocblas_status_to_string(s));
  return ITG_OK;
}

int itg_libsyrk_lower(itg_ctx* ctx, char trans, int64_t n, int64_t k, double alpha,
                      const double* A, int64_t lda, double* C, int64_t ldc) {
  if (!fits32(lda) || !fits32(ldc))
    return itg_fail(ctx, ITG_EUNSUPPORTED, "leading dimension beyond 32 bits");
  rocblas_handle h = handle_of(ctx);
  if (!h) return itg_fail(ctx, ITG_EHIP, "rocblas_create_handle failed");
  rocblas_set_stream(h, ctx->stream);
  const double beta = 0.0;
  const int t = itg_tstart(ctx, ITG_K_GETT);
  rocblas_status s = rocblas_dsyrk(h, rocblas_fill_lower, op_of(trans), (rocblas_int)n,
                                   (rocblas_int)k, &alpha, A, (rocblas_int)lda, &beta, C,
                                   (rocblas_int)ldc);
  itg_tstop(ctx, t, ITG_K_GETT, 1.0 * n * (n + 1) * k, 8.0 * (n * k + 0.5 * n * (n + 1)));
  if (s != rocblas_status_success)
    return itg_fail(ctx, ITG_EHIP, "rocblas_dsyrk: %s", rocblas_status_to_string(s));
  return ITG_OK;
}

void itg_libgemm_free(itg_ctx* ctx) {
  if (ctx->blas) {
    rocblas_destroy_handle((rocblas_handle)ctx->blas);
    ctx->blas = nullptr;
  }
}
