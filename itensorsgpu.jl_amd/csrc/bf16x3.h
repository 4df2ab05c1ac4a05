// Split-BF16 operands and the GEMM over them (bf16x3.hip).
#pragma once

#include <cstdint>

#include "itg_internal.h"

// Leading dimension of a split operand with K columns (multiple of 64).
int64_t itg_split_ldk(int64_t K);

// Split the R x K matrix with element (r, k) at src[r*sr + k*sk] (FP64 or
// FP32 per dtype) into bf16 planes hi, lo of R rows x ldk (zero padded).
// Planes must be 16-byte aligned.
int itg_split_bf16(itg_ctx* ctx, int dtype, const void* src, int64_t R, int64_t K, int64_t sr,
                   int64_t sk, int64_t ldk, uint16_t* hi, uint16_t* lo);

// C(m, n) = alpha * sum_k A(m, k) B(n, k) + beta * C(m, n) for m < M, n < N
// over split planes A (M rows, lda) and B (N rows, ldb); Kp a multiple of 64
// covering K (padding is zero); C column-major (ldc) of type dtype.
int itg_gemm_split(itg_ctx* ctx, int dtype, int64_t M, int64_t N, int64_t K, int64_t Kp,
                   double alpha, const uint16_t* Ah, const uint16_t* Al, int64_t lda,
                   const uint16_t* Bh, const uint16_t* Bl, int64_t ldb, double beta, void* C,
                   int64_t ldc);
