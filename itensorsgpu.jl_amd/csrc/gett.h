// GEMM-coordinate description of a contraction, shared by the contraction
// entry points and library-internal callers (eigensolver merges) that bring
// their own device-resident offset tables.
#pragma once

#include <cstdint>

struct itg_ctx;

struct GettArgs {
  const void* A;
  const void* B;
  void* C;
  int64_t M, N, K;
  // affine strides (>= 0) or -1 to use the table
  int64_t sAm, sAk, sBk, sBn, sCm, sCn;
  const int64_t* tAm;
  const int64_t* tAk;
  const int64_t* tBk;
  const int64_t* tBn;
  const int64_t* tCm;
  const int64_t* tCn;
  double alpha, beta;
  int64_t kchunk;   // K range per blockIdx.z (set by the launcher)
  void* partial;    // split-K partials [z][n][m] in the compute type, or null
  int64_t tilesM, tilesN;
  // 1: skip output tiles (tm, tn) with tm > tn + 1 (callers that need only
  // the lower triangle plus one tile band of a square C; no split-K)
  int band;
};

// Launch C(m,n) = alpha * sum_k A(m,k) B(k,n) + beta * C(m,n) with the
// addressing in `g` (tables, if any, must be device pointers). ak/bk select
// k-contiguous operand loads.
int itg_gett_launch(itg_ctx* ctx, int dtype, int mode, GettArgs* g, bool ak, bool bk);
