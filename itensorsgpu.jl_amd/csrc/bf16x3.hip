// Split-BF16 ("bf16x3") GEMM on operands split once into bf16 planes.
//
// x = hi + lo with hi = bf16_rne(x), lo = bf16_rne(x - hi); a product is
// ah*bh + ah*bl + al*bh with FP32 accumulation on MFMA (lo*lo dropped).
// A split operand is two planes, hi and lo, each row-major [row][k] with
// leading dimension ldk (a multiple of 64, zero padded for k >= K), so an
// MFMA lane's 8 consecutive k are one 16-byte load and the GEMM's inner loop
// does no conversion and reads 4 bytes per element instead of 8.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "bf16x3.h"
#include "itg_internal.h"

namespace {

typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
typedef float f4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// ---------------------------------------------------------------- split
// One block: a 64 (rows) x 64 (k) tile. Element (r, k) of the source is at
// src[r*sr + k*sk]. Reads run along whichever of r, k has unit stride; the
// tile is transposed through LDS when needed; writes are 16 bytes per lane.
template <typename T>
__global__ __launch_bounds__(256) void split_kernel(const T* __restrict__ src, int64_t R,
                                                    int64_t K, int64_t sr, int64_t sk,
                                                    int64_t ldk, int64_t tilesK,
                                                    uint16_t* __restrict__ hi,
                                                    uint16_t* __restrict__ lo) {
  __shared__ float tile[64][65];
  const int t = threadIdx.x;
  const int64_t r0 = (blockIdx.x / tilesK) * 64;
  const int64_t k0 = (blockIdx.x % tilesK) * 64;
  if (sk == 1 || sr != 1) {
    // k fastest: thread -> (k = t % 64, r = t / 64 + 4 i)
    const int kk = t & 63;
#pragma unroll 4
    for (int i = 0; i < 16; ++i) {
      int rr = (t >> 6) + 4 * i;
      int64_t r = r0 + rr, k = k0 + kk;
      tile[rr][kk] = (r < R && k < K) ? (float)src[r * sr + k * sk] : 0.0f;
    }
  } else {
    // r fastest
    const int rr = t & 63;
#pragma unroll 4
    for (int i = 0; i < 16; ++i) {
      int kk = (t >> 6) + 4 * i;
      int64_t r = r0 + rr, k = k0 + kk;
      tile[rr][kk] = (r < R && k < K) ? (float)src[r * sr + k * sk] : 0.0f;
    }
  }
  __syncthreads();
  // 64 rows x 8 chunks of 8 k: two chunks per thread
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    int c = t + 256 * i;
    int rr = c >> 3, kc = (c & 7) * 8;
    int64_t r = r0 + rr;
    if (r >= R) continue;
    uint16_t h[8], l[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      float x = tile[rr][kc + e];
      h[e] = f2bf(x);
      l[e] = f2bf(x - bf2f(h[e]));
    }
    uint4 hv, lv;
    hv.x = h[0] | ((uint32_t)h[1] << 16); hv.y = h[2] | ((uint32_t)h[3] << 16);
    hv.z = h[4] | ((uint32_t)h[5] << 16); hv.w = h[6] | ((uint32_t)h[7] << 16);
    lv.x = l[0] | ((uint32_t)l[1] << 16); lv.y = l[2] | ((uint32_t)l[3] << 16);
    lv.z = l[4] | ((uint32_t)l[5] << 16); lv.w = l[6] | ((uint32_t)l[7] << 16);
    *(uint4*)(hi + r * ldk + k0 + kc) = hv;
    *(uint4*)(lo + r * ldk + k0 + kc) = lv;
  }
}

// ----------------------------------------------------------------- GEMM
// C(m, n) = alpha * sum_k A(m, k) B(n, k) + beta * C(m, n), C column-major
// with leading dimension ldc. 128 x 128 tile per 256-thread workgroup, 2 x 2
// waves of 64 x 64, K-tile 64 (two 16x16x32 MFMA steps). The MFMA is issued
// as D = B_frag . A_frag so that a lane's output column index runs along m:
// sixteen consecutive m per store, C's unit-stride direction.
constexpr int SB_M = 128, SB_N = 128, SB_K = 64, SB_PAD = 8, SB_LD = SB_K + SB_PAD;

template <typename T>
__global__ __launch_bounds__(256, 2) void gemm_split_kernel(
    const uint16_t* __restrict__ Ah, const uint16_t* __restrict__ Al, int64_t lda,
    const uint16_t* __restrict__ Bh, const uint16_t* __restrict__ Bl, int64_t ldb, int64_t M,
    int64_t N, int64_t Kp, T* __restrict__ C, int64_t ldc, double alpha, double beta,
    int64_t tilesM, int64_t tilesN) {
  extern __shared__ __attribute__((aligned(16))) uint16_t lds[];
  uint16_t* sAh = lds;
  uint16_t* sAl = sAh + SB_M * SB_LD;
  uint16_t* sBh = sAl + SB_M * SB_LD;
  uint16_t* sBl = sBh + SB_N * SB_LD;

  // XCD-aware grouped tile order (8 XCDs, groups of 8 tile rows)
  int64_t nb = (int64_t)gridDim.x, bid = blockIdx.x;
  int64_t lid = (nb % 8 == 0) ? (bid % 8) * (nb / 8) + bid / 8 : bid;
  const int64_t GM = 8;
  int64_t per_group = GM * tilesN;
  int64_t group = lid / per_group;
  int64_t first = group * GM;
  int64_t gsize = std::min<int64_t>(tilesM - first, GM);
  int64_t in = lid % per_group;
  const int64_t m0 = (first + in % gsize) * SB_M;
  const int64_t n0 = (in / gsize) * SB_N;

  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int lr = lane & 15, lg = lane >> 4;

  // loader: chunk c = t + 256 i (i < 4) -> row c / 8 = t / 8 + 32 i, k-chunk
  // (t % 8) * 8; named registers (no arrays) for the prefetched chunks
  const int lrow0 = t >> 3, lkc = (t & 7) * 8;
  const int64_t ar0 = std::min<int64_t>(m0 + lrow0, M - 1) * lda + lkc;
  const int64_t ar1 = std::min<int64_t>(m0 + lrow0 + 32, M - 1) * lda + lkc;
  const int64_t ar2 = std::min<int64_t>(m0 + lrow0 + 64, M - 1) * lda + lkc;
  const int64_t ar3 = std::min<int64_t>(m0 + lrow0 + 96, M - 1) * lda + lkc;
  const int64_t br0 = std::min<int64_t>(n0 + lrow0, N - 1) * ldb + lkc;
  const int64_t br1 = std::min<int64_t>(n0 + lrow0 + 32, N - 1) * ldb + lkc;
  const int64_t br2 = std::min<int64_t>(n0 + lrow0 + 64, N - 1) * ldb + lkc;
  const int64_t br3 = std::min<int64_t>(n0 + lrow0 + 96, N - 1) * ldb + lkc;
  const int so = lrow0 * SB_LD + lkc, sstep = 32 * SB_LD;
  uint4 ah0, ah1, ah2, ah3, al0, al1, al2, al3, bh0, bh1, bh2, bh3, bl0, bl1, bl2, bl3;
#define SPLIT_LOAD(k0)                                                    \
  {                                                                       \
    ah0 = *(const uint4*)(Ah + ar0 + (k0)); al0 = *(const uint4*)(Al + ar0 + (k0)); \
    ah1 = *(const uint4*)(Ah + ar1 + (k0)); al1 = *(const uint4*)(Al + ar1 + (k0)); \
    ah2 = *(const uint4*)(Ah + ar2 + (k0)); al2 = *(const uint4*)(Al + ar2 + (k0)); \
    ah3 = *(const uint4*)(Ah + ar3 + (k0)); al3 = *(const uint4*)(Al + ar3 + (k0)); \
    bh0 = *(const uint4*)(Bh + br0 + (k0)); bl0 = *(const uint4*)(Bl + br0 + (k0)); \
    bh1 = *(const uint4*)(Bh + br1 + (k0)); bl1 = *(const uint4*)(Bl + br1 + (k0)); \
    bh2 = *(const uint4*)(Bh + br2 + (k0)); bl2 = *(const uint4*)(Bl + br2 + (k0)); \
    bh3 = *(const uint4*)(Bh + br3 + (k0)); bl3 = *(const uint4*)(Bl + br3 + (k0)); \
  }
#define SPLIT_STORE()                                                     \
  {                                                                       \
    *(uint4*)(sAh + so) = ah0; *(uint4*)(sAl + so) = al0;                 \
    *(uint4*)(sAh + so + sstep) = ah1; *(uint4*)(sAl + so + sstep) = al1; \
    *(uint4*)(sAh + so + 2 * sstep) = ah2; *(uint4*)(sAl + so + 2 * sstep) = al2; \
    *(uint4*)(sAh + so + 3 * sstep) = ah3; *(uint4*)(sAl + so + 3 * sstep) = al3; \
    *(uint4*)(sBh + so) = bh0; *(uint4*)(sBl + so) = bl0;                 \
    *(uint4*)(sBh + so + sstep) = bh1; *(uint4*)(sBl + so + sstep) = bl1; \
    *(uint4*)(sBh + so + 2 * sstep) = bh2; *(uint4*)(sBl + so + 2 * sstep) = bl2; \
    *(uint4*)(sBh + so + 3 * sstep) = bh3; *(uint4*)(sBl + so + 3 * sstep) = bl3; \
  }

  f4_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};

  SPLIT_LOAD(0);
  for (int64_t k0 = 0; k0 < Kp; k0 += SB_K) {
    __syncthreads();
    SPLIT_STORE();
    __syncthreads();
    if (k0 + SB_K < Kp) { SPLIT_LOAD(k0 + SB_K); }
#pragma unroll
    for (int ks = 0; ks < SB_K; ks += 32) {
      bf8_t ah[4], al[4], bh[4], bl[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        int ma = (wm * 64 + i * 16 + lr) * SB_LD + ks + lg * 8;
        int nb_ = (wn * 64 + i * 16 + lr) * SB_LD + ks + lg * 8;
        ah[i] = *(const bf8_t*)(sAh + ma);
        al[i] = *(const bf8_t*)(sAl + ma);
        bh[i] = *(const bf8_t*)(sBh + nb_);
        bl[i] = *(const bf8_t*)(sBl + nb_);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], al[i], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[j], ah[i], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], ah[i], acc[i][j], 0, 0, 0);
        }
    }
  }

  // lane holds D[n = 4 lg + r][m = lr] of block (i along m, j along n)
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    int64_t m = m0 + wm * 64 + i * 16 + lr;
    if (m >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int64_t n = n0 + wn * 64 + j * 16 + 4 * lg + r;
        if (n >= N) continue;
        T* c = C + m + n * ldc;
        double v = alpha * (double)acc[i][j][r];
        if (beta != 0.0) v += beta * (double)*c;
        *c = (T)v;
      }
  }
}

// ------------------------------------------------------------ wide GEMM
// The same product on 256 x 256 tiles: one 512-thread workgroup per CU,
// eight waves of 64 (m) x 128 (n), K staged 32 deep through two LDS buffers
// (each 4 planes x 256 rows x 32 bf16; 128 KB in all), one barrier per
// stage. Against the 128 x 128 kernel a wave reads a third fewer LDS bytes
// per MFMA (24 fragment loads per 96 products instead of 16 per 48) and the
// workgroup stages half the operand bytes per product.
//
// A row is 64 bytes, four 16-byte k-chunks, stored at chunk c ^ sw(row) with
// sw(row) = ((row >> 1) ^ (row >> 3)) & 3. A fragment load (16 rows, one
// logical chunk) then covers every bank once per 8 lanes on 32 banks and
// per 16 lanes on 64, and so does a staging store (4 rows x 4 chunks).
constexpr int SW_M = 256, SW_N = 256, SW_K = 32, SW_LD = SW_K;
constexpr int SW_PLANE = 256 * SW_LD;
constexpr int SW_STAGE = 4 * SW_PLANE;
constexpr size_t SW_LDS = 2 * (size_t)SW_STAGE * sizeof(uint16_t);

__device__ __forceinline__ int wide_sw(int row) { return ((row >> 1) ^ (row >> 3)) & 3; }

template <typename T>
__global__ __launch_bounds__(512) void gemm_split_wide_kernel(
    const uint16_t* __restrict__ Ah, const uint16_t* __restrict__ Al, int64_t lda,
    const uint16_t* __restrict__ Bh, const uint16_t* __restrict__ Bl, int64_t ldb, int64_t M,
    int64_t N, int64_t Kp, T* __restrict__ C, int64_t ldc, double alpha, double beta,
    int64_t tilesM, int64_t tilesN) {
  extern __shared__ __attribute__((aligned(16))) uint16_t lds[];

  // XCD-aware grouped tile order (8 XCDs, groups of 8 tile rows)
  int64_t nb = (int64_t)gridDim.x, bid = blockIdx.x;
  int64_t lid = (nb % 8 == 0) ? (bid % 8) * (nb / 8) + bid / 8 : bid;
  const int64_t GM = 8;
  int64_t per_group = GM * tilesN;
  int64_t group = lid / per_group;
  int64_t first = group * GM;
  int64_t gsize = std::min<int64_t>(tilesM - first, GM);
  int64_t in = lid % per_group;
  const int64_t m0 = (first + in % gsize) * SW_M;
  const int64_t n0 = (in / gsize) * SW_N;

  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave & 3, wn = wave >> 2;
  const int lr = lane & 15, lg = lane >> 4;

  // loader: rows t / 4 and t / 4 + 128 of each operand, k-chunk (t % 4) * 8
  const int lrow = t >> 2, lkc = (t & 3) * 8;
  const int64_t ar0 = std::min<int64_t>(m0 + lrow, M - 1) * lda + lkc;
  const int64_t ar1 = std::min<int64_t>(m0 + lrow + 128, M - 1) * lda + lkc;
  const int64_t br0 = std::min<int64_t>(n0 + lrow, N - 1) * ldb + lkc;
  const int64_t br1 = std::min<int64_t>(n0 + lrow + 128, N - 1) * ldb + lkc;
  // the swizzle reads row bits 1..4, so rows lrow and lrow + 128 share it
  const int so0 = lrow * SW_LD + (((t & 3) ^ wide_sw(lrow)) * 8), so1 = so0 + 128 * SW_LD;
  uint4 ah0, ah1, al0, al1, bh0, bh1, bl0, bl1;
#define WIDE_LOAD(k0)                                                               \
  {                                                                                 \
    ah0 = *(const uint4*)(Ah + ar0 + (k0)); ah1 = *(const uint4*)(Ah + ar1 + (k0)); \
    al0 = *(const uint4*)(Al + ar0 + (k0)); al1 = *(const uint4*)(Al + ar1 + (k0)); \
    bh0 = *(const uint4*)(Bh + br0 + (k0)); bh1 = *(const uint4*)(Bh + br1 + (k0)); \
    bl0 = *(const uint4*)(Bl + br0 + (k0)); bl1 = *(const uint4*)(Bl + br1 + (k0)); \
  }
#define WIDE_STORE(s)                                                               \
  {                                                                                 \
    uint16_t* st = lds + (s) * SW_STAGE;                                            \
    *(uint4*)(st + so0) = ah0; *(uint4*)(st + so1) = ah1;                           \
    *(uint4*)(st + SW_PLANE + so0) = al0; *(uint4*)(st + SW_PLANE + so1) = al1;     \
    *(uint4*)(st + 2 * SW_PLANE + so0) = bh0; *(uint4*)(st + 2 * SW_PLANE + so1) = bh1; \
    *(uint4*)(st + 3 * SW_PLANE + so0) = bl0; *(uint4*)(st + 3 * SW_PLANE + so1) = bl1; \
  }

  f4_t acc[4][8];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};

  const int64_t nk = Kp / SW_K;
  WIDE_LOAD(0);
  WIDE_STORE(0);
  __syncthreads();
  for (int64_t kt = 0; kt < nk; ++kt) {
    const int s = (int)(kt & 1);
    if (kt + 1 < nk) { WIDE_LOAD((kt + 1) * SW_K); }
    const uint16_t* st = lds + s * SW_STAGE;
    // Fragments are read one step ahead of the products that use them and
    // in use order (the scheduling barriers keep that order), so that after
    // the stage barrier a wave's first products wait for four reads rather
    // than for the whole set while all eight waves read at once.
    bf8_t ah[4], al[4], bh[2], bl[2];
    auto read_a = [&](int i) {
      const int row = wm * 64 + i * 16 + lr;
      const int ma = row * SW_LD + ((lg ^ wide_sw(row)) * 8);
      ah[i] = *(const bf8_t*)(st + ma);
      al[i] = *(const bf8_t*)(st + SW_PLANE + ma);
    };
    auto read_b = [&](int j) {
      const int row = wn * 128 + j * 16 + lr;
      const int nb_ = row * SW_LD + ((lg ^ wide_sw(row)) * 8);
      bh[j & 1] = *(const bf8_t*)(st + 2 * SW_PLANE + nb_);
      bl[j & 1] = *(const bf8_t*)(st + 3 * SW_PLANE + nb_);
    };
    read_b(0);
    read_a(0);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i < 3) read_a(i + 1); else read_b(1);
      __builtin_amdgcn_sched_barrier(0);
      acc[i][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[0], al[i], acc[i][0], 0, 0, 0);
      acc[i][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[0], ah[i], acc[i][0], 0, 0, 0);
      acc[i][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[0], ah[i], acc[i][0], 0, 0, 0);
    }
#pragma unroll
    for (int j = 1; j < 8; ++j) {
      if (j < 7) read_b(j + 1);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j & 1], al[i], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[j & 1], ah[i], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j & 1], ah[i], acc[i][j], 0, 0, 0);
      }
    }
    // the other buffer was last read before the previous barrier
    if (kt + 1 < nk) { WIDE_STORE(s ^ 1); }
    __syncthreads();
  }
#undef WIDE_LOAD
#undef WIDE_STORE

  // lane holds D[n = 4 lg + r][m = lr] of block (i along m, j along n)
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    int64_t m = m0 + wm * 64 + i * 16 + lr;
    if (m >= M) continue;
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int64_t n = n0 + wn * 128 + j * 16 + 4 * lg + r;
        if (n >= N) continue;
        T* c = C + m + n * ldc;
        double v = alpha * (double)acc[i][j][r];
        if (beta != 0.0) v += beta * (double)*c;
        *c = (T)v;
      }
  }
}

bool lds_ready = false;
bool wide_ok = false;

// One round of the wide kernel (one 256 x 256 tile per CU) takes as long as
// ~1.66 rounds of the 128 x 128 kernel (two tiles per CU) at the same K
// (MI355X, K = 2048 and 10240), so the wide tile is chosen when its rounds,
// so weighted, are fewer: C1 and C4 at chi = 2048 (5 against 10 and 1 against
// 2 rounds) but not C1 at chi = 1024 (2 against 3).
bool use_wide(int64_t tiles_wide, int64_t tiles_narrow, int num_cu) {
  if (num_cu <= 0) num_cu = 256;
  int64_t rw = (tiles_wide + num_cu - 1) / num_cu;
  int64_t rn = (tiles_narrow + 2 * num_cu - 1) / (2 * num_cu);
  return 5 * rw < 3 * rn;
}

}  // namespace

int64_t itg_split_ldk(int64_t K) { return (K + 63) / 64 * 64; }

int itg_split_bf16(itg_ctx* ctx, int dtype, const void* src, int64_t R, int64_t K, int64_t sr,
                   int64_t sk, int64_t ldk, uint16_t* hi, uint16_t* lo) {
  if (R <= 0 || K <= 0) return ITG_OK;
  if (ldk % 64 != 0 || ldk < K) return itg_fail(ctx, ITG_EARG, "split: ldk %lld", (long long)ldk);
  int64_t tilesR = (R + 63) / 64, tilesK = ldk / 64;
  int64_t blocks = tilesR * tilesK;
  if (blocks > 0x7fffffff) return itg_fail(ctx, ITG_EUNSUPPORTED, "split: too large");
  if (dtype == ITG_F64)
    split_kernel<double><<<(unsigned)blocks, 256, 0, ctx->stream>>>(
        (const double*)src, R, K, sr, sk, ldk, tilesK, hi, lo);
  else
    split_kernel<float><<<(unsigned)blocks, 256, 0, ctx->stream>>>(
        (const float*)src, R, K, sr, sk, ldk, tilesK, hi, lo);
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}

int itg_gemm_split(itg_ctx* ctx, int dtype, int64_t M, int64_t N, int64_t K, int64_t Kp,
                   double alpha, const uint16_t* Ah, const uint16_t* Al, int64_t lda,
                   const uint16_t* Bh, const uint16_t* Bl, int64_t ldb, double beta, void* C,
                   int64_t ldc) {
  if (M <= 0 || N <= 0) return ITG_OK;
  if (Kp <= 0 || Kp % SB_K != 0 || lda < Kp || ldb < Kp || lda % 8 != 0 || ldb % 8 != 0)
    return itg_fail(ctx, ITG_EARG, "split gemm: K %lld, lda %lld, ldb %lld", (long long)Kp,
                    (long long)lda, (long long)ldb);
  const size_t lds = (size_t)(2 * SB_M + 2 * SB_N) * SB_LD * sizeof(uint16_t);
  if (!lds_ready) {
    ITG_HIP(ctx, hipFuncSetAttribute((const void*)gemm_split_kernel<double>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    ITG_HIP(ctx, hipFuncSetAttribute((const void*)gemm_split_kernel<float>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    // the wide tile needs 128 KB of a CU's 160 KB of LDS; without it the
    // 128 x 128 kernel serves every shape
    wide_ok = getenv("ITG_SPLIT_WIDE") == nullptr || strcmp(getenv("ITG_SPLIT_WIDE"), "0") != 0;
    if (wide_ok &&
        (hipFuncSetAttribute((const void*)gemm_split_wide_kernel<double>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)SW_LDS) != hipSuccess ||
         hipFuncSetAttribute((const void*)gemm_split_wide_kernel<float>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)SW_LDS) != hipSuccess)) {
      (void)hipGetLastError();
      wide_ok = false;
    }
    lds_ready = true;
  }
  const int64_t wM = (M + SW_M - 1) / SW_M, wN = (N + SW_N - 1) / SW_N;
  if (wide_ok && use_wide(wM * wN, ((M + SB_M - 1) / SB_M) * ((N + SB_N - 1) / SB_N),
                          ctx->num_cu)) {
    if (wM * wN > 0x7fffffff) return itg_fail(ctx, ITG_EUNSUPPORTED, "split gemm: too large");
    itg_tev tk = itg_tevents(ctx, ITG_K_GEMM_SPLIT);
    if (dtype == ITG_F64)
      hipExtLaunchKernelGGL(gemm_split_wide_kernel<double>, dim3((unsigned)(wM * wN)), dim3(512),
                            (uint32_t)SW_LDS, ctx->stream, tk.a, tk.b, 0, Ah, Al, lda, Bh, Bl,
                            ldb, M, N, Kp, (double*)C, ldc, alpha, beta, wM, wN);
    else
      hipExtLaunchKernelGGL(gemm_split_wide_kernel<float>, dim3((unsigned)(wM * wN)), dim3(512),
                            (uint32_t)SW_LDS, ctx->stream, tk.a, tk.b, 0, Ah, Al, lda, Bh, Bl,
                            ldb, M, N, Kp, (float*)C, ldc, alpha, beta, wM, wN);
    ITG_LAUNCH_CHECK(ctx);
    double es = (double)itg_dtype_size(dtype);
    itg_tcommit(ctx, tk, ITG_K_GEMM_SPLIT, 2.0 * M * N * K,
                4.0 * (M + N) * K + es * M * N * (beta != 0.0 ? 2.0 : 1.0));
    return ITG_OK;
  }
  int64_t tilesM = (M + SB_M - 1) / SB_M, tilesN = (N + SB_N - 1) / SB_N;
  int64_t tiles = tilesM * tilesN;
  if (tiles > 0x7fffffff) return itg_fail(ctx, ITG_EUNSUPPORTED, "split gemm: too large");
  itg_tev tk = itg_tevents(ctx, ITG_K_GEMM_SPLIT);
  if (dtype == ITG_F64)
    hipExtLaunchKernelGGL(gemm_split_kernel<double>, dim3((unsigned)tiles), dim3(256),
                          (uint32_t)lds, ctx->stream, tk.a, tk.b, 0, Ah, Al, lda, Bh, Bl, ldb, M,
                          N, Kp, (double*)C, ldc, alpha, beta, tilesM, tilesN);
  else
    hipExtLaunchKernelGGL(gemm_split_kernel<float>, dim3((unsigned)tiles), dim3(256),
                          (uint32_t)lds, ctx->stream, tk.a, tk.b, 0, Ah, Al, lda, Bh, Bl, ldb, M,
                          N, Kp, (float*)C, ldc, alpha, beta, tilesM, tilesN);
  ITG_LAUNCH_CHECK(ctx);
  // algorithmic: 2MNK flops; split planes read once, C written (and read)
  double es = (double)itg_dtype_size(dtype);
  itg_tcommit(ctx, tk, ITG_K_GEMM_SPLIT, 2.0 * M * N * K,
            4.0 * (M + N) * K + es * M * N * (beta != 0.0 ? 2.0 : 1.0));
  return ITG_OK;
}
