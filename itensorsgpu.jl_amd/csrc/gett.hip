// Tensor contraction C = alpha * A.B + beta * C on gfx950 MFMA, with every
// index permutation folded into the operand loads and the epilogue.
//
// The contraction is reduced to GEMM coordinates (m, n, k): m runs over the
// free modes of A, n over those of B, k over the contracted modes. The
// address of A(m,k) is offAm[m] + offAk[k] (likewise B, C); each offset is
// either affine (one merged mode: idx * stride) or read from a small table
// built on the host and cached per contraction signature, the analogue of
// the reference's ContractionPlans cache (src/ITensorsGPU.jl:55,
// src/tensor/cudense.jl:285-326). Tables replace the permutedims passes of
// the TTGT route (cudense.jl:170-236): tiles are gathered straight from the
// tensors into LDS, in whichever direction (m/n or k) is contiguous in HBM.
//
// Tile: 128x128 per 256-thread workgroup, 2x2 waves of 64x64, each wave a
// 4x4 grid of 16x16 MFMA blocks (FP64 launches with fewer tiles than
// workgroup slots: 64x64, 2x2 waves of 32x32).
//   FP64   : v_mfma_f64_16x16x4_f64, FP64 accumulate          (BK = 16)
//   FP32   : v_mfma_f32_16x16x4_f32, FP32 accumulate          (BK = 16)
//   BF16x3 : a = ah + al, b = bh + bl (bf16, round-to-nearest-even);
//            ah*bh + ah*bl + al*bh on v_mfma_f32_16x16x32_bf16,
//            FP32 accumulate (Henry, Tang & Heinecke, ARITH 2019) (BK = 32)
// Workgroup ids are remapped so that each XCD (ids congruent mod 8) works on
// a contiguous group of tiles sharing A/B panels in its own L2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <type_traits>

#include "itg_internal.h"

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));

#include "gett.h"

__device__ __forceinline__ int64_t goff(const int64_t* t, int64_t s, int64_t i) {
  return s >= 0 ? i * s : t[i];
}

__device__ __forceinline__ uint16_t f2bf_rne(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

template <typename T, int MODE, int TS = 128>
struct GettTraits {
  static constexpr bool BF = (MODE == ITG_MODE_BF16X3);
  static constexpr bool F64 = (MODE == ITG_MODE_NATIVE) && std::is_same<T, double>::value;
  using CT = typename std::conditional<F64, double, float>::type;
  static constexpr int BM = TS, BN = TS;   // 128, or 64 for FP64 launches with few tiles
  static constexpr int BKT = BF ? 32 : 16;
  static constexpr int NPT = BM * BKT / 256;   // elements per thread per operand
};

// map workgroup id -> (tm, tn) with XCD-aware placement and grouped order
__device__ __forceinline__ void tile_coords(int64_t tilesM, int64_t tilesN, int64_t& tm,
                                            int64_t& tn) {
  int64_t nb = (int64_t)gridDim.x, bid = blockIdx.x;
  int64_t lid = bid;
  if (nb % 8 == 0) lid = (bid % 8) * (nb / 8) + bid / 8;
  const int64_t GM = 8;
  int64_t per_group = GM * tilesN;
  int64_t group = lid / per_group;
  int64_t first = group * GM;
  int64_t gsize = tilesM - first < GM ? tilesM - first : GM;
  int64_t in = lid % per_group;
  tm = first + in % gsize;
  tn = in / gsize;
}

template <typename T, int MODE, bool AK, bool BKF, int TS = 128>
__global__ __launch_bounds__(256, 2) void gett_kernel(GettArgs g) {
  using TR = GettTraits<T, MODE, TS>;
  using CT = typename TR::CT;
  constexpr int BM = TR::BM, BN = TR::BN, BKT = TR::BKT, NPT = TR::NPT;
  // four waves in 2 x 2, each a WH x WH sub-tile of WT x WT 16 x 16 blocks
  constexpr int WH = BM / 2, WT = BM / 32;
  constexpr bool BF = TR::BF;
  constexpr int PAD = BF ? 8 : 1;

  // LDS: FP64/FP32 k-major [k][m]; BF16 m-major [m][k] (8 consecutive k per lane)
  __shared__ __attribute__((aligned(16))) char smem[BF ? (4 * BM * (BKT + PAD) * 2)
                                                       : (2 * BKT * (BM + PAD) * sizeof(CT))];
  CT* As = (CT*)smem;
  CT* Bs = As + BKT * (BM + PAD);
  uint16_t* Ah = (uint16_t*)smem;
  uint16_t* Al = Ah + BM * (BKT + PAD);
  uint16_t* Bh = Al + BM * (BKT + PAD);
  uint16_t* Bl = Bh + BN * (BKT + PAD);

  const T* __restrict__ A = (const T*)g.A;
  const T* __restrict__ B = (const T*)g.B;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  int64_t tm, tn;
  tile_coords(g.tilesM, g.tilesN, tm, tn);
  if (g.band && tm > tn + 1) return;   // whole workgroup, before any barrier
  const int64_t m0 = tm * BM, n0 = tn * BN;
  const int64_t kbeg = (int64_t)blockIdx.z * g.kchunk;
  const int64_t kend = (kbeg + g.kchunk < g.K) ? kbeg + g.kchunk : g.K;

  // fixed per-thread coordinates for the loaders
  // A: AK ? (k = t % BKT, m = t / BKT + (256/BKT)*i) : (m = t % BM, k = t / BM + 2*i)
  constexpr int ASTEP = AK ? 256 / BKT : 256 / BM;
  constexpr int BSTEP = BKF ? 256 / BKT : 256 / BN;
  const int a_fix = AK ? (t % BKT) : (t % BM);
  const int a_var0 = AK ? (t / BKT) : (t / BM);
  const int b_fix = BKF ? (t % BKT) : (t % BN);
  const int b_var0 = BKF ? (t / BKT) : (t / BN);

  // offsets constant over the K loop
  int64_t a_offm[AK ? NPT : 1];
  int64_t b_offn[BKF ? NPT : 1];
  bool a_okm[AK ? NPT : 1];
  bool b_okn[BKF ? NPT : 1];
  if (AK) {
#pragma unroll
    for (int i = 0; i < NPT; ++i) {
      int64_t m = m0 + a_var0 + ASTEP * i;
      a_okm[i] = m < g.M;
      a_offm[i] = a_okm[i] ? goff(g.tAm, g.sAm, m) : 0;
    }
  } else {
    int64_t m = m0 + a_fix;
    a_okm[0] = m < g.M;
    a_offm[0] = a_okm[0] ? goff(g.tAm, g.sAm, m) : 0;
  }
  if (BKF) {
#pragma unroll
    for (int i = 0; i < NPT; ++i) {
      int64_t n = n0 + b_var0 + BSTEP * i;
      b_okn[i] = n < g.N;
      b_offn[i] = b_okn[i] ? goff(g.tBn, g.sBn, n) : 0;
    }
  } else {
    int64_t n = n0 + b_fix;
    b_okn[0] = n < g.N;
    b_offn[0] = b_okn[0] ? goff(g.tBn, g.sBn, n) : 0;
  }

  T ra[NPT], rb[NPT];
  auto load = [&](int64_t k0) {
    if (AK) {
      int64_t k = k0 + a_fix;
      bool okk = k < kend;
      int64_t ok_ = okk ? goff(g.tAk, g.sAk, k) : 0;
#pragma unroll
      for (int i = 0; i < NPT; ++i) ra[i] = (okk && a_okm[i]) ? A[a_offm[i] + ok_] : (T)0;
    } else {
#pragma unroll
      for (int i = 0; i < NPT; ++i) {
        int64_t k = k0 + a_var0 + ASTEP * i;
        bool ok = a_okm[0] && k < kend;
        ra[i] = ok ? A[a_offm[0] + goff(g.tAk, g.sAk, k)] : (T)0;
      }
    }
    if (BKF) {
      int64_t k = k0 + b_fix;
      bool okk = k < kend;
      int64_t ok_ = okk ? goff(g.tBk, g.sBk, k) : 0;
#pragma unroll
      for (int i = 0; i < NPT; ++i) rb[i] = (okk && b_okn[i]) ? B[b_offn[i] + ok_] : (T)0;
    } else {
#pragma unroll
      for (int i = 0; i < NPT; ++i) {
        int64_t k = k0 + b_var0 + BSTEP * i;
        bool ok = b_okn[0] && k < kend;
        rb[i] = ok ? B[b_offn[0] + goff(g.tBk, g.sBk, k)] : (T)0;
      }
    }
  };

  auto store = [&]() {
#pragma unroll
    for (int i = 0; i < NPT; ++i) {
      int am = AK ? a_var0 + ASTEP * i : a_fix;
      int ak = AK ? a_fix : a_var0 + ASTEP * i;
      int bn = BKF ? b_var0 + BSTEP * i : b_fix;
      int bk = BKF ? b_fix : b_var0 + BSTEP * i;
      if constexpr (BF) {
        float xa = (float)ra[i];
        uint16_t ha = f2bf_rne(xa);
        Ah[am * (BKT + PAD) + ak] = ha;
        Al[am * (BKT + PAD) + ak] = f2bf_rne(xa - bf2f(ha));
        float xb = (float)rb[i];
        uint16_t hb = f2bf_rne(xb);
        Bh[bn * (BKT + PAD) + bk] = hb;
        Bl[bn * (BKT + PAD) + bk] = f2bf_rne(xb - bf2f(hb));
      } else {
        As[ak * (BM + PAD) + am] = (CT)ra[i];
        Bs[bk * (BN + PAD) + bn] = (CT)rb[i];
      }
    }
  };

  typedef typename std::conditional<TR::F64, d4_t, f4_t>::type acc_t;
  acc_t acc[WT][WT];
#pragma unroll
  for (int i = 0; i < WT; ++i)
#pragma unroll
    for (int j = 0; j < WT; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  const int lr = lane & 15, lg = lane >> 4;
  if (kbeg < kend) load(kbeg);
  for (int64_t k0 = kbeg; k0 < kend; k0 += BKT) {
    __syncthreads();
    store();
    __syncthreads();
    if (k0 + BKT < kend) load(k0 + BKT);
    if constexpr (BF) {
      bf8_t ah[WT], al[WT], bh[WT], bl[WT];
#pragma unroll
      for (int i = 0; i < WT; ++i) {
        int m = wm * WH + i * 16 + lr;
        ah[i] = *(const bf8_t*)&Ah[m * (BKT + PAD) + lg * 8];
        al[i] = *(const bf8_t*)&Al[m * (BKT + PAD) + lg * 8];
        int n = wn * WH + i * 16 + lr;
        bh[i] = *(const bf8_t*)&Bh[n * (BKT + PAD) + lg * 8];
        bl[i] = *(const bf8_t*)&Bl[n * (BKT + PAD) + lg * 8];
      }
#pragma unroll
      for (int i = 0; i < WT; ++i)
#pragma unroll
        for (int j = 0; j < WT; ++j) {
          if constexpr (BF) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
          }
        }
    } else {
#pragma unroll
      for (int kk = 0; kk < BKT / 4; ++kk) {
        CT a[WT], b[WT];
#pragma unroll
        for (int i = 0; i < WT; ++i) {
          a[i] = As[(kk * 4 + lg) * (BM + PAD) + wm * WH + i * 16 + lr];
          b[i] = Bs[(kk * 4 + lg) * (BN + PAD) + wn * WH + i * 16 + lr];
        }
#pragma unroll
        for (int i = 0; i < WT; ++i)
#pragma unroll
          for (int j = 0; j < WT; ++j) {
            if constexpr (TR::F64)
              acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
            else if constexpr (!BF)
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
          }
      }
    }
  }

  // epilogue: lane holds D[row][lr] of each 16x16 block, with row = 4*lg + r
  // for the FP32-accumulating MFMAs (4-row groups) and row = 4*r + lg for
  // v_mfma_f64_16x16x4_f64 (1-row groups)
  if (g.partial) {
    CT* P = (CT*)g.partial + (int64_t)blockIdx.z * g.M * g.N;
#pragma unroll
    for (int i = 0; i < WT; ++i)
#pragma unroll
      for (int j = 0; j < WT; ++j) {
        int64_t n = n0 + wn * WH + j * 16 + lr;
        if (n >= g.N) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int64_t m = m0 + wm * WH + i * 16 + (TR::F64 ? 4 * r + lg : 4 * lg + r);
          if (m < g.M) P[n * g.M + m] = acc[i][j][r];
        }
      }
    return;
  }
  T* C = (T*)g.C;
  if (g.sCm == 1 && g.sCn >= 0 && m0 + BM <= g.M && n0 + BN <= g.N) {
    // interior tile of a column-major C: one column pointer per j and
    // compile-time row offsets, no per-element bounds or offset arithmetic
    // (a short epilogue: this code runs once per tile, from a cold
    // instruction cache on small launches)
    T* cb = C + (n0 + wn * WH + lr) * g.sCn + m0 + wm * WH + (TR::F64 ? lg : 4 * lg);
    const double al = g.alpha, be = g.beta;
    if (be != 0.0) {
#pragma unroll
      for (int j = 0; j < WT; ++j) {
        T* cj = cb + (int64_t)(j * 16) * g.sCn;
#pragma unroll
        for (int i = 0; i < WT; ++i)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int mo = i * 16 + (TR::F64 ? 4 * r : r);
            cj[mo] = (T)(al * (double)acc[i][j][r] + be * (double)cj[mo]);
          }
      }
    } else {
#pragma unroll
      for (int j = 0; j < WT; ++j) {
        T* cj = cb + (int64_t)(j * 16) * g.sCn;
#pragma unroll
        for (int i = 0; i < WT; ++i)
#pragma unroll
          for (int r = 0; r < 4; ++r) cj[i * 16 + (TR::F64 ? 4 * r : r)] =
              (T)(al * (double)acc[i][j][r]);
      }
    }
    return;
  }
  if (g.sCn == 1 && g.sCm >= 0 && m0 + BM <= g.M && n0 + BN <= g.N) {
    // the same for a row-major C (itg_gemm's C^T): one row pointer per
    // (i, r), compile-time column offsets
    T* cb = C + (m0 + wm * WH + (TR::F64 ? lg : 4 * lg)) * g.sCm + n0 + wn * WH + lr;
    const double al = g.alpha, be = g.beta;
#pragma unroll
    for (int i = 0; i < WT; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        T* ci = cb + (int64_t)(i * 16 + (TR::F64 ? 4 * r : r)) * g.sCm;
        if (be != 0.0) {
#pragma unroll
          for (int j = 0; j < WT; ++j)
            ci[j * 16] = (T)(al * (double)acc[i][j][r] + be * (double)ci[j * 16]);
        } else {
#pragma unroll
          for (int j = 0; j < WT; ++j) ci[j * 16] = (T)(al * (double)acc[i][j][r]);
        }
      }
    return;
  }
#pragma unroll
  for (int i = 0; i < WT; ++i)
#pragma unroll
    for (int j = 0; j < WT; ++j) {
      int64_t n = n0 + wn * WH + j * 16 + lr;
      if (n >= g.N) continue;
      int64_t on = goff(g.tCn, g.sCn, n);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int64_t m = m0 + wm * WH + i * 16 + (TR::F64 ? 4 * r + lg : 4 * lg + r);
        if (m >= g.M) continue;
        int64_t o = on + goff(g.tCm, g.sCm, m);
        double v = g.alpha * (double)acc[i][j][r];
        if (g.beta != 0.0) v += g.beta * (double)C[o];
        C[o] = (T)v;
      }
    }
}

// split-K reduction + epilogue
template <typename T, typename CT>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(GettArgs g, int nsplit) {
  int64_t MN = g.M * g.N;
  const CT* P = (const CT*)g.partial;
  T* C = (T*)g.C;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < MN;
       e += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int z = 0; z < nsplit; ++z) s += (double)P[z * MN + e];
    int64_t m = e % g.M, n = e / g.M;
    int64_t o = goff(g.tCm, g.sCm, m) + goff(g.tCn, g.sCn, n);
    double v = g.alpha * s;
    if (g.beta != 0.0) v += g.beta * (double)C[o];
    C[o] = (T)v;
  }
}

// Contractions with at most 16 outputs (scalars, short vectors; dot-like):
// every block reduces a K-range for all outputs in FP64, then one reduction.
template <typename T>
__global__ __launch_bounds__(256) void small_out_kernel(GettArgs g, double* part) {
  const T* A = (const T*)g.A;
  const T* B = (const T*)g.B;
  int mn = (int)(g.M * g.N);
  double acc[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) acc[e] = 0.0;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < g.K;
       k += (int64_t)gridDim.x * blockDim.x) {
    int64_t ak = goff(g.tAk, g.sAk, k), bk = goff(g.tBk, g.sBk, k);
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      if (e < mn) {
        int64_t m = e % g.M, n = e / g.M;
        acc[e] += (double)A[goff(g.tAm, g.sAm, m) + ak] * (double)B[bk + goff(g.tBn, g.sBn, n)];
      }
    }
  }
  __shared__ double sh[4][16];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    double v = acc[e];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) sh[w][e] = v;
  }
  __syncthreads();
  if (threadIdx.x < 16) {
    int e = threadIdx.x;
    part[blockIdx.x * 16 + e] = sh[0][e] + sh[1][e] + sh[2][e] + sh[3][e];
  }
}

template <typename T>
__global__ void small_out_finish(GettArgs g, const double* part, int nblk) {
  int e = threadIdx.x;
  int mn = (int)(g.M * g.N);
  if (e >= mn) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += part[b * 16 + e];
  int64_t m = e % g.M, n = e / g.M;
  T* C = (T*)g.C;
  int64_t o = goff(g.tCm, g.sCm, m) + goff(g.tCn, g.sCn, n);
  double v = g.alpha * s;
  if (g.beta != 0.0) v += g.beta * (double)C[o];
  C[o] = (T)v;
}

// ======================================================================
// host side: plans
// A contraction with one small operand: the big operand X carries the free
// modes F (at most 3) and the summed modes K, the small one Y the free modes
// G and K, with K and G small (K <= 64, G <= 32, K G <= 1024). It streams X
// once through sic_kernel, which forms every offset from the modes' extents
// and strides: no offset tables (which for a non-mergeable F would have F
// entries, built on the host for every new shape).
constexpr int SIC_KMAX = 64, SIC_GMAX = 32;
constexpr int64_t SIC_FMIN = 4096;   // smaller F: the GEMM kernel is fine
struct SicArgs {
  const void* X;
  const void* Y;
  void* C;
  int64_t F;
  int nf, ng, nk, G, K;
  int64_t fd[3], fx[3], fc[3];   // F: extent, stride in X, stride in C
  int64_t gd[4], gy[4], gc[4];   // G: extent, stride in Y, stride in C
  int64_t kd[4], kx[4], ky[4];   // K: extent, stride in X, stride in Y
  double alpha, beta;
};

struct itg_plan {
  int64_t M = 1, N = 1, K = 1;
  bool swap = false;
  int64_t sAm = 0, sAk = 0, sBk = 0, sBn = 0, sCm = 0, sCn = 0;
  int64_t* dtab = nullptr;   // device tables (concatenated)
  int64_t oAm = -1, oAk = -1, oBk = -1, oBn = -1, oCm = -1, oCn = -1;  // offsets into dtab
  bool akfast = false, bkfast = false;
  size_t bytes = 0;
  // small-operand form (no tables): X is the A-role operand if sic_xa
  bool sic = false, sic_xa = true;
  SicArgs sa = {};
};

template <typename T, int GM>
__global__ __launch_bounds__(256) void sic_kernel(SicArgs a) {
  __shared__ T ys[SIC_KMAX * SIC_GMAX];   // Y(k, g) at k G + g
  __shared__ int64_t kox[SIC_KMAX];       // offset of k in X
  __shared__ int64_t goc[SIC_GMAX];       // offset of g in C
  const int t = threadIdx.x;
  const T* __restrict__ Y = (const T*)a.Y;
  for (int e = t; e < a.K * a.G; e += 256) {
    int64_t rk = e / a.G, rg = e % a.G, o = 0;
    for (int q = 0; q < a.nk; ++q) { o += (rk % a.kd[q]) * a.ky[q]; rk /= a.kd[q]; }
    for (int q = 0; q < a.ng; ++q) { o += (rg % a.gd[q]) * a.gy[q]; rg /= a.gd[q]; }
    ys[e] = Y[o];
  }
  for (int k = t; k < a.K; k += 256) {
    int64_t r = k, o = 0;
    for (int q = 0; q < a.nk; ++q) { o += (r % a.kd[q]) * a.kx[q]; r /= a.kd[q]; }
    kox[k] = o;
  }
  for (int g = t; g < a.G; g += 256) {
    int64_t r = g, o = 0;
    for (int q = 0; q < a.ng; ++q) { o += (r % a.gd[q]) * a.gc[q]; r /= a.gd[q]; }
    goc[g] = o;
  }
  __syncthreads();
  const int64_t f = (int64_t)blockIdx.x * 256 + t;
  if (f >= a.F) return;
  int64_t ox = 0, oc = 0, r = f;
  for (int q = 0; q < a.nf; ++q) {
    const int64_t i = r % a.fd[q];
    r /= a.fd[q];
    ox += i * a.fx[q];
    oc += i * a.fc[q];
  }
  const T* __restrict__ X = (const T*)a.X + ox;
  T acc[GM];
#pragma unroll
  for (int g = 0; g < GM; ++g) acc[g] = (T)0;
#pragma unroll 4
  for (int k = 0; k < a.K; ++k) {
    const T xv = X[kox[k]];
#pragma unroll
    for (int g = 0; g < GM; ++g)
      if (g < a.G) acc[g] += xv * ys[k * a.G + g];
  }
  T* __restrict__ C = (T*)a.C + oc;
#pragma unroll
  for (int g = 0; g < GM; ++g) {
    if (g < a.G) {
      T v = (T)a.alpha * acc[g];
      if (a.beta != 0.0) v += (T)a.beta * C[goc[g]];
      C[goc[g]] = v;
    }
  }
}

static int launch_sic(itg_ctx* ctx, int dtype, const itg_plan* p, const void* A, const void* B,
                      void* C, double alpha, double beta) {
  SicArgs a = p->sa;
  a.X = p->sic_xa ? A : B;
  a.Y = p->sic_xa ? B : A;
  a.C = C;
  a.alpha = alpha;
  a.beta = beta;
  const int64_t nblk = (a.F + 255) / 256;
  if (nblk > 0x7fffffff) return itg_fail(ctx, ITG_EUNSUPPORTED, "contraction too large");
  const int tk = itg_tstart(ctx, ITG_K_GETT);
  const bool small = a.G <= 16;
  if (dtype == ITG_F64) {
    if (small) sic_kernel<double, 16><<<(unsigned)nblk, 256, 0, ctx->stream>>>(a);
    else sic_kernel<double, SIC_GMAX><<<(unsigned)nblk, 256, 0, ctx->stream>>>(a);
  } else {
    if (small) sic_kernel<float, 16><<<(unsigned)nblk, 256, 0, ctx->stream>>>(a);
    else sic_kernel<float, SIC_GMAX><<<(unsigned)nblk, 256, 0, ctx->stream>>>(a);
  }
  ITG_LAUNCH_CHECK(ctx);
  const double es = dtype == ITG_F64 ? 8.0 : 4.0;
  itg_tstop(ctx, tk, ITG_K_GETT, 2.0 * a.F * a.G * a.K,
            es * (a.F * a.K + a.F * a.G * (beta != 0.0 ? 2 : 1)));
  return ITG_OK;
}

void itg_plan_cache_clear(itg_ctx* ctx) {
  for (auto& kv : ctx->plans) {
    if (kv.second->dtab) hipFree(kv.second->dtab);
    delete kv.second;
  }
  ctx->plans.clear();
  ctx->plan_bytes = 0;
}

namespace {
struct Mode {
  int64_t d;
  int64_t s1, s2;  // strides in the two tensors that carry this mode
};

// order by s1 ascending and merge modes that are contiguous continuations in
// both tensors
std::vector<Mode> canon(std::vector<Mode> v, bool by_second) {
  std::vector<Mode> w;
  for (auto& x : v)
    if (x.d != 1) w.push_back(x);
  std::stable_sort(w.begin(), w.end(), [&](const Mode& a, const Mode& b) {
    return by_second ? a.s2 < b.s2 : a.s1 < b.s1;
  });
  std::vector<Mode> out;
  for (auto& x : w) {
    if (!out.empty() && out.back().s1 * out.back().d == x.s1 &&
        out.back().s2 * out.back().d == x.s2) {
      out.back().d *= x.d;
    } else {
      out.push_back(x);
    }
  }
  return out;
}

// offsets of a group for tensor `which` (1 or 2); returns stride or -1 with table
int64_t group_offsets(const std::vector<Mode>& g, int which, std::vector<int64_t>& tab) {
  if (g.empty()) return 0;
  if (g.size() == 1) return which == 1 ? g[0].s1 : g[0].s2;
  int64_t n = 1;
  for (auto& x : g) n *= x.d;
  tab.resize(n);
  for (int64_t i = 0; i < n; ++i) {
    int64_t rem = i, o = 0;
    for (auto& x : g) {
      int64_t j = rem % x.d;
      rem /= x.d;
      o += j * (which == 1 ? x.s1 : x.s2);
    }
    tab[i] = o;
  }
  return -1;
}
}  // namespace

static int build_plan(itg_ctx* ctx, const std::string& key, int rA, const int64_t* dA,
                      const int32_t* lA, int rB, const int64_t* dB, const int32_t* lB, int rC,
                      const int64_t* dC, const int32_t* lC, itg_plan** out) {
  // strides
  int64_t sA[ITG_MAX_RANK], sB[ITG_MAX_RANK], sC[ITG_MAX_RANK];
  int64_t s = 1;
  for (int i = 0; i < rA; ++i) { sA[i] = s; s *= dA[i]; }
  s = 1;
  for (int i = 0; i < rB; ++i) { sB[i] = s; s *= dB[i]; }
  s = 1;
  for (int i = 0; i < rC; ++i) { sC[i] = s; s *= dC[i]; }
  auto find = [](const int32_t* l, int r, int32_t x) {
    for (int i = 0; i < r; ++i) if (l[i] == x) return i;
    return -1;
  };
  // C's fastest non-unit mode decides the operand roles: it must be an n-mode
  bool swap = false;
  for (int i = 0; i < rC; ++i) {
    if (dC[i] == 1) continue;
    swap = find(lA, rA, lC[i]) >= 0;
    break;
  }
  const int32_t *la = lA, *lb = lB;
  const int64_t *da = dA, *db = dB, *psA = sA, *psB = sB;
  int ra = rA, rb = rB;
  if (swap) {
    std::swap(la, lb); std::swap(da, db); std::swap(psA, psB); std::swap(ra, rb);
  }
  std::vector<Mode> gm, gn, gk;
  for (int i = 0; i < ra; ++i) {
    int jb = find(lb, rb, la[i]);
    if (jb >= 0) {
      gk.push_back({da[i], psA[i], psB[jb]});
    } else {
      int jc = find(lC, rC, la[i]);
      gm.push_back({da[i], psA[i], sC[jc]});
    }
  }
  for (int i = 0; i < rb; ++i) {
    if (find(la, ra, lb[i]) >= 0) continue;
    int jc = find(lC, rC, lb[i]);
    gn.push_back({db[i], psB[i], sC[jc]});
  }
  // K order: by A's strides if A's unit-stride mode is contracted, else by B's
  bool a_unit_in_k = false, b_unit_in_k = false;
  for (auto& x : gk) {
    if (x.d > 1 && x.s1 == 1) a_unit_in_k = true;
    if (x.d > 1 && x.s2 == 1) b_unit_in_k = true;
  }
  std::vector<Mode> cm = canon(gm, false), cn = canon(gn, false);
  std::vector<Mode> ck = canon(gk, !a_unit_in_k && b_unit_in_k);
  itg_plan* p = new itg_plan();
  p->swap = swap;
  for (auto& x : cm) p->M *= x.d;
  for (auto& x : cn) p->N *= x.d;
  for (auto& x : ck) p->K *= x.d;
  // the small-operand form: the big free group of at most 3 modes, the
  // other free group and K small
  if (ck.size() <= 4 && p->K <= SIC_KMAX) {
    const bool xa = cm.size() <= 3 && cn.size() <= 4 && p->N <= SIC_GMAX &&
                    p->N * p->K <= SIC_KMAX * SIC_GMAX && p->M >= SIC_FMIN;
    const bool xb = !xa && cn.size() <= 3 && cm.size() <= 4 && p->M <= SIC_GMAX &&
                    p->M * p->K <= SIC_KMAX * SIC_GMAX && p->N >= SIC_FMIN;
    if (xa || xb) {
      // Mode strides: M {A, C}, N {B, C}, K {A, B}
      const std::vector<Mode>& F = xa ? cm : cn;
      const std::vector<Mode>& G = xa ? cn : cm;
      SicArgs& a = p->sa;
      a.F = xa ? p->M : p->N;
      a.G = (int)(xa ? p->N : p->M);
      a.K = (int)p->K;
      a.nf = (int)F.size();
      a.ng = (int)G.size();
      a.nk = (int)ck.size();
      for (int q = 0; q < a.nf; ++q) { a.fd[q] = F[q].d; a.fx[q] = F[q].s1; a.fc[q] = F[q].s2; }
      for (int q = 0; q < a.ng; ++q) { a.gd[q] = G[q].d; a.gy[q] = G[q].s1; a.gc[q] = G[q].s2; }
      for (int q = 0; q < a.nk; ++q) {
        a.kd[q] = ck[q].d;
        a.kx[q] = xa ? ck[q].s1 : ck[q].s2;
        a.ky[q] = xa ? ck[q].s2 : ck[q].s1;
      }
      p->sic = true;
      p->sic_xa = xa;
      if (ctx->plans.size() > 4096) {
        hipStreamSynchronize(ctx->stream);
        itg_plan_cache_clear(ctx);
      }
      ctx->plans[key] = p;
      *out = p;
      return ITG_OK;
    }
  }
  std::vector<int64_t> tAm, tAk, tBk, tBn, tCm, tCn;
  p->sAm = group_offsets(cm, 1, tAm);
  p->sCm = group_offsets(cm, 2, tCm);
  p->sBn = group_offsets(cn, 1, tBn);
  p->sCn = group_offsets(cn, 2, tCn);
  p->sAk = group_offsets(ck, 1, tAk);
  p->sBk = group_offsets(ck, 2, tBk);
  // load directions: contiguous along the free index if its first mode has
  // unit stride, else along k if k's first mode has unit stride
  p->akfast = !(!cm.empty() && cm[0].s1 == 1) && (!ck.empty() && ck[0].s1 == 1);
  p->bkfast = !(!cn.empty() && cn[0].s1 == 1) && (!ck.empty() && ck[0].s2 == 1);
  size_t total = tAm.size() + tAk.size() + tBk.size() + tBn.size() + tCm.size() + tCn.size();
  if (total) {
    std::vector<int64_t> all;
    all.reserve(total);
    auto put = [&](std::vector<int64_t>& v, int64_t& o) {
      if (!v.empty()) { o = (int64_t)all.size(); all.insert(all.end(), v.begin(), v.end()); }
    };
    put(tAm, p->oAm); put(tAk, p->oAk); put(tBk, p->oBk);
    put(tBn, p->oBn); put(tCm, p->oCm); put(tCn, p->oCn);
    if (hipMalloc(&p->dtab, total * sizeof(int64_t)) != hipSuccess) {
      delete p;
      return itg_fail(ctx, ITG_EOOM, "plan table allocation failed");
    }
    hipError_t e = hipMemcpyAsync(p->dtab, all.data(), total * sizeof(int64_t),
                                  hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
      hipFree(p->dtab);
      delete p;
      return itg_fail(ctx, ITG_EHIP, "plan upload failed: %s", hipGetErrorString(e));
    }
    p->bytes = total * sizeof(int64_t);
  }
  // bound the cache
  if (ctx->plans.size() > 4096 || ctx->plan_bytes + p->bytes > ((size_t)1 << 30)) {
    hipStreamSynchronize(ctx->stream);
    itg_plan_cache_clear(ctx);
  }
  ctx->plans[key] = p;
  ctx->plan_bytes += p->bytes;
  *out = p;
  return ITG_OK;
}

template <typename T, int MODE, int TS = 128>
static int launch_gett_t(itg_ctx* ctx, GettArgs& g, bool ak, bool bk) {
  using TR = GettTraits<T, MODE, TS>;
  using CT = typename TR::CT;
  g.tilesM = (g.M + TR::BM - 1) / TR::BM;
  g.tilesN = (g.N + TR::BN - 1) / TR::BN;
  int64_t tiles = g.tilesM * g.tilesN;
  if (tiles > 0x7fffffff) return itg_fail(ctx, ITG_EUNSUPPORTED, "contraction too large");
  if constexpr (TR::F64 && TS == 128) {
    // FP64 with fewer 128 x 128 tiles than workgroup slots, or a short K:
    // 64 x 64 tiles give each wave a quarter of the MFMA chain per k step
    // and four times the workgroups, more of them resident per CU (the
    // K = 64 trailing updates of the eigensolver)
    if (tiles < 2 * (int64_t)std::max(ctx->num_cu, 1) || g.K <= 128)
      return launch_gett_t<T, MODE, 64>(ctx, g, ak, bk);
  }
  // split K when the output has too few tiles to fill the chip
  int nsplit = 1;
  int64_t target = 2 * (int64_t)std::max(ctx->num_cu, 1);
  if (!g.band && tiles < target && g.K >= TR::BKT * 16) {
    // at least 4 k-tiles per split
    int64_t want = (target + tiles - 1) / tiles;
    int64_t maxs = g.K / (TR::BKT * 4);
    nsplit = (int)std::max<int64_t>(1, std::min<int64_t>({want, maxs, 64}));
  }
  if (nsplit > 1 && ctx->fixed_ws) {
    // a no-alloc call: no more splits than its fixed scratch holds
    const size_t per = (size_t)g.M * g.N * sizeof(CT);
    nsplit = (int)std::max<size_t>(1, std::min<size_t>((size_t)nsplit, ctx->fixed_ws_bytes / per));
  }
  g.partial = nullptr;
  g.kchunk = g.K;
  if (nsplit > 1) {
    int64_t chunk = (g.K + nsplit - 1) / nsplit;
    chunk = (chunk + TR::BKT - 1) / TR::BKT * TR::BKT;
    nsplit = (int)((g.K + chunk - 1) / chunk);
    g.kchunk = chunk;
    void* ws;
    if (ctx->fixed_ws) {
      ws = ctx->fixed_ws;
    } else {
      int rc = itg_workspace(ctx, (size_t)nsplit * g.M * g.N * sizeof(CT), &ws);
      if (rc) return rc;
    }
    g.partial = ws;
  }
  dim3 grid((unsigned)tiles, 1, (unsigned)nsplit);
  int tk = itg_tstart(ctx, ITG_K_GETT);
  if (ak && bk) gett_kernel<T, MODE, true, true, TS><<<grid, 256, 0, ctx->stream>>>(g);
  else if (ak) gett_kernel<T, MODE, true, false, TS><<<grid, 256, 0, ctx->stream>>>(g);
  else if (bk) gett_kernel<T, MODE, false, true, TS><<<grid, 256, 0, ctx->stream>>>(g);
  else gett_kernel<T, MODE, false, false, TS><<<grid, 256, 0, ctx->stream>>>(g);
  ITG_LAUNCH_CHECK(ctx);
  itg_tstop(ctx, tk, ITG_K_GETT, 2.0 * g.M * g.N * g.K,
            (double)sizeof(T) * (g.M * g.K + g.K * g.N + g.M * g.N * (g.beta != 0.0 ? 2 : 1)));
  if (nsplit > 1) {
    splitk_reduce_kernel<T, CT><<<itg_blocks(g.M * g.N, 256, 4096), 256, 0, ctx->stream>>>(g, nsplit);
    ITG_LAUNCH_CHECK(ctx);
  }
  return ITG_OK;
}

static int launch_gett(itg_ctx* ctx, int dtype, int mode, GettArgs& g, bool ak, bool bk) {
  if (g.M * g.N <= 16) {
    int nblk = (int)std::min<int64_t>(std::max<int64_t>((g.K + 255) / 256, 1), 512);
    double* part = ctx->red + 64;
    if (dtype == ITG_F64) {
      small_out_kernel<double><<<nblk, 256, 0, ctx->stream>>>(g, part);
      small_out_finish<double><<<1, 64, 0, ctx->stream>>>(g, part, nblk);
    } else {
      small_out_kernel<float><<<nblk, 256, 0, ctx->stream>>>(g, part);
      small_out_finish<float><<<1, 64, 0, ctx->stream>>>(g, part, nblk);
    }
    ITG_LAUNCH_CHECK(ctx);
    return ITG_OK;
  }
  if (dtype == ITG_F64) {
    if (mode == ITG_MODE_NATIVE) return launch_gett_t<double, ITG_MODE_NATIVE>(ctx, g, ak, bk);
    if (mode == ITG_MODE_F32) return launch_gett_t<double, ITG_MODE_F32>(ctx, g, ak, bk);
    return launch_gett_t<double, ITG_MODE_BF16X3>(ctx, g, ak, bk);
  }
  if (mode == ITG_MODE_BF16X3) return launch_gett_t<float, ITG_MODE_BF16X3>(ctx, g, ak, bk);
  return launch_gett_t<float, ITG_MODE_NATIVE>(ctx, g, ak, bk);
}

// A contraction whose M, N and K modes each merge into one stride is a plain
// product C(m, n) = alpha sum_k A(m, k) B(k, n) + beta C(m, n), elements at
// m sAm + k sAk, k sBk + n sBn, m sCm + n sCn. It runs as a column-major
// GEMM (the vendor library's) when each matrix has a unit stride on one
// side: X(r, c) at r sr + c sc is op 'N' with ld = sc when sr = 1, op 'T'
// with ld = sr when sc = 1. If C's unit stride is along n, C^T = B^T A^T is
// computed instead. Returns 1 for a layout that is not of this kind.
static int contract_on_lib(itg_ctx* ctx, const GettArgs& g) {
  auto as_col = [](int64_t rows, int64_t cols, int64_t sr, int64_t sc, char& op,
                   int64_t& ld) -> bool {
    if ((rows == 1 || sr == 1) && (cols == 1 || sc >= rows)) {
      op = 'N';
      ld = cols == 1 ? rows : sc;
      return true;
    }
    if ((cols == 1 || sc == 1) && (rows == 1 || sr >= cols)) {
      op = 'T';
      ld = rows == 1 ? cols : sr;
      return true;
    }
    return false;
  };
  char oa, ob, oc;
  int64_t la, lb, lc;
  if (!as_col(g.M, g.N, g.sCm, g.sCn, oc, lc)) return 1;
  if (oc == 'N') {
    if (as_col(g.M, g.K, g.sAm, g.sAk, oa, la) && as_col(g.K, g.N, g.sBk, g.sBn, ob, lb))
      return itg_libgemm(ctx, oa, ob, g.M, g.N, g.K, g.alpha, (const double*)g.A, la,
                         (const double*)g.B, lb, g.beta, (double*)g.C, lc);
  } else {
    // C^T (n x m, column-major with ld lc) = B^T A^T
    if (as_col(g.N, g.K, g.sBn, g.sBk, ob, lb) && as_col(g.K, g.M, g.sAk, g.sAm, oa, la))
      return itg_libgemm(ctx, ob, oa, g.N, g.M, g.K, g.alpha, (const double*)g.B, lb,
                         (const double*)g.A, la, g.beta, (double*)g.C, lc);
  }
  return 1;
}

static int scale_output(itg_ctx* ctx, int dtype, int rC, const int64_t* dC, double beta, void* C) {
  int64_t n = 1;
  for (int i = 0; i < rC; ++i) n *= dC[i];
  if (beta == 0.0) return itg_fill(ctx, dtype, n, 0.0, C);
  return itg_scal(ctx, dtype, n, beta, C);
}

extern "C" int itg_contract(itg_ctx* ctx, int dtype, int mode, int rA, const int64_t* dA,
                            const int32_t* lA, const void* A, int rB, const int64_t* dB,
                            const int32_t* lB, const void* B, int rC, const int64_t* dC,
                            const int32_t* lC, void* C, double alpha, double beta) {
  ITG_CHECK_CTX(ctx);
  if (dtype == ITG_C64)
    return itg_c64_contract(ctx, mode, rA, dA, lA, A, rB, dB, lB, B, rC, dC, lC, C, alpha, beta);
  ITG_CHECK_DTYPE(ctx, dtype);
  if (mode < 0 || mode > 2) return itg_fail(ctx, ITG_EARG, "unknown mode %d", mode);
  if (rA < 0 || rB < 0 || rC < 0 || rA > ITG_MAX_RANK || rB > ITG_MAX_RANK || rC > ITG_MAX_RANK)
    return itg_fail(ctx, ITG_EDIM, "rank out of range");
  // validate labels
  auto find = [](const int32_t* l, int r, int32_t x) {
    for (int i = 0; i < r; ++i) if (l[i] == x) return i;
    return -1;
  };
  for (int i = 0; i < rA; ++i) {
    if (find(lA, i, lA[i]) >= 0) return itg_fail(ctx, ITG_EDIM, "repeated label in A");
    if (dA[i] < 0) return itg_fail(ctx, ITG_EDIM, "negative dimension");
    int jb = find(lB, rB, lA[i]), jc = find(lC, rC, lA[i]);
    if (jb >= 0 && jc >= 0)
      return itg_fail(ctx, ITG_EUNSUPPORTED, "label %d in A, B and C (batch modes)", lA[i]);
    if (jb < 0 && jc < 0) return itg_fail(ctx, ITG_EDIM, "label %d of A appears nowhere else", lA[i]);
    if (jb >= 0 && dB[jb] != dA[i]) return itg_fail(ctx, ITG_EDIM, "contracted dims differ for label %d", lA[i]);
    if (jc >= 0 && dC[jc] != dA[i]) return itg_fail(ctx, ITG_EDIM, "dimension mismatch for label %d", lA[i]);
  }
  for (int i = 0; i < rB; ++i) {
    if (find(lB, i, lB[i]) >= 0) return itg_fail(ctx, ITG_EDIM, "repeated label in B");
    if (dB[i] < 0) return itg_fail(ctx, ITG_EDIM, "negative dimension");
    int ja = find(lA, rA, lB[i]), jc = find(lC, rC, lB[i]);
    if (ja < 0 && jc < 0) return itg_fail(ctx, ITG_EDIM, "label %d of B appears nowhere else", lB[i]);
    if (jc >= 0 && dC[jc] != dB[i]) return itg_fail(ctx, ITG_EDIM, "dimension mismatch for label %d", lB[i]);
  }
  for (int i = 0; i < rC; ++i) {
    if (find(lC, i, lC[i]) >= 0) return itg_fail(ctx, ITG_EDIM, "repeated label in C");
    if (find(lA, rA, lC[i]) < 0 && find(lB, rB, lC[i]) < 0)
      return itg_fail(ctx, ITG_EDIM, "output label %d missing from inputs", lC[i]);
  }
  int64_t nC = 1, nK = 1;
  for (int i = 0; i < rC; ++i) nC *= dC[i];
  for (int i = 0; i < rA; ++i) if (find(lB, rB, lA[i]) >= 0) nK *= dA[i];
  if (nC == 0) return ITG_OK;
  if (nK == 0 || alpha == 0.0) return scale_output(ctx, dtype, rC, dC, beta, C);
  // plan lookup: key = ranks, dims and the label pattern
  std::string key;
  key.reserve(16 + 12 * (rA + rB + rC));
  auto app = [&](const void* p, size_t n) { key.append((const char*)p, n); };
  int32_t hdr[3] = {rA, rB, rC};
  app(hdr, sizeof hdr);
  app(dA, rA * sizeof(int64_t)); app(dB, rB * sizeof(int64_t)); app(dC, rC * sizeof(int64_t));
  for (int i = 0; i < rB; ++i) { int32_t x = find(lA, rA, lB[i]); app(&x, 4); }
  for (int i = 0; i < rC; ++i) {
    int32_t x = find(lA, rA, lC[i]);
    if (x < 0) x = 100 + find(lB, rB, lC[i]);
    app(&x, 4);
  }
  itg_plan* p;
  auto it = ctx->plans.find(key);
  if (it != ctx->plans.end()) {
    p = it->second;
  } else {
    int rc = build_plan(ctx, key, rA, dA, lA, rB, dB, lB, rC, dC, lC, &p);
    if (rc) return rc;
  }
  if (p->sic) return launch_sic(ctx, dtype, p, p->swap ? B : A, p->swap ? A : B, C, alpha, beta);
  GettArgs g;
  memset(&g, 0, sizeof g);
  g.A = p->swap ? B : A;
  g.B = p->swap ? A : B;
  g.C = C;
  g.M = p->M; g.N = p->N; g.K = p->K;
  g.sAm = p->sAm; g.sAk = p->sAk; g.sBk = p->sBk; g.sBn = p->sBn; g.sCm = p->sCm; g.sCn = p->sCn;
  auto tp = [&](int64_t o) { return o >= 0 ? p->dtab + o : (const int64_t*)nullptr; };
  g.tAm = tp(p->oAm); g.tAk = tp(p->oAk); g.tBk = tp(p->oBk);
  g.tBn = tp(p->oBn); g.tCm = tp(p->oCm); g.tCn = tp(p->oCn);
  g.alpha = alpha;
  g.beta = beta;
  if (!g.tAm && !g.tAk && !g.tBk && !g.tBn && !g.tCm && !g.tCn &&
      itg_libgemm_ok(ctx, dtype, mode, g.M, g.N, g.K)) {
    int rc = contract_on_lib(ctx, g);
    if (rc != 1) return rc;
  }
  return launch_gett(ctx, dtype, mode, g, p->akfast, p->bkfast);
}

int itg_gemm_internal(itg_ctx* ctx, int dtype, int mode, char ta, char tb, int64_t m, int64_t n,
                      int64_t k, double alpha, const void* A, int64_t lda, const void* B,
                      int64_t ldb, double beta, void* C, int64_t ldc, bool lower_band) {
  if (m == 0 || n == 0) return ITG_OK;
  if (k == 0 || alpha == 0.0) {
    // C = beta*C on an m x n window with leading dimension ldc
    if (ldc == m) {
      if (beta == 0.0) return itg_fill(ctx, dtype, m * n, 0.0, C);
      return itg_scal(ctx, dtype, m * n, beta, C);
    }
    return itg_fail(ctx, ITG_EUNSUPPORTED, "k == 0 with ldc != m");
  }
  bool tA = (ta == 'T' || ta == 't'), tB = (tb == 'T' || tb == 't');
  if (!lower_band && lda >= (tA ? k : m) && ldb >= (tB ? n : k) && ldc >= m &&
      itg_libgemm_ok(ctx, dtype, mode, m, n, k))
    return itg_libgemm(ctx, ta, tb, m, n, k, alpha, (const double*)A, lda, (const double*)B, ldb,
                       beta, (double*)C, ldc);
  GettArgs g;
  memset(&g, 0, sizeof g);
  // C(m,n) at m + n*ldc: its unit stride is m, so compute C^T = B^T A^T
  g.A = B; g.B = A; g.C = C;
  g.M = n; g.N = m; g.K = k;
  // new A = op(B)^T: element (n, k) of op(B)^T = op(B)(k, n)
  g.sAm = tB ? 1 : ldb;  g.sAk = tB ? ldb : 1;
  // new B = op(A)^T: element (k, m) = op(A)(m, k)
  g.sBn = tA ? lda : 1;  g.sBk = tA ? 1 : lda;
  g.sCm = ldc; g.sCn = 1;
  g.alpha = alpha;
  g.beta = beta;
  // the kernel's (tm, tn) is the caller's (column block, row block): the
  // caller's rows >= columns - 1 tile band is tm <= tn + 1
  g.band = (lower_band && m == n) ? 1 : 0;
  bool ak = (g.sAm != 1) && (g.sAk == 1);
  bool bk = (g.sBn != 1) && (g.sBk == 1);
  return launch_gett(ctx, dtype, mode, g, ak, bk);
}

// C(r, c) = C(c, r) for r < c from the lower triangle, through 32 x 32 LDS
// tiles (block (bi, bj) writes the upper tile at rows bi, columns bj >= bi)
template <typename T>
__global__ __launch_bounds__(256) void mirror_lower_kernel(int64_t n, T* __restrict__ C,
                                                           int64_t ldc) {
  __shared__ T t[32][33];
  const int64_t bi = blockIdx.x, bj = blockIdx.y;
  if (bj < bi) return;
  for (int yy = threadIdx.y; yy < 32; yy += 8) {
    int64_t r = bj * 32 + threadIdx.x, c = bi * 32 + yy;
    if (r < n && c < n) t[yy][threadIdx.x] = C[r + c * ldc];
  }
  __syncthreads();
  for (int yy = threadIdx.y; yy < 32; yy += 8) {
    int64_t r = bi * 32 + threadIdx.x, c = bj * 32 + yy;
    if (r < n && c < n && r < c) C[r + c * ldc] = t[threadIdx.x][yy];
  }
}

extern "C" int itg_gemm(itg_ctx* ctx, int dtype, int mode, char ta, char tb, int64_t m, int64_t n,
                        int64_t k, double alpha, const void* A, int64_t lda, const void* B,
                        int64_t ldb, double beta, void* C, int64_t ldc) {
  ITG_CHECK_CTX(ctx);
  ITG_CHECK_DTYPE(ctx, dtype);
  if (mode < 0 || mode > 2) return itg_fail(ctx, ITG_EARG, "unknown mode %d", mode);
  if (m < 0 || n < 0 || k < 0 || ldc < std::max<int64_t>(1, m))
    return itg_fail(ctx, ITG_EDIM, "bad GEMM dimensions");
  const bool nt = (ta == 'N' || ta == 'n') && (tb == 'T' || tb == 't');
  const bool tn = (ta == 'T' || ta == 't') && (tb == 'N' || tb == 'n');
  if (A == B && lda == ldb && m == n && m >= 256 && k > 0 && beta == 0.0 && alpha != 0.0 &&
      (nt || tn)) {
    // C = alpha A A^T (or A^T A), the density matrix of a bond: the lower
    // triangle (and one tile band) on MFMA, the rest mirrored, so C is
    // exactly symmetric
    int rc;
    if (lda >= (nt ? m : k) && itg_libgemm_ok(ctx, dtype, mode, m, n, k))
      rc = itg_libsyrk_lower(ctx, nt ? 'N' : 'T', m, k, alpha, (const double*)A, lda, (double*)C,
                             ldc);
    else
      rc = itg_gemm_internal(ctx, dtype, mode, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C,
                             ldc, true);
    if (rc) return rc;
    dim3 grid((unsigned)((m + 31) / 32), (unsigned)((m + 31) / 32));
    if (dtype == ITG_F64)
      mirror_lower_kernel<double><<<grid, dim3(32, 8), 0, ctx->stream>>>(m, (double*)C, ldc);
    else
      mirror_lower_kernel<float><<<grid, dim3(32, 8), 0, ctx->stream>>>(m, (float*)C, ldc);
    ITG_LAUNCH_CHECK(ctx);
    return ITG_OK;
  }
  return itg_gemm_internal(ctx, dtype, mode, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

int itg_gett_launch(itg_ctx* ctx, int dtype, int mode, GettArgs* g, bool ak, bool bk) {
  if (g->M <= 0 || g->N <= 0) return ITG_OK;
  return launch_gett(ctx, dtype, mode, *g, ak, bk);
}
