// Diag storage on the device (src/tensor/cudiag.jl) and the block
// embedding of MPS/MPO addition.
//
// A Diag tensor of rank r stores only its diagonal: D(i1, .., ir) = d[k]
// when i1 = .. = ir = k < n, else 0; a uniform Diag is one value on that
// diagonal (NDTensors.Diag(2.0), cudiag.jl's UniformDiag). Its contraction
// with a dense tensor is an element-wise pass: every output element picks
// the diagonal position k from its own coordinates (when some label of D
// survives) or sums over k (when every label of D is contracted).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "itg_internal.h"

namespace {

struct DiagDesc {
  int rC;
  int64_t dC[ITG_MAX_RANK];
  int nDC;                    // D's labels that survive into C
  int posDC[ITG_MAX_RANK];    // their positions in C
  int64_t sBC[ITG_MAX_RANK];  // B stride of each C mode (0: not a B mode)
  int64_t sBk;                // B stride step per diagonal position k
  int64_t nd;                 // diagonal length
};

template <typename T>
struct Ops {
  static __device__ __forceinline__ T mul(T a, T b) { return a * b; }
  static __device__ __forceinline__ T zero() { return T(0); }
  static __device__ __forceinline__ T add(T a, T b) { return a + b; }
  static __device__ __forceinline__ T scale(double s, T a) { return (T)(s * (double)a); }
};
template <>
struct Ops<double2> {
  static __device__ __forceinline__ double2 mul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
  }
  static __device__ __forceinline__ double2 zero() { return make_double2(0.0, 0.0); }
  static __device__ __forceinline__ double2 add(double2 a, double2 b) {
    return make_double2(a.x + b.x, a.y + b.y);
  }
  static __device__ __forceinline__ double2 scale(double s, double2 a) {
    return make_double2(s * a.x, s * a.y);
  }
};

// C = alpha * (D . B) + beta * C; d == nullptr: uniform diagonal value `u`
template <typename T>
__global__ __launch_bounds__(256) void diag_contract_kernel(int64_t total, DiagDesc g,
                                                            const T* __restrict__ d, T u,
                                                            const T* __restrict__ B, double alpha,
                                                            double beta, T* __restrict__ C) {
  using O = Ops<T>;
  for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < total;
       o += (int64_t)gridDim.x * blockDim.x) {
    int64_t rem = o, offB = 0;
    int64_t coord[ITG_MAX_RANK];
    for (int q = 0; q < g.rC; ++q) {
      coord[q] = rem % g.dC[q];
      rem /= g.dC[q];
      offB += coord[q] * g.sBC[q];
    }
    T v = O::zero();
    if (g.nDC > 0) {
      const int64_t k = coord[g.posDC[0]];
      bool on = k < g.nd;
      for (int q = 1; q < g.nDC; ++q) on = on && coord[g.posDC[q]] == k;
      if (on) v = O::mul(d ? d[k] : u, B[offB + k * g.sBk]);
    } else {
      for (int64_t k = 0; k < g.nd; ++k) v = O::add(v, O::mul(d ? d[k] : u, B[offB + k * g.sBk]));
    }
    v = O::scale(alpha, v);
    if (beta != 0.0) v = O::add(v, O::scale(beta, C[o]));
    C[o] = v;
  }
}

template <typename T>
__global__ void hadamard_kernel(int64_t n, const T* __restrict__ x, const T* __restrict__ y,
                                T* __restrict__ z) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    z[i] = Ops<T>::mul(x[i], y[i]);
}

struct EmbedDesc {
  int rank;
  int64_t dS[ITG_MAX_RANK];   // source dims
  int64_t sD[ITG_MAX_RANK];   // destination strides
  int64_t off;                // destination offset of the block's origin
};

// dst[block at offsets] = src, element size ES bytes (moved as words)
template <typename W>
__global__ __launch_bounds__(256) void embed_kernel(int64_t total, EmbedDesc g,
                                                    const W* __restrict__ src, W* __restrict__ dst) {
  for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < total;
       o += (int64_t)gridDim.x * blockDim.x) {
    int64_t rem = o, od = g.off;
    for (int q = 0; q < g.rank; ++q) {
      od += (rem % g.dS[q]) * g.sD[q];
      rem /= g.dS[q];
    }
    dst[od] = src[o];
  }
}

}  // namespace

extern "C" int itg_diag_contract(itg_ctx* ctx, int dtype, int rD, const int64_t* dD,
                                 const int32_t* lD, const void* ddata, int64_t nd,
                                 const double* uniform, int rB, const int64_t* dB,
                                 const int32_t* lB, const void* B, int rC, const int64_t* dC,
                                 const int32_t* lC, void* C, double alpha, double beta) {
  ITG_CHECK_CTX(ctx);
  if (dtype != ITG_F64 && dtype != ITG_F32 && dtype != ITG_C64)
    return itg_fail(ctx, ITG_EDTYPE, "unsupported dtype %d", dtype);
  if (rD < 1 || rB < 0 || rC < 0 || rD > ITG_MAX_RANK || rB > ITG_MAX_RANK || rC > ITG_MAX_RANK)
    return itg_fail(ctx, ITG_EDIM, "diag contract: rank out of range");
  if ((ddata == nullptr) == (uniform == nullptr))
    return itg_fail(ctx, ITG_EARG, "diag contract: give the diagonal or a uniform value");
  auto find = [](const int32_t* l, int r, int32_t x) {
    for (int i = 0; i < r; ++i) if (l[i] == x) return i;
    return -1;
  };
  int64_t mind = dD[0];
  for (int i = 0; i < rD; ++i) {
    if (find(lD, i, lD[i]) >= 0) return itg_fail(ctx, ITG_EDIM, "repeated label in D");
    mind = std::min(mind, dD[i]);
  }
  if (nd < 0 || (ddata && nd != mind))
    return itg_fail(ctx, ITG_EDIM, "diag contract: diagonal length %lld for dims min %lld",
                    (long long)nd, (long long)mind);
  if (!ddata) nd = mind;
  DiagDesc g = {};
  g.rC = rC;
  g.nd = nd;
  // B strides
  int64_t sB[ITG_MAX_RANK];
  int64_t s = 1;
  for (int i = 0; i < rB; ++i) {
    if (find(lB, i, lB[i]) >= 0) return itg_fail(ctx, ITG_EDIM, "repeated label in B");
    sB[i] = s;
    s *= dB[i];
  }
  for (int q = 0; q < rC; ++q) {
    g.dC[q] = dC[q];
    int jd = find(lD, rD, lC[q]), jb = find(lB, rB, lC[q]);
    if ((jd >= 0) == (jb >= 0))
      return itg_fail(ctx, ITG_EDIM, "label %d of C must come from exactly one of D, B", lC[q]);
    if (jd >= 0) {
      if (dD[jd] != dC[q]) return itg_fail(ctx, ITG_EDIM, "dimension mismatch for label %d", lC[q]);
      g.posDC[g.nDC++] = q;
    } else {
      if (dB[jb] != dC[q]) return itg_fail(ctx, ITG_EDIM, "dimension mismatch for label %d", lC[q]);
      g.sBC[q] = sB[jb];
    }
  }
  for (int i = 0; i < rB; ++i) {
    int jd = find(lD, rD, lB[i]), jc = find(lC, rC, lB[i]);
    if (jd >= 0 && jc >= 0)
      return itg_fail(ctx, ITG_EUNSUPPORTED, "label %d in D, B and C (batch modes)", lB[i]);
    if (jd < 0 && jc < 0) return itg_fail(ctx, ITG_EDIM, "label %d of B appears nowhere else", lB[i]);
    if (jd >= 0) {
      if (dD[jd] != dB[i]) return itg_fail(ctx, ITG_EDIM, "contracted dims differ for label %d", lB[i]);
      g.sBk += sB[i];
    }
  }
  for (int i = 0; i < rD; ++i)
    if (find(lB, rB, lD[i]) < 0 && find(lC, rC, lD[i]) < 0)
      return itg_fail(ctx, ITG_EDIM, "label %d of D appears nowhere else", lD[i]);
  int64_t total = 1;
  for (int q = 0; q < rC; ++q) total *= dC[q];
  if (total == 0) return ITG_OK;
  const unsigned nb = itg_blocks(total, 256, 8192);
  if (dtype == ITG_F64) {
    diag_contract_kernel<double><<<nb, 256, 0, ctx->stream>>>(
        total, g, (const double*)ddata, uniform ? uniform[0] : 0.0, (const double*)B, alpha, beta,
        (double*)C);
  } else if (dtype == ITG_F32) {
    diag_contract_kernel<float><<<nb, 256, 0, ctx->stream>>>(
        total, g, (const float*)ddata, uniform ? (float)uniform[0] : 0.0f, (const float*)B, alpha,
        beta, (float*)C);
  } else {
    diag_contract_kernel<double2><<<nb, 256, 0, ctx->stream>>>(
        total, g, (const double2*)ddata,
        uniform ? make_double2(uniform[0], uniform[1]) : make_double2(0.0, 0.0),
        (const double2*)B, alpha, beta, (double2*)C);
  }
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}

extern "C" int itg_hadamard(itg_ctx* ctx, int dtype, int64_t n, const void* x, const void* y,
                            void* z) {
  ITG_CHECK_CTX(ctx);
  if (n <= 0) return ITG_OK;
  const unsigned nb = itg_blocks(n, 256, 8192);
  if (dtype == ITG_F64)
    hadamard_kernel<double><<<nb, 256, 0, ctx->stream>>>(n, (const double*)x, (const double*)y,
                                                         (double*)z);
  else if (dtype == ITG_F32)
    hadamard_kernel<float><<<nb, 256, 0, ctx->stream>>>(n, (const float*)x, (const float*)y,
                                                        (float*)z);
  else if (dtype == ITG_C64)
    hadamard_kernel<double2><<<nb, 256, 0, ctx->stream>>>(n, (const double2*)x,
                                                          (const double2*)y, (double2*)z);
  else
    return itg_fail(ctx, ITG_EDTYPE, "unsupported dtype %d", dtype);
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}

extern "C" int itg_embed(itg_ctx* ctx, int dtype, int rank, const int64_t* dimsS, const void* S,
                         const int64_t* dimsD, const int64_t* offsets, void* D) {
  ITG_CHECK_CTX(ctx);
  if (dtype != ITG_F64 && dtype != ITG_F32 && dtype != ITG_C64)
    return itg_fail(ctx, ITG_EDTYPE, "unsupported dtype %d", dtype);
  if (rank < 0 || rank > ITG_MAX_RANK) return itg_fail(ctx, ITG_EDIM, "embed: rank out of range");
  EmbedDesc g = {};
  g.rank = rank;
  int64_t s = 1, total = 1;
  for (int q = 0; q < rank; ++q) {
    if (dimsS[q] < 0 || offsets[q] < 0 || offsets[q] + dimsS[q] > dimsD[q])
      return itg_fail(ctx, ITG_EDIM, "embed: block of mode %d does not fit", q);
    g.dS[q] = dimsS[q];
    g.sD[q] = s;
    g.off += offsets[q] * s;
    s *= dimsD[q];
    total *= dimsS[q];
  }
  if (total == 0) return ITG_OK;
  const unsigned nb = itg_blocks(total, 256, 8192);
  if (dtype == ITG_F32)
    embed_kernel<float><<<nb, 256, 0, ctx->stream>>>(total, g, (const float*)S, (float*)D);
  else if (dtype == ITG_F64)
    embed_kernel<double><<<nb, 256, 0, ctx->stream>>>(total, g, (const double*)S, (double*)D);
  else
    embed_kernel<double2><<<nb, 256, 0, ctx->stream>>>(total, g, (const double2*)S, (double2*)D);
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}
