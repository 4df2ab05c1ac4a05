// 3-layer GELU MLP (in -> h -> h -> out) for the learned Davidson
// preconditioner: forward, MSE backward and Adam, FP32. No counterpart in the
// reference. Activations are feature-major (column-major features x rows, i.e.
// row-major rows x features); weights are packed row-major
// [W1(h x in), b1, W2(h x h), b2, W3(out x h), b3]. The matrix products run
// on FP32 MFMA through the contraction kernel; bias, GELU and the loss are
// fused elementwise kernels.
#include <hip/hip_runtime.h>

#include <cmath>

#include "itg_internal.h"

namespace {

__device__ __forceinline__ float gelu_f(float x) {
  return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_grad_f(float x) {
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) +
         x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// Z (f x rows) += b[f]; A = gelu(Z) (A may be null: bias only)
// (A may alias Z: each element is read once, then written)
__global__ void bias_act_kernel(int f, int64_t rows, float* Z, const float* __restrict__ b,
                                float* A) {
  int64_t n = (int64_t)f * rows;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    float z = Z[e] + b[e % f];
    Z[e] = z;
    if (A) A[e] = gelu_f(z);
  }
}

// dY = 2 (Y - T) / (rows * out); per-block partial sums of (Y - T)^2
__global__ __launch_bounds__(256) void mse_grad_kernel(int64_t n, const float* __restrict__ Y,
                                                       const float* __restrict__ T, float scale,
                                                       float* __restrict__ dY, double* part) {
  __shared__ double sh[4];
  double s = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    float dlt = Y[e] - T[e];
    dY[e] = 2.0f * dlt * scale;
    s += (double)dlt * dlt;
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

__global__ void sum_partials_kernel(int n, const double* part, double* out) {
  __shared__ double sh[4];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += part[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = sh[0] + sh[1] + sh[2] + sh[3];
}

// dZ = dA * gelu'(Z), in place on dA
__global__ void gelu_back_kernel(int64_t n, const float* __restrict__ Z, float* __restrict__ dA) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    dA[e] *= gelu_grad_f(Z[e]);
}

// gb[f] = sum over rows of dZ(f, row): one block per feature
__global__ __launch_bounds__(256) void bias_grad_kernel(int f, int64_t rows, const float* __restrict__ dZ,
                                                        float* __restrict__ gb) {
  __shared__ double sh[4];
  int feat = blockIdx.x;
  double s = 0.0;
  for (int64_t r = threadIdx.x; r < rows; r += blockDim.x) s += dZ[feat + (int64_t)f * r];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) gb[feat] = (float)(sh[0] + sh[1] + sh[2] + sh[3]);
}

__global__ void adam_kernel(int64_t n, float* __restrict__ p, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    float gi = g[i];
    float mi = b1 * m[i] + (1.0f - b1) * gi;
    float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

struct Layout {
  int64_t W1, b1, W2, b2, W3, b3, total;
};

Layout layout(int in, int h, int out) {
  Layout L;
  L.W1 = 0;
  L.b1 = L.W1 + (int64_t)h * in;
  L.W2 = L.b1 + h;
  L.b2 = L.W2 + (int64_t)h * h;
  L.W3 = L.b2 + h;
  L.b3 = L.W3 + (int64_t)out * h;
  L.total = L.b3 + out;
  return L;
}

int check_shape(itg_ctx* ctx, int rows, int in, int h, int out) {
  if (rows <= 0 || in <= 0 || h <= 0 || out <= 0)
    return itg_fail(ctx, ITG_EDIM, "mlp: non-positive dimension");
  return ITG_OK;
}

}  // namespace

extern "C" int64_t itg_mlp_param_count(int in, int hidden, int out) {
  return layout(in, hidden, out).total;
}

extern "C" int itg_mlp_forward(itg_ctx* ctx, int rows, int in, int h, int out, const float* P,
                               const float* X, float* Y, float* acts) {
  ITG_CHECK_CTX(ctx);
  int rc = check_shape(ctx, rows, in, h, out);
  if (rc) return rc;
  Layout L = layout(in, h, out);
  int64_t rh = (int64_t)rows * h;
  float* Z1 = acts;
  float* A1 = acts + rh;
  float* Z2 = acts + 2 * rh;
  float* A2 = acts + 3 * rh;
  rc = itg_gemm_internal(ctx, ITG_F32, ITG_MODE_NATIVE, 'T', 'N', h, rows, in, 1.0, P + L.W1, in,
                         X, in, 0.0, Z1, h);
  if (rc) return rc;
  bias_act_kernel<<<itg_blocks(rh, 256, 8192), 256, 0, ctx->stream>>>(h, rows, Z1, P + L.b1, A1);
  rc = itg_gemm_internal(ctx, ITG_F32, ITG_MODE_NATIVE, 'T', 'N', h, rows, h, 1.0, P + L.W2, h, A1,
                         h, 0.0, Z2, h);
  if (rc) return rc;
  bias_act_kernel<<<itg_blocks(rh, 256, 8192), 256, 0, ctx->stream>>>(h, rows, Z2, P + L.b2, A2);
  rc = itg_gemm_internal(ctx, ITG_F32, ITG_MODE_NATIVE, 'T', 'N', out, rows, h, 1.0, P + L.W3, h,
                         A2, h, 0.0, Y, out);
  if (rc) return rc;
  bias_act_kernel<<<itg_blocks((int64_t)rows * out, 256, 8192), 256, 0, ctx->stream>>>(
      out, rows, Y, P + L.b3, nullptr);
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}

// Fused inference for h = 256 and (in, out) within (16, 4) or (32, 9) (the
// preconditioner's d = 2 and d = 3 networks): a persistent grid of 256-thread
// blocks, each walking tiles of 64 rows with every activation on chip. Each
// wave keeps its 64 columns of W2 as bf16 MFMA fragments in registers for the
// whole launch, and b2 and W3 sit in LDS (re-reading and converting W2 from
// L2 in every tile's K loop took 7.0 against 4.65 ms for 4.2M rows, with the
// same results bit for bit). Layer 1 on VALU (thread = hidden unit) into LDS
// as bf16; layer 2 on v_mfma_f32_16x16x32_bf16 (wave w: hidden outputs
// [64 w, 64 w + 64)); bias and GELU in registers; layer 3 as per-lane
// partial dots reduced over the 16 lanes of a column group and then over the
// four waves in LDS.
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
typedef float f4v_t __attribute__((ext_vector_type(4)));
constexpr int MI_H = 256, MI_RB = 64, MI_LD = MI_H + 8;

template <int MI_IN, int MI_OUT>
__global__ __launch_bounds__(256) void mlp_infer_kernel(int64_t rows, int in, int out,
                                                        const float* __restrict__ W1,
                                                        const float* __restrict__ b1,
                                                        const float* __restrict__ W2,
                                                        const float* __restrict__ b2,
                                                        const float* __restrict__ W3,
                                                        const float* __restrict__ b3,
                                                        const float* __restrict__ X,
                                                        float* __restrict__ Y) {
  __shared__ float xs[MI_RB][MI_IN + 1];
  __shared__ __attribute__((aligned(16))) __bf16 h1[MI_RB][MI_LD];
  __shared__ float yp[4][MI_RB][MI_OUT + 1];
  __shared__ float b2s[MI_H];
  __shared__ float w3s[MI_H][MI_OUT];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  // per block, once: layer 1's weights of hidden unit t, layer 3's and b2 in
  // LDS, and this wave's W2 fragments (bw[kk][jj] holds W2 rows
  // 64 wv + 16 jj + lr, columns 32 kk + 8 lg .. + 8, rounded to bf16)
  float w[MI_IN];
#pragma unroll
  for (int k = 0; k < MI_IN; ++k) w[k] = k < in ? W1[k + (int64_t)t * in] : 0.0f;
  const float bj1 = b1[t];
  b2s[t] = b2[t];
#pragma unroll
  for (int o = 0; o < MI_OUT; ++o) w3s[t][o] = o < out ? W3[t + (int64_t)o * MI_H] : 0.0f;
  bf8_t bw[MI_H / 32][4];
#pragma unroll
  for (int kk = 0; kk < MI_H / 32; ++kk)
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      const int j = 64 * wv + 16 * jj + lr;
      const float4* src = (const float4*)(W2 + (int64_t)j * MI_H + 32 * kk + 8 * lg);
      const float4 u = src[0], v = src[1];
      bf8_t bb;
      bb[0] = (__bf16)u.x; bb[1] = (__bf16)u.y; bb[2] = (__bf16)u.z; bb[3] = (__bf16)u.w;
      bb[4] = (__bf16)v.x; bb[5] = (__bf16)v.y; bb[6] = (__bf16)v.z; bb[7] = (__bf16)v.w;
      bw[kk][jj] = bb;
    }
  const int64_t ntiles = (rows + MI_RB - 1) / MI_RB;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t r0 = tile * MI_RB;
    // the whole tile up to MI_IN: layer 1 multiplies every column (by a zero
    // weight past `in`), so none may hold stale LDS contents
    for (int e = t; e < MI_RB * MI_IN; e += 256) {
      const int r = e / MI_IN, k = e % MI_IN;
      xs[r][k] = (k < in && r0 + r < rows) ? X[(r0 + r) * in + k] : 0.0f;
    }
    __syncthreads();
    // layer 1: hidden unit j = t over the 64 rows
    for (int r = 0; r < MI_RB; ++r) {
      float z = bj1;
#pragma unroll
      for (int k = 0; k < MI_IN; ++k) z += xs[r][k] * w[k];
      h1[r][t] = (__bf16)gelu_f(z);
    }
    __syncthreads();
    // layer 2: D[row][col] over rows 64, cols 64 wv + [0, 64); issued as
    // W2-fragment x H1-fragment, so acc[i][jj][e] is element
    // (row 16 i + lr, col 64 wv + 16 jj + 4 lg + e)
    f4v_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) acc[i][jj] = f4v_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < MI_H / 32; ++kk) {
      bf8_t a[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = *(const bf8_t*)&h1[16 * i + lr][32 * kk + 8 * lg];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj)
          acc[i][jj] =
              __builtin_amdgcn_mfma_f32_16x16x32_bf16(bw[kk][jj], a[i], acc[i][jj], 0, 0, 0);
    }
    // bias, GELU and this lane's share of layer 3 for its rows 16 i + lr
    float part[4][MI_OUT];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int o = 0; o < MI_OUT; ++o) part[i][o] = 0.0f;
#pragma unroll
    for (int jj = 0; jj < 4; ++jj)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int j = 64 * wv + 16 * jj + 4 * lg + e;
        const float bj = b2s[j];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float a2 = gelu_f(acc[i][jj][e] + bj);
#pragma unroll
          for (int o = 0; o < MI_OUT; ++o) part[i][o] += a2 * w3s[j][o];
        }
      }
    // sum over the four column groups lg (lanes lr + 16 lg) of the wave
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int o = 0; o < MI_OUT; ++o) {
        float v = part[i][o];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        part[i][o] = v;
      }
    if (lg == 0) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int o = 0; o < MI_OUT; ++o) yp[wv][16 * i + lr][o] = part[i][o];
    }
    __syncthreads();
    for (int e = t; e < MI_RB * out; e += 256) {
      const int r = e / out, o = e % out;
      if (r0 + r < rows)
        Y[(r0 + r) * out + o] =
            b3[o] + ((yp[0][r][o] + yp[1][r][o]) + (yp[2][r][o] + yp[3][r][o]));
    }
    __syncthreads();   // the next tile rewrites xs, h1 and yp
  }
}

extern "C" int itg_mlp_infer(itg_ctx* ctx, int rows, int in, int h, int out, const float* P,
                             const float* X, float* Y, float* work) {
  ITG_CHECK_CTX(ctx);
  int rc = check_shape(ctx, rows, in, h, out);
  if (rc) return rc;
  Layout L = layout(in, h, out);
  // a persistent grid: one block per CU (the kernel runs one wave per SIMD)
  const int64_t ntiles = ((int64_t)rows + MI_RB - 1) / MI_RB;
  const unsigned nblk = (unsigned)std::max<int64_t>(
      1, std::min<int64_t>(ntiles, (int64_t)std::max(ctx->num_cu, 1)));
  // the fused kernel reads W2 in 16-byte vectors
  const bool aligned = ((uintptr_t)(P + L.W2) & 15) == 0;
  if (aligned && h == MI_H && in <= 16 && out <= 4) {
    mlp_infer_kernel<16, 4><<<nblk, 256, 0, ctx->stream>>>(
        rows, in, out, P + L.W1, P + L.b1, P + L.W2, P + L.b2, P + L.W3, P + L.b3, X, Y);
    ITG_LAUNCH_CHECK(ctx);
    return ITG_OK;
  }
  if (aligned && h == MI_H && in <= 32 && out <= 9) {
    mlp_infer_kernel<32, 9><<<nblk, 256, 0, ctx->stream>>>(
        rows, in, out, P + L.W1, P + L.b1, P + L.W2, P + L.b2, P + L.W3, P + L.b3, X, Y);
    ITG_LAUNCH_CHECK(ctx);
    return ITG_OK;
  }
  // other shapes: the forward's GEMMs with activations in `work`
  if (!work) return itg_fail(ctx, ITG_EARG, "mlp_infer: this shape needs a work buffer");
  int64_t rh = (int64_t)rows * h;
  float* A1 = work;
  float* A2 = work + rh;
  rc = itg_gemm_internal(ctx, ITG_F32, ITG_MODE_NATIVE, 'T', 'N', h, rows, in, 1.0, P + L.W1, in,
                         X, in, 0.0, A1, h);
  if (rc) return rc;
  bias_act_kernel<<<itg_blocks(rh, 256, 8192), 256, 0, ctx->stream>>>(h, rows, A1, P + L.b1, A1);
  rc = itg_gemm_internal(ctx, ITG_F32, ITG_MODE_NATIVE, 'T', 'N', h, rows, h, 1.0, P + L.W2, h, A1,
                         h, 0.0, A2, h);
  if (rc) return rc;
  bias_act_kernel<<<itg_blocks(rh, 256, 8192), 256, 0, ctx->stream>>>(h, rows, A2, P + L.b2, A2);
  rc = itg_gemm_internal(ctx, ITG_F32, ITG_MODE_NATIVE, 'T', 'N', out, rows, h, 1.0, P + L.W3, h,
                         A2, h, 0.0, Y, out);
  if (rc) return rc;
  bias_act_kernel<<<itg_blocks((int64_t)rows * out, 256, 8192), 256, 0, ctx->stream>>>(
      out, rows, Y, P + L.b3, nullptr);
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}

static int mlp_workspace(itg_ctx* ctx, size_t bytes, void** out) {
  if (bytes > ctx->mlp_ws_bytes) {
    if (ctx->mlp_ws) {
      ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
      ITG_HIP(ctx, hipFree(ctx->mlp_ws));
      ctx->mlp_ws = nullptr;
      ctx->mlp_ws_bytes = 0;
    }
    if (hipMalloc(&ctx->mlp_ws, bytes) != hipSuccess) {
      hipGetLastError();
      return itg_fail(ctx, ITG_EOOM, "mlp workspace of %zu bytes failed", bytes);
    }
    ctx->mlp_ws_bytes = bytes;
  }
  *out = ctx->mlp_ws;
  return ITG_OK;
}

extern "C" int itg_mlp_backward(itg_ctx* ctx, int rows, int in, int h, int out, const float* P,
                                const float* X, const float* T, const float* acts, const float* Y,
                                float* G, double* loss) {
  ITG_CHECK_CTX(ctx);
  int rc = check_shape(ctx, rows, in, h, out);
  if (rc) return rc;
  Layout L = layout(in, h, out);
  int64_t rh = (int64_t)rows * h, ro = (int64_t)rows * out;
  const float* Z1 = acts;
  const float* A1 = acts + rh;
  const float* Z2 = acts + 2 * rh;
  const float* A2 = acts + 3 * rh;
  void* ws;
  size_t nb = 1024;
  rc = mlp_workspace(ctx, (size_t)(ro + 2 * rh) * sizeof(float) + nb * sizeof(double) + 256, &ws);
  if (rc) return rc;
  float* dY = (float*)ws;
  float* dA2 = dY + ro;
  float* dA1 = dA2 + rh;
  double* part = (double*)(((uintptr_t)(dA1 + rh) + 255) & ~(uintptr_t)255);
  unsigned nblk = itg_blocks(ro, 256, nb);
  mse_grad_kernel<<<nblk, 256, 0, ctx->stream>>>(ro, Y, T, 1.0f / (float)ro, dY, part);
  sum_partials_kernel<<<1, 256, 0, ctx->stream>>>((int)nblk, part, ctx->red);
  ITG_LAUNCH_CHECK(ctx);
  // layer 3
  rc = itg_gemm_internal(ctx, ITG_F32, ITG_MODE_NATIVE, 'N', 'T', h, out, rows, 1.0, A2, h, dY, out,
                         0.0, G + L.W3, h);
  if (rc) return rc;
  bias_grad_kernel<<<out, 256, 0, ctx->stream>>>(out, rows, dY, G + L.b3);
  rc = itg_gemm_internal(ctx, ITG_F32, ITG_MODE_NATIVE, 'N', 'N', h, rows, out, 1.0, P + L.W3, h,
                         dY, out, 0.0, dA2, h);
  if (rc) return rc;
  gelu_back_kernel<<<itg_blocks(rh, 256, 8192), 256, 0, ctx->stream>>>(rh, Z2, dA2);
  // layer 2
  rc = itg_gemm_internal(ctx, ITG_F32, ITG_MODE_NATIVE, 'N', 'T', h, h, rows, 1.0, A1, h, dA2, h,
                         0.0, G + L.W2, h);
  if (rc) return rc;
  bias_grad_kernel<<<h, 256, 0, ctx->stream>>>(h, rows, dA2, G + L.b2);
  rc = itg_gemm_internal(ctx, ITG_F32, ITG_MODE_NATIVE, 'N', 'N', h, rows, h, 1.0, P + L.W2, h,
                         dA2, h, 0.0, dA1, h);
  if (rc) return rc;
  gelu_back_kernel<<<itg_blocks(rh, 256, 8192), 256, 0, ctx->stream>>>(rh, Z1, dA1);
  // layer 1
  rc = itg_gemm_internal(ctx, ITG_F32, ITG_MODE_NATIVE, 'N', 'T', in, h, rows, 1.0, X, in, dA1, h,
                         0.0, G + L.W1, in);
  if (rc) return rc;
  bias_grad_kernel<<<h, 256, 0, ctx->stream>>>(h, rows, dA1, G + L.b1);
  ITG_LAUNCH_CHECK(ctx);
  if (loss) {
    ITG_HIP(ctx, hipMemcpyAsync(ctx->red_host, ctx->red, sizeof(double), hipMemcpyDeviceToHost,
                                ctx->stream));
    ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *loss = ctx->red_host[0] / (double)ro;
  }
  return ITG_OK;
}

extern "C" int itg_mlp_adam(itg_ctx* ctx, int64_t n, float* p, const float* g, float* m, float* v,
                            int step, float lr, float b1, float b2, float eps) {
  ITG_CHECK_CTX(ctx);
  if (n <= 0) return ITG_OK;
  if (step < 1) return itg_fail(ctx, ITG_EARG, "adam: step must be >= 1");
  float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  adam_kernel<<<itg_blocks(n, 256, 4096), 256, 0, ctx->stream>>>(n, p, g, m, v, lr, b1, b2, eps, c1,
                                                                 c2);
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}

// ------------------------------------------------------ residual features
// The network sees a two-site vector (chiL, d, d, chiR) as chiL*chiR rows, one
// per bond pair (a, b), each holding the d*d physical block. Row k has 3*d*d
// features: r/rms(r), delta/rms(delta) and (r/delta)/rms(r/delta), with
// delta = diag(Heff) - lambda clipped away from zero like the diagonal
// preconditioner. Scales are global, so the map is the same at every chi.
namespace {

__device__ __forceinline__ double clip_den(double den, double fl) {
  if (fabs(den) < fl) den = den < 0.0 ? -fl : fl;
  return den;
}

__global__ __launch_bounds__(256) void feat_stats_kernel(int64_t n, const double* __restrict__ r,
                                                         const double* __restrict__ dg, double lam,
                                                         double fl, double* part) {
  __shared__ double sh[3][4];
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    double re = r[e], de = dg[e] - lam, dc = clip_den(de, fl);
    double p = re / dc;
    s0 += re * re;
    s1 += de * de;
    s2 += p * p;
  }
  for (int o = 32; o > 0; o >>= 1) {
    s0 += __shfl_down(s0, o, 64);
    s1 += __shfl_down(s1, o, 64);
    s2 += __shfl_down(s2, o, 64);
  }
  int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sh[0][w] = s0;
    sh[1][w] = s1;
    sh[2][w] = s2;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    int c = threadIdx.x;
    part[3 * blockIdx.x + c] = sh[c][0] + sh[c][1] + sh[c][2] + sh[c][3];
  }
}

// inverse rms of each statistic (1 where it vanishes)
__global__ void feat_scale_kernel(int nb, int64_t n, const double* part, double* inv) {
  __shared__ double sh[3][4];
  double s[3] = {0.0, 0.0, 0.0};
  for (int i = threadIdx.x; i < nb; i += blockDim.x)
    for (int c = 0; c < 3; ++c) s[c] += part[3 * i + c];
  int w = threadIdx.x >> 6;
  for (int c = 0; c < 3; ++c) {
    for (int o = 32; o > 0; o >>= 1) s[c] += __shfl_down(s[c], o, 64);
    if ((threadIdx.x & 63) == 0) sh[c][w] = s[c];
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    int c = threadIdx.x;
    double t = sh[c][0] + sh[c][1] + sh[c][2] + sh[c][3];
    double rms = sqrt(t / (double)n);
    inv[c] = rms > 0.0 ? 1.0 / rms : 1.0;
  }
}

__global__ void features_kernel(int64_t chiL, int dd, int64_t chiR, const double* __restrict__ r,
                                const double* __restrict__ dg, double lam, double fl,
                                const double* __restrict__ inv, int64_t row0, int64_t step,
                                int64_t nrows, float* __restrict__ X) {
  double i0 = inv[0], i1 = inv[1], i2 = inv[2];
  int64_t total = nrows * dd;
  int nf = 3 * dd;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    int64_t k = t / dd;
    int j = (int)(t - k * dd);
    int64_t row = row0 + k * step;
    int64_t a = row % chiL, b = row / chiL;
    int64_t e = a + chiL * (j + (int64_t)dd * b);
    double re = r[e], de = dg[e] - lam;
    float* x = X + k * nf;
    x[j] = (float)(re * i0);
    x[dd + j] = (float)(de * i1);
    x[2 * dd + j] = (float)(re / clip_den(de, fl) * i2);
  }
}

__global__ void gather_rows_kernel(int64_t chiL, int dd, int64_t chiR, const double* __restrict__ v,
                                   double scale, int64_t row0, int64_t step, int64_t nrows,
                                   float* __restrict__ T) {
  int64_t total = nrows * dd;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    int64_t k = t / dd;
    int j = (int)(t - k * dd);
    int64_t row = row0 + k * step;
    int64_t a = row % chiL, b = row / chiL;
    T[t] = (float)(v[a + chiL * (j + (int64_t)dd * b)] * scale);
  }
}

__global__ void scatter_rows_kernel(int64_t chiL, int dd, int64_t chiR, const float* __restrict__ Y,
                                    int64_t row0, int64_t nrows, double* __restrict__ v) {
  int64_t total = nrows * dd;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    int64_t k = t / dd;
    int j = (int)(t - k * dd);
    int64_t row = row0 + k;
    int64_t a = row % chiL, b = row / chiL;
    v[a + chiL * (j + (int64_t)dd * b)] = (double)Y[t];
  }
}

int check_rows(itg_ctx* ctx, int64_t chiL, int64_t d, int64_t chiR, int64_t row0, int64_t step,
               int64_t nrows) {
  if (chiL <= 0 || d <= 0 || chiR <= 0 || d * d > 1024)
    return itg_fail(ctx, ITG_EDIM, "mlp rows: bad shape (%lld,%lld,%lld)", (long long)chiL,
                    (long long)d, (long long)chiR);
  int64_t rows = chiL * chiR;
  if (nrows < 0 || row0 < 0 || step < 1 || (nrows > 0 && row0 + (nrows - 1) * step >= rows))
    return itg_fail(ctx, ITG_EDIM, "mlp rows: rows %lld + k*%lld (k < %lld) outside [0,%lld)",
                    (long long)row0, (long long)step, (long long)nrows, (long long)rows);
  return ITG_OK;
}

}  // namespace

extern "C" int itg_mlp_features(itg_ctx* ctx, int64_t chiL, int64_t d, int64_t chiR,
                                const double* r, const double* diag, double lambda, double floor,
                                int64_t row0, int64_t row_step, int64_t nrows, float* X) {
  ITG_CHECK_CTX(ctx);
  int rc = check_rows(ctx, chiL, d, chiR, row0, row_step, nrows);
  if (rc) return rc;
  if (nrows == 0) return ITG_OK;
  int64_t n = chiL * d * d * chiR;
  unsigned nb = itg_blocks(n, 256, 1024);
  double* part = ctx->red + 16;
  double* inv = ctx->red + 8;
  feat_stats_kernel<<<nb, 256, 0, ctx->stream>>>(n, r, diag, lambda, floor, part);
  feat_scale_kernel<<<1, 256, 0, ctx->stream>>>((int)nb, n, part, inv);
  features_kernel<<<itg_blocks(nrows * d * d, 256, 8192), 256, 0, ctx->stream>>>(
      chiL, (int)(d * d), chiR, r, diag, lambda, floor, inv, row0, row_step, nrows, X);
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}

extern "C" int itg_mlp_gather(itg_ctx* ctx, int64_t chiL, int64_t d, int64_t chiR, const double* v,
                              double scale, int64_t row0, int64_t row_step, int64_t nrows,
                              float* T) {
  ITG_CHECK_CTX(ctx);
  int rc = check_rows(ctx, chiL, d, chiR, row0, row_step, nrows);
  if (rc) return rc;
  if (nrows == 0) return ITG_OK;
  gather_rows_kernel<<<itg_blocks(nrows * d * d, 256, 8192), 256, 0, ctx->stream>>>(
      chiL, (int)(d * d), chiR, v, scale, row0, row_step, nrows, T);
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}

extern "C" int itg_mlp_scatter(itg_ctx* ctx, int64_t chiL, int64_t d, int64_t chiR, const float* Y,
                               int64_t row0, int64_t nrows, double* v) {
  ITG_CHECK_CTX(ctx);
  int rc = check_rows(ctx, chiL, d, chiR, row0, 1, nrows);
  if (rc) return rc;
  if (nrows == 0) return ITG_OK;
  scatter_rows_kernel<<<itg_blocks(nrows * d * d, 256, 8192), 256, 0, ctx->stream>>>(
      chiL, (int)(d * d), chiR, Y, row0, nrows, v);
  ITG_LAUNCH_CHECK(ctx);
  return ITG_OK;
}
