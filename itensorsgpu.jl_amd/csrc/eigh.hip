// Symmetric eigendecomposition with truncation, FP64 throughout: the
// `eigen(Hermitian)` step of factorize (src/tensor/culinearalgebra.jl:74-108:
// syevd, eigenvalues reversed to descending, truncate!, columns reversed and
// sliced) used by replacebond! on the density matrix rho = phi phi^T.
//
// 1. Householder tridiagonalisation, blocked as in LAPACK dsytrd/dlatrd
//    (panels of NB columns): per column two kernels spanning the GPU (the
//    reflector fused into the symmetric matrix-vector product over the
//    trailing block, then w and the next column's update; see below); per
//    panel the rank-2*NB trailing update A -= V W^T + W V^T on FP64 MFMA.
// 2. Divide and conquer on the tridiagonal (Cuppen tearing at every split,
//    bottom-up merges of blocks 1,2,4,...): per merge a sort + deflation
//    workgroup, rotation, secular-root, Gu-Eisenstat and eigenvector kernels,
//    and Q_new = Q_old[:, nondeflated] V on MFMA (gathered columns through
//    a device offset table) for large merges.
// 3. Eigenvalues sorted descending, truncate! decided on the device, and only
//    the k kept eigenvectors back-transformed by the compact-WY block
//    reflectors (Schreiber & Van Loan 1989) on MFMA.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <vector>

#include "dc_core.h"
#include "eigh2.h"
#include "gett.h"
#include "itg_internal.h"
#include "truncate_core.h"

int itg_truncate_device_async(itg_ctx* ctx, const double* P, int64_t n, int64_t maxdim,
                              int64_t mindim, double cutoff, int abs_cut, int rel_cut,
                              int semantics, double* out3);

namespace {

constexpr int NB = 32;          // tridiagonalisation panel width
constexpr int TWO_STAGE_MIN = 256;   // n from which the two-stage reduction runs
constexpr int SMALL_MERGE = 256;

__device__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  int nw = (blockDim.x + 63) >> 6;
  for (int i = 0; i < nw; ++i) r += sh[i];
  return r;
}

// ------------------------------------------------------- tridiagonalisation
// Per column j (panel column i) two kernels run across the whole GPU:
//  trd_symv: one block per 64 x 64 tile on or below the diagonal of A22
//    leaves row and column partials of A22 a (a = column j below the
//    diagonal, one writer per slot) and a partial of a^T A22 a; these need
//    neither the reflector's norm nor its scale. Extra blocks form the
//    reflector from the norm partials left by the previous kernel and the
//    panel dots a = W^T v and b = V^T v; the grid stores v into the panel.
//  trd_post: y = A22 v = s (A22 a - beta A22(:, 0)), then
//    w = tau (y - V a - W b) + alpha v, where
//    alpha = -tau/2 w^T v = -tau^2/2 (y^T v - 2 a^T b) needs no extra
//    reduction pass; v goes into A; then the next panel column's update
//    A(:,j+1) -= V W(j+1,:)^T + W V(j+1,:)^T and the partial sums of its
//    squared tail for the next reflector.
constexpr int TRD_ROWS = 64;   // rows per block of the row-parallel kernels

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// sum over a 256-thread block, the same value returned to every thread and
// computed in the same order in every block (bitwise identical results)
__device__ __forceinline__ double block_sum256(double v, double* sh4) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh4[threadIdx.x >> 6] = v;
  __syncthreads();
  return (sh4[0] + sh4[1]) + (sh4[2] + sh4[3]);
}

__device__ __forceinline__ double sum_partials(const double* __restrict__ p, int np, double* sh4) {
  double s = 0.0;
  for (int b = threadIdx.x; b < np; b += 256) s += p[b];
  return block_sum256(s, sh4);
}

// partial sums of A[r, j]^2, r >= j+2, one per TRD_ROWS-row block
__global__ __launch_bounds__(256) void trd_colnorm_kernel(const double* __restrict__ A, int n, int j,
                                                          double* __restrict__ pnorm) {
  __shared__ double sh4[4];
  int r = blockIdx.x * TRD_ROWS + threadIdx.x;
  double s = 0.0;
  if (threadIdx.x < TRD_ROWS && r < n && r >= j + 2) {
    double a = A[r + (size_t)j * n];
    s = a * a;
  }
  s = block_sum256(s, sh4);
  if (threadIdx.x == 0) pnorm[blockIdx.x] = s;
}

// Symmetric product of the trailing matrix A22 (rows/cols o = j+1 .. n-1,
// local index l = r - o, size m) with column j itself, a(l) = A[o + l, j],
// reading only the 64 x 64 tiles on and below the diagonal (both triangles
// of a diagonal tile are stored, so it is used whole). The reflector is
// v = s (a - beta e0), s = 1 / (alpha - beta), so
//   A22 v = s (A22 a - beta A22(:, 0)),
//   v^T A22 v = s^2 (a^T A22 a - 2 beta (A22 a)(0) + beta^2 A22(0, 0)),
// and the tiles need neither the norm nor beta: the post kernel applies s
// and beta. Tile (I, J), I >= J, writes its row partial
// sum_{c in J} A[r,c] a[c] to Ypart[J][r] and, below the diagonal, its
// column partial sum_{r in I} A[r,c] a[r] to Ypart[I][c]: every (t, l) of
// Ypart (t < T tile rows) has exactly one writer and (A22 a)(l) =
// sum_t Ypart[t][l]. Its share of a^T A22 a goes to pyv[tile].
//
// Thread (wave w, lane l) keeps A[r0 + l, c0 + w + 4q], q < 16, in
// registers. Row partials are 16 products per thread summed over the four
// waves in LDS; column partials are sums over the 64 lanes, done for all 16
// columns of a wave at once by a transposing butterfly (each exchange halves
// the values a lane carries), after which lanes with l % 4 == 0 hold column
// w + 4 q(l), q(l) = 8 b5 + 4 b4 + 2 b3 + b2 from the lane's bits. With no
// tile in LDS eight blocks fit on a CU, so a 4096-row trailing block (2080
// tiles) is read in one round.
//
// The first i blocks form the panel dots W(:,q).v and V(:,q).v, q < i; they,
// and the blocks that write v into the panel (and its mirror V2), form the
// reflector from the norm partials (identically in every block). Block 0
// stores d, e, tau and sc = {s, A22(0, 0)}.
constexpr int TS = 64;

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void trd_symv_kernel(
    const double* __restrict__ A, int n, int j, int i, const double* __restrict__ pnorm, int npn,
    double* __restrict__ Vp, const double* __restrict__ Wp, double* __restrict__ V2,
    double* __restrict__ Ypart, double* __restrict__ ab, double* __restrict__ pyv, int ntiles,
    double* d, double* e, double* tau, double* __restrict__ sc) {
  __shared__ double vr[TS], vc[TS];
  __shared__ double red[4][TS];
  __shared__ double sh4[4];
  const size_t ld = (size_t)n;
  const double* a = A + (size_t)j * ld;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int m = n - j - 1, o = j + 1;
  // the panel-dot blocks come first: each walks a whole column, so they
  // start before the tiles fill the machine
  const int bid = (int)blockIdx.x - i;   // tile index, negative for the dots
  const int rv0 = blockIdx.x * 256 + t;
  const bool writes_v = (int)blockIdx.x * 256 < n;
  if (bid >= 0) {
    int I = (int)((sqrt(8.0 * bid + 1.0) - 1.0) * 0.5);
    while ((I + 1) * (I + 2) / 2 <= bid) ++I;
    while (I * (I + 1) / 2 > bid) --I;
    const int J = bid - I * (I + 1) / 2;
    const int r0 = TS * I, c0 = TS * J;
    const bool rok = r0 + lane < m;
    // wave-uniform column pointers with a per-lane row offset
    const int wu = __builtin_amdgcn_readfirstlane(wv);
    const double* base = A + (size_t)(o + r0) + (size_t)(o + c0 + wu) * ld;
    double x[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int c = c0 + wu + 4 * q;
      x[q] = (rok && c < m) ? base[(size_t)(4 * q) * ld + lane] : 0.0;
    }
    if (t < TS) {
      vr[t] = r0 + t < m ? a[o + r0 + t] : 0.0;
    } else if (t < 2 * TS) {
      vc[t - TS] = c0 + t - TS < m ? a[o + c0 + t - TS] : 0.0;
    }
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < 16; ++q) s += x[q] * vc[wv + 4 * q];
    red[wv][lane] = s;
    if (I > J) {
      const double vl = vr[lane];
#pragma unroll
      for (int q = 0; q < 16; ++q) x[q] *= vl;
      // 16 values -> 8 -> 4 -> 2 -> 1 across lane bits 5, 4, 3, 2
      {
        const bool hi = (lane & 32) != 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const double send = hi ? x[k] : x[k + 8];
          const double keep = hi ? x[k + 8] : x[k];
          x[k] = keep + __shfl_xor(send, 32, 64);
        }
      }
      {
        const bool hi = (lane & 16) != 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const double send = hi ? x[k] : x[k + 4];
          const double keep = hi ? x[k + 4] : x[k];
          x[k] = keep + __shfl_xor(send, 16, 64);
        }
      }
      {
        const bool hi = (lane & 8) != 0;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          const double send = hi ? x[k] : x[k + 2];
          const double keep = hi ? x[k + 2] : x[k];
          x[k] = keep + __shfl_xor(send, 8, 64);
        }
      }
      {
        const bool hi = (lane & 4) != 0;
        const double send = hi ? x[0] : x[1];
        const double keep = hi ? x[1] : x[0];
        x[0] = keep + __shfl_xor(send, 4, 64);
      }
      double z = x[0];
      z += __shfl_xor(z, 2, 64);
      z += __shfl_xor(z, 1, 64);
      const int q = ((lane >> 5) & 1) * 8 + ((lane >> 4) & 1) * 4 + ((lane >> 3) & 1) * 2 +
                    ((lane >> 2) & 1);
      const int c = c0 + wv + 4 * q;
      if ((lane & 3) == 0 && c < m) Ypart[(size_t)I * m + c] = z;
    }
    __syncthreads();
    if (wv == 0) {
      const double rowp = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
      if (r0 + lane < m) Ypart[(size_t)J * m + r0 + lane] = rowp;
      double pv = wave_sum(rowp * vr[lane]);
      if (lane == 0) pyv[bid] = (I > J ? 2.0 : 1.0) * pv;
    }
    if (!writes_v) return;
  }
  // the reflector: panel dots, v into the panel, d, e, tau, sc
  const double alpha = a[j + 1];
  const double av0 = (rv0 < n && rv0 > j + 1) ? a[rv0] : 0.0;
  const double ajj = (blockIdx.x == 0 && t == 0) ? a[j] : 0.0;
  const double a00 = (blockIdx.x == 0 && t == 0) ? A[(size_t)o + (size_t)o * ld] : 0.0;
  double xn2 = sum_partials(pnorm, npn, sh4);
  double beta = alpha, tv = 0.0, scale = 1.0;
  if (xn2 != 0.0) {
    beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
    tv = (beta - alpha) / beta;
    scale = 1.0 / (alpha - beta);
  }
  if (blockIdx.x == 0 && t == 0) {
    d[j] = ajj;
    e[j] = beta;
    tau[j] = tv;
    sc[0] = scale;
    sc[1] = a00;
  }
  for (int r = rv0; r < n; r += gridDim.x * 256) {
    double v = r <= j ? 0.0 : (r == j + 1 ? 1.0 : (r == rv0 ? av0 : a[r]) * scale);
    Vp[r + (size_t)i * ld] = v;
    V2[r + (size_t)i * ld] = v;
  }
  if (bid < 0) {
    const int q = (int)blockIdx.x;   // q < i
    const double* wq = Wp + (size_t)q * ld + o;
    const double* vq = Vp + (size_t)q * ld + o;
    const double* ac = a + o;
    double s0 = 0.0, s1 = 0.0, u0 = 0.0, u1 = 0.0;
    int l = t;
    // two rows per thread in flight
    for (; l + 256 < m; l += 2 * 256) {
      const double a0 = ac[l], a1 = ac[l + 256];
      const double w0 = wq[l], w1 = wq[l + 256];
      const double p0 = vq[l], p1 = vq[l + 256];
      const double v0 = l == 0 ? 1.0 : a0 * scale;
      const double v1 = a1 * scale;
      s0 += w0 * v0;
      u0 += w1 * v1;
      s1 += p0 * v0;
      u1 += p1 * v1;
    }
    for (; l < m; l += 256) {
      const double vv = l == 0 ? 1.0 : ac[l] * scale;
      s0 += wq[l] * vv;
      s1 += vq[l] * vv;
    }
    s0 = block_sum256(s0 + u0, sh4);
    s1 = block_sum256(s1 + u1, sh4);
    if (t == 0) {
      ab[q] = s0;
      ab[i + q] = s1;
    }
  }
}

// sum of p[b] over b = t (mod 256) for thread t, eight loads in flight
__device__ __forceinline__ double strided_partial(const double* __restrict__ p, int np, int t) {
  double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0, q4 = 0.0, q5 = 0.0, q6 = 0.0, q7 = 0.0;
  int b = t;
  for (; b + 7 * 256 < np; b += 8 * 256) {
    q0 += p[b];
    q1 += p[b + 256];
    q2 += p[b + 512];
    q3 += p[b + 768];
    q4 += p[b + 1024];
    q5 += p[b + 1280];
    q6 += p[b + 1536];
    q7 += p[b + 1792];
  }
  for (; b < np; b += 256) q0 += p[b];
  return ((q0 + q1) + (q2 + q3)) + ((q4 + q5) + (q6 + q7));
}


constexpr int POST_ROWS = 16;   // rows per block of the post kernel

// Rows [r0, r0 + 16), r0 = 16 (b0 + blockIdx.x), b0 = (j + 1) / 16 (rows up
// to j are finished): w = tau (y - V a - W b) + alpha v, v -> A(:, j), the
// next panel column's update and its squared-tail partial sum. Thread
// (row rr = t % 16, part = t / 16) sums its share t = part (mod 16) of the
// symv row partials and q = part (mod 16) of the panel terms; the 16 parts
// meet in LDS once. The symv partials are of A22 a; with the reflector's
// scale s = sc[0] and beta = e[j], y(r) = s ((A22 a)(r) - beta A(r, j+1)),
// and v^T A22 v = s^2 (a^T A22 a - 2 beta (A22 a)(0) + beta^2 sc[1]),
// sc[1] = A(j+1, j+1). The scalars (v^T A22 v, a^T b, y(j+1), the panel
// dot at j+1) are formed the same way in every block, so all blocks agree
// bit for bit.
__global__ __launch_bounds__(256) void trd_post_kernel(
    double* __restrict__ A, int n, int j, int i, int nsym, const double* __restrict__ Vp,
    double* __restrict__ Wp, const double* __restrict__ Ypart, int T,
    const double* __restrict__ ab, const double* __restrict__ pyv,
    const double* __restrict__ tau, const double* __restrict__ e,
    const double* __restrict__ sc, int next, double* __restrict__ pnorm, int b0) {
  __shared__ double red[3][POST_ROWS][POST_ROWS + 1];
  __shared__ double scr[4][4];
  const size_t ld = (size_t)n;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int rr = t % POST_ROWS, part = t / POST_ROWS;
  const int r = (b0 + (int)blockIdx.x) * POST_ROWS + rr;
  const int jn = j + 1;
  const int m = n - j - 1;
  const bool live = r > j && r < n;   // rows of v, w and the update
  // symv row partials: y(r) = sum_t Ypart[t][r - jn]
  double yp = 0.0;
  if (live) {
    const double* yc = Ypart + (r - jn);
    double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0;
    int tt = part;
    for (; tt + 3 * POST_ROWS < T; tt += 4 * POST_ROWS) {
      q0 += yc[(size_t)tt * m];
      q1 += yc[(size_t)(tt + POST_ROWS) * m];
      q2 += yc[(size_t)(tt + 2 * POST_ROWS) * m];
      q3 += yc[(size_t)(tt + 3 * POST_ROWS) * m];
    }
    for (; tt < T; tt += POST_ROWS) q0 += yc[(size_t)tt * m];
    yp = (q0 + q1) + (q2 + q3);
  }
  // panel terms over q < i: V(r,q) a_q + W(r,q) b_q and, for the next
  // column, V(r,q) W(jn,q) + W(r,q) V(jn,q)
  double sw = 0.0, sn = 0.0;
  if (live) {
    for (int q = part; q < i; q += POST_ROWS) {
      const double vq = Vp[r + (size_t)q * ld], wq = Wp[r + (size_t)q * ld];
      sw += vq * ab[q] + wq * ab[i + q];
      if (next) sn += vq * Wp[jn + (size_t)q * ld] + wq * Vp[jn + (size_t)q * ld];
    }
  }
  // block scalars: every thread a strided share of the v^T A22 v partials;
  // wave 1: y(jn); wave 2: a^T b; wave 3: the panel dot at row jn
  double pv = strided_partial(pyv, nsym, t);
  double s1 = 0.0;
  if (wv == 1) {
    for (int tt = lane; tt < T; tt += 64) s1 += Ypart[(size_t)tt * m];
  } else if (wv == 2) {
    if (lane < i) s1 = ab[lane] * ab[i + lane];
  } else if (wv == 3) {
    if (next && lane < i)
      s1 = Vp[jn + (size_t)lane * ld] * ab[lane] + Wp[jn + (size_t)lane * ld] * ab[i + lane];
  }
  // wave 0's operands of the final step, loaded before the barrier
  const double tj = tau[j], bj = e[j], sj = sc[0], a00 = sc[1];
  double vr0 = 0.0, ajn0 = 0.0;
  if (wv == 0 && lane < POST_ROWS && r < n) {
    vr0 = Vp[r + (size_t)i * ld];
    if (r >= jn) ajn0 = A[r + (size_t)jn * ld];
  }
  pv = wave_sum(pv);
  s1 = wave_sum(s1);
  red[0][part][rr] = yp;
  red[1][part][rr] = sw;
  red[2][part][rr] = sn;
  if (lane == 0) {
    scr[0][wv] = pv;
    scr[1][wv] = s1;
  }
  __syncthreads();
  if (wv != 0) return;
  const double ata = (scr[0][0] + scr[0][1]) + (scr[0][2] + scr[0][3]);
  const double y0a = scr[1][1], atb = scr[1][2], djn = scr[1][3];
  const double ytv = sj * sj * ((ata - 2.0 * bj * y0a) + bj * bj * a00);
  const double y0 = sj * (y0a - bj * a00);
  const double alpha = -0.5 * tj * tj * (ytv - 2.0 * atb);
  double s = 0.0;
  if (lane < POST_ROWS && r < n) {
    double ya = 0.0, swr = 0.0, snr = 0.0;
#pragma unroll
    for (int p = 0; p < POST_ROWS; ++p) {
      ya += red[0][p][lane];
      swr += red[1][p][lane];
      snr += red[2][p][lane];
    }
    const double yr = sj * (ya - bj * ajn0);
    const double vr = vr0;
    double wr = 0.0;
    if (r > j && tj != 0.0) wr = tj * (yr - swr) + alpha * vr;
    Wp[r + (size_t)i * ld] = wr;
    if (r > j) A[r + (size_t)j * ld] = vr;
    if (next && r >= jn) {
      // w(jn) for the panel column i (V(jn, i) = 1)
      const double wj1 = (tj != 0.0) ? tj * (y0 - djn) + alpha : 0.0;
      double acc = ajn0 - snr;
      acc -= vr * wj1 + wr;
      A[r + (size_t)jn * ld] = acc;
      if (r >= jn + 2) s = acc * acc;
    }
  }
  if (next) {
    s = wave_sum(s);
    if (lane == 0) pnorm[b0 + blockIdx.x] = s;
  }
}

constexpr int TRD_TAIL = 128;   // trailing blocks this small are reduced in LDS

__device__ __forceinline__ double block_sum1024(double v, double* red) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int q = 0; q < 16; ++q) s += red[q];
  return s;
}

// The trailing block of m = n - p <= TRD_TAIL rows and columns (every
// earlier update applied, its lower triangle current) reduced to tridiagonal
// form in LDS by one 1024-thread block, column by column (dsytd2): y = tau
// A22 v, w = y - tau/2 (y^T v) v, A22 -= v w^T + w v^T. d, e, tau and the
// reflectors (v(j+1) = 1 stored) go where the per-column kernels put them.
__global__ __launch_bounds__(1024) void trd_tail_kernel(double* __restrict__ A, int n, int p,
                                                        double* __restrict__ d,
                                                        double* __restrict__ e,
                                                        double* __restrict__ tau) {
  extern __shared__ double S[];
  __shared__ double yp[8][TRD_TAIL];
  __shared__ double red[16];
  const int m = n - p, ld = m + 1;
  double* a = S;                    // a[r + c ld], both triangles
  double* v = S + (size_t)m * ld;   // [m]
  double* w = v + m;                // [m]
  const int t = threadIdx.x;
  for (int q = t; q < m * m; q += 1024) {
    const int r = q % m, c = q / m;
    const int hi = r > c ? r : c, lo = r > c ? c : r;
    a[r + c * ld] = A[(size_t)(p + hi) + (size_t)(p + lo) * n];
  }
  __syncthreads();
  for (int k = 0; k + 1 < m; ++k) {
    double s = 0.0;
    for (int r = k + 2 + t; r < m; r += 1024) {
      const double x = a[r + k * ld];
      s += x * x;
    }
    const double xn2 = block_sum1024(s, red);
    const double alpha = a[k + 1 + k * ld];
    double beta = alpha, tv = 0.0, scale = 1.0;
    if (xn2 != 0.0) {
      beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
      tv = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
    if (t == 0) {
      d[p + k] = a[k + k * ld];
      e[p + k] = beta;
      tau[p + k] = tv;
    }
    for (int r = t; r < m; r += 1024) {
      const double vr = r <= k ? 0.0 : (r == k + 1 ? 1.0 : a[r + k * ld] * scale);
      v[r] = vr;
      if (r > k) A[(size_t)(p + r) + (size_t)(p + k) * n] = vr;
    }
    __syncthreads();
    {
      // A22 v: row r = t % TRD_TAIL, columns c = k + 1 + part (mod 8)
      const int r = t % TRD_TAIL, part = t / TRD_TAIL;
      double acc = 0.0;
      if (r < m && r > k)
        for (int c = k + 1 + part; c < m; c += 8) acc += a[r + c * ld] * v[c];
      yp[part][r] = acc;
    }
    __syncthreads();
    double yv = 0.0;
    if (t < m) {
      double y = 0.0;
      if (t > k) {
        for (int q = 0; q < 8; ++q) y += yp[q][t];
        y *= tv;
        yv = y * v[t];
      }
      w[t] = y;
    }
    const double ytv = block_sum1024(yv, red);
    if (t < m && t > k) w[t] += -0.5 * tv * ytv * v[t];
    __syncthreads();
    for (int q = t; q < m * m; q += 1024) {
      const int r = q % m, c = q / m;
      if (r > k && c > k) a[r + c * ld] -= v[r] * w[c] + w[r] * v[c];
    }
    __syncthreads();
  }
  if (t == 0) {
    d[n - 1] = a[(m - 1) + (m - 1) * ld];
    tau[n - 1] = 0.0;
  }
}

__global__ void trd_last_kernel(const double* __restrict__ A, int n, double* d, double* tau) {
  d[n - 1] = A[(size_t)(n - 1) * n + (n - 1)];
  tau[n - 1] = 0.0;
}

// --------------------------------------------------------- back-transform
// reflector j has v(j+off) = 1 and v(r) = A[r, j] below it (off = 1 for the
// tridiagonal reduction, 0 for QR)
__global__ void bt_build_v_kernel(const double* __restrict__ A, int n, int p, int nbp, int off,
                                  double* __restrict__ Vb) {
  int64_t tot = (int64_t)n * nbp;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    int r = (int)(e % n), i = (int)(e / n);
    int j = p + i;
    double v = 0.0;
    if (r == j + off) v = 1.0;
    else if (r > j + off) v = A[r + (size_t)j * n];
    Vb[e] = v;
  }
}

// compact-WY triangular factor, forward columnwise (dlarft 'F','C'); G and
// tau are staged in LDS first so the column recurrence touches no HBM
__global__ void bt_larft_kernel(int nbp, const double* __restrict__ G, const double* __restrict__ tau,
                                double* __restrict__ T) {
  __shared__ double Ts[NB][NB + 1];
  __shared__ double Gs[NB][NB + 1];
  __shared__ double ts[NB];
  __shared__ double tmp[NB];
  int l = threadIdx.x;
  for (int e = l; e < NB * NB; e += blockDim.x) {
    int r = e % NB, c = e / NB;
    Ts[r][c] = 0.0;
    Gs[r][c] = (r < nbp && c < nbp) ? G[r + c * nbp] : 0.0;
  }
  if (l < nbp) ts[l] = tau[l];
  __syncthreads();
  for (int i = 0; i < nbp; ++i) {
    if (l < i) tmp[l] = -ts[i] * Gs[l][i];
    __syncthreads();
    if (l < i) {
      double s = 0.0;
      for (int m = l; m < i; ++m) s += Ts[l][m] * tmp[m];
      Ts[l][i] = s;
    }
    if (l == i) Ts[i][i] = ts[i];
    __syncthreads();
  }
  for (int e = l; e < nbp * nbp; e += blockDim.x) T[e] = Ts[e % nbp][e / nbp];
}

// After the tridiagonalisation column j of A holds v_j below the diagonal
// (v_j(j+1) = 1 stored); clearing rows 0..j makes A[:, p..p+NB) the panel's
// reflector block V with no copy.
__global__ void bt_clear_upper_kernel(double* __restrict__ A, int n) {
  int64_t tot = (int64_t)n * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    int r = (int)(e % n), j = (int)(e / n);
    if (r <= j) A[e] = 0.0;
  }
}

// All compact-WY factors at once: block b forms G = V_b^T V_b of panel b
// (columns p = NB b .. p+nb, rows p+1..n-1, staged through LDS in 64-row
// chunks) and runs the dlarft recurrence on it in LDS; T_b goes to
// Tall[b * NB * NB] (nb x nb, column-major with leading dimension nb).
__global__ __launch_bounds__(256) void bt_tall_kernel(const double* __restrict__ A, int n,
                                                      int nrefl, const double* __restrict__ tau,
                                                      double* __restrict__ Tall) {
  __shared__ double Vs[64][NB + 1];
  __shared__ double Gs[NB][NB + 1];
  __shared__ double Ts[NB][NB + 1];
  __shared__ double ts[NB];
  __shared__ double tmp[NB];
  const int t = threadIdx.x;
  const int p = blockIdx.x * NB;
  const int nb = min(NB, nrefl - p);
  if (nb <= 0) return;
  double g0 = 0.0, g1 = 0.0, g2 = 0.0, g3 = 0.0;   // pairs t, t+256, t+512, t+768
  for (int r0 = p + 1; r0 < n; r0 += 64) {
    __syncthreads();
    for (int x = t; x < 64 * NB; x += 256) {
      int rr = x % 64, c = x / 64, r = r0 + rr;
      Vs[rr][c] = (r < n && c < nb) ? A[r + (size_t)(p + c) * n] : 0.0;
    }
    __syncthreads();
    int rows = min(64, n - r0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int pr = t + 256 * q, ca = pr % NB, cb = pr / NB;
      double s = 0.0;
      for (int rr = 0; rr < rows; ++rr) s += Vs[rr][ca] * Vs[rr][cb];
      if (q == 0) g0 += s; else if (q == 1) g1 += s; else if (q == 2) g2 += s; else g3 += s;
    }
  }
  {
    double gq[4] = {g0, g1, g2, g3};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int pr = t + 256 * q;
      Gs[pr % NB][pr / NB] = gq[q];
      Ts[pr % NB][pr / NB] = 0.0;
    }
  }
  if (t < nb) ts[t] = tau[p + t];
  __syncthreads();
  for (int i = 0; i < nb; ++i) {
    if (t < i) tmp[t] = -ts[i] * Gs[t][i];
    __syncthreads();
    if (t < i) {
      double acc = 0.0;
      for (int m = t; m < i; ++m) acc += Ts[t][m] * tmp[m];
      Ts[t][i] = acc;
    }
    if (t == i) Ts[i][i] = ts[i];
    __syncthreads();
  }
  double* T = Tall + (size_t)blockIdx.x * NB * NB;
  for (int e = t; e < nb * nb; e += 256) T[e] = Ts[e % nb][e / nb];
}

// X2 = T X for one panel: T nb x nb upper triangular, X nb x k (ld nb);
// one thread per element, eight columns per block (rows along lanes)
__global__ __launch_bounds__(256) void bt_apply_t_kernel(int nb, int k, const double* __restrict__ T,
                                                         const double* __restrict__ X,
                                                         double* __restrict__ X2) {
  __shared__ double Tsh[NB * NB];
  __shared__ double Xs[8][NB];
  const int t = threadIdx.x, r = t % NB, cl = t / NB;
  const int c = blockIdx.x * 8 + cl;
  for (int e = t; e < nb * nb; e += 256) Tsh[e] = T[e];
  Xs[cl][r] = (c < k && r < nb) ? X[(size_t)c * nb + r] : 0.0;
  __syncthreads();
  if (c >= k || r >= nb) return;
  double s = 0.0;
  for (int q = r; q < nb; ++q) s += Tsh[r + q * nb] * Xs[cl][q];
  X2[(size_t)c * nb + r] = s;
}

// Back-transform of a group of BTG consecutive panels (up to BTG*NB
// reflectors) at once. With Y = V_G^T U (the group's rows of U before the
// group is applied) and the Gram matrix Gg = V_G^T V_G, the panels applied
// last-to-first give X_b = Y_b - sum_{b' > b} Gg(b, b') X2_b' and
// X2_b = T_b X_b; then U -= V_G X2 in one product. One block: 8 columns of
// U x the 32 rows of the current panel; X2 of the later panels in LDS.
constexpr int BTG = 4;

__global__ __launch_bounds__(256) void bt_group_solve_kernel(int nbg, int k,
                                                             const double* __restrict__ Y,
                                                             const double* __restrict__ Gg,
                                                             const double* __restrict__ Tall,
                                                             double* __restrict__ X2) {
  __shared__ double X2s[8][BTG * NB];
  __shared__ double xs[8][NB];
  __shared__ double Gs[NB][(BTG - 1) * NB + 1];   // Gg rows of panel b, later columns
  __shared__ double Ts[NB][NB + 1];
  const int t = threadIdx.x, r = t % NB, cl = t / NB;
  const int c = blockIdx.x * 8 + cl;
  const bool live = c < k;
  const int npb = (nbg + NB - 1) / NB;
  for (int b = npb - 1; b >= 0; --b) {
    const int r0 = b * NB, nb = min(NB, nbg - r0), q0 = r0 + nb, nq = nbg - q0;
    // stage Gg(r0.., q0..) and T_b (group-relative panel b)
    for (int e = t; e < NB * nq; e += 256) {
      int rr = e % NB, qq = e / NB;
      Gs[rr][qq] = rr < nb ? Gg[(r0 + rr) + (size_t)(q0 + qq) * nbg] : 0.0;
    }
    const double* T = Tall + (size_t)b * NB * NB;
    for (int e = t; e < nb * nb; e += 256) Ts[e % nb][e / nb] = T[e];
    __syncthreads();
    double x = 0.0;
    if (live && r < nb) {
      x = Y[(size_t)c * nbg + r0 + r];
      for (int qq = 0; qq < nq; ++qq) x -= Gs[r][qq] * X2s[cl][q0 + qq];
    }
    xs[cl][r] = x;
    __syncthreads();
    if (live && r < nb) {
      double v = 0.0;
      for (int q = r; q < nb; ++q) v += Ts[r][q] * xs[cl][q];
      X2s[cl][r0 + r] = v;
      X2[(size_t)c * nbg + r0 + r] = v;
    }
    __syncthreads();
  }
}

// --------------------------------------------------------- divide & conquer
// Q = I and Qn = 0: every merge rewrites whole block squares, and entries
// outside the current squares of both ping-pong buffers must stay zero
__global__ void dc_init_kernel(int n, const double* d, const double* e, double* D, double* Q,
                               double* Qn) {
  int64_t tot = (int64_t)n * n;
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < tot;
       x += (int64_t)gridDim.x * blockDim.x) {
    int r = (int)(x % n), c = (int)(x / n);
    Q[x] = (r == c) ? 1.0 : 0.0;
    Qn[x] = 0.0;
    if (r == 0) {
      double v = d[c];
      if (c > 0) v -= fabs(e[c - 1]);
      if (c < n - 1) v -= fabs(e[c]);
      D[c] = v;
    }
  }
}

__device__ void bitonic_sort(double* key, int* val, int np2) {
  for (int k = 2; k <= np2; k <<= 1) {
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
      for (int i = threadIdx.x; i < np2; i += blockDim.x) {
        int ixj = i ^ jj;
        if (ixj > i) {
          bool up = ((i & k) == 0);
          double a = key[i], b = key[ixj];
          if ((a > b) == up) {
            key[i] = b;
            key[ixj] = a;
            int tv = val[i];
            val[i] = val[ixj];
            val[ixj] = tv;
          }
        }
      }
      __syncthreads();
    }
  }
}

struct DcBufs {
  double *ds, *zs, *dl, *wv, *dval, *rcs, *tau, *zh, *mrho;
  int *cs, *gcol, *dcol, *rcol, *org, *meta;
};

constexpr int DC_ZLDS = 4096;   // largest merge whose z is staged in LDS

__global__ __launch_bounds__(1024) void dc_prepare_kernel(int n, int s_lvl, int half,
                                                          const double* __restrict__ D,
                                                          const double* __restrict__ Q,
                                                          const double* __restrict__ e, DcBufs b) {
  extern __shared__ double lds_d[];
  int m = blockIdx.x;
  int r0 = m * s_lvl;
  int s = min(s_lvl, n - r0);
  if (s <= half) return;
  int s1 = half;
  int np2 = 1;
  while (np2 < s) np2 <<= 1;
  double* key = lds_d;
  int* val = (int*)(key + np2);
  for (int i = threadIdx.x; i < np2; i += blockDim.x) {
    key[i] = i < s ? D[r0 + i] : INFINITY;
    val[i] = r0 + (i < s ? i : 0);
  }
  __syncthreads();
  bitonic_sort(key, val, np2);
  double beta = e[r0 + s1 - 1];
  double sg = beta >= 0.0 ? 1.0 : -1.0;
  // the sequential deflation pass reads d, z and the columns from LDS when
  // z fits beside them (s_lvl <= DC_ZLDS), else z from global memory
  double* z = s_lvl <= DC_ZLDS ? (double*)(val + np2) : b.zs + r0;
  for (int i = threadIdx.x; i < s; i += blockDim.x) {
    int c = val[i];
    double zi = (c < r0 + s1) ? Q[(size_t)c * n + (r0 + s1 - 1)] : sg * Q[(size_t)c * n + (r0 + s1)];
    z[i] = zi * 0.7071067811865476;
  }
  __syncthreads();
  __shared__ int ksh, serial, isum[16];
  __shared__ double wsum[16], amax[2][16];
  const double rho = 2.0 * fabs(beta);
  if (s_lvl <= DC_ZLDS) {
    // Without a Givens rotation the sequential deflation pass is a pure
    // compaction: entries with rho |z| <= tol go to the deflated list, the
    // rest to the kept list, both in sorted order. Every pair test of
    // consecutive kept entries is made here with the original values; if
    // none calls for a rotation the two lists are formed in parallel,
    // else the sequential pass below runs as before.
    int* kidx = (int*)(z + np2);
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    double dm = 0.0, zm = 0.0;
    for (int i = t; i < s; i += blockDim.x) {
      dm = fmax(dm, fabs(key[i]));
      zm = fmax(zm, fabs(z[i]));
    }
    for (int o = 32; o > 0; o >>= 1) {
      dm = fmax(dm, __shfl_xor(dm, o, 64));
      zm = fmax(zm, __shfl_xor(zm, o, 64));
    }
    if (lane == 0) {
      amax[0][wv] = dm;
      amax[1][wv] = zm;
    }
    if (t == 0) serial = 0;
    __syncthreads();
    dm = 0.0;
    zm = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) {
      dm = fmax(dm, amax[0][q]);
      zm = fmax(zm, amax[1][q]);
    }
    const double tol = 8.0 * ITG_DC_EPS * fmax(dm, zm);
    // thread t owns the contiguous entries [t cs, t cs + cs)
    const int cs = (s + (int)blockDim.x - 1) / (int)blockDim.x;
    const int i0 = min(s, t * cs), i1 = min(s, i0 + cs);
    int cnt = 0;
    for (int i = i0; i < i1; ++i) cnt += (rho * fabs(z[i]) <= tol) ? 0 : 1;
    // exclusive scan of the kept counts over the block
    int inc = cnt;
    for (int o = 1; o < 64; o <<= 1) {
      int u = __shfl_up(inc, o, 64);
      if (lane >= o) inc += u;
    }
    if (lane == 63) isum[wv] = inc;
    __syncthreads();
    int base = 0, total = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) {
      if (q < wv) base += isum[q];
      total += isum[q];
    }
    int pos = base + inc - cnt;
    for (int i = i0; i < i1; ++i) {
      if (rho * fabs(z[i]) <= tol) {
        b.dval[r0 + (i - pos)] = key[i];
        b.dcol[r0 + (i - pos)] = val[i];
      } else {
        kidx[pos++] = i;
      }
    }
    __syncthreads();
    if (rho * zm <= tol) serial = 1;   // everything deflates: the sequential pass
    for (int p = t + 1; p < total; p += blockDim.x) {
      const int pj = kidx[p - 1], j = kidx[p];
      double S = z[pj], Cc = z[j];
      double tau = sqrt(Cc * Cc + S * S);
      double tt = key[j] - key[pj];
      double it = 1.0 / tau;
      double c = Cc * it, sn = -S * it;
      if (fabs(tt * c * sn) <= tol) serial = 1;
    }
    __syncthreads();
    if (!serial) {
      for (int p = t; p < total; p += blockDim.x) {
        const int j = kidx[p];
        b.dl[r0 + p] = key[j];
        b.wv[r0 + p] = z[j];
        b.gcol[r0 + p] = val[j];
      }
      if (t == 0) {
        ksh = total;
        b.meta[3 * m] = total;
        b.meta[3 * m + 1] = s - total;
        b.meta[3 * m + 2] = 0;
        b.mrho[2 * m] = rho;
      }
    }
  } else if (threadIdx.x == 0) {
    serial = 1;
  }
  __syncthreads();
  if (serial && threadIdx.x == 0) {
    int nd = 0, nrot = 0;
    int k = dc_deflate(s, key, z, val, rho, b.dl + r0, b.wv + r0, b.gcol + r0, b.dval + r0,
                       b.dcol + r0, &nd, b.rcol + 2 * r0, b.rcs + 2 * r0, &nrot);
    ksh = k;
    b.meta[3 * m] = k;
    b.meta[3 * m + 1] = nd;
    b.meta[3 * m + 2] = nrot;
    b.mrho[2 * m] = rho;
  }
  __syncthreads();
  // sum of w^2 over the k kept entries: strided per thread, per wave, then
  // the 16 wave sums in order
  const int k = ksh;
  double w2 = 0.0;
  for (int j = threadIdx.x; j < k; j += blockDim.x) w2 += b.wv[r0 + j] * b.wv[r0 + j];
  for (int o = 32; o > 0; o >>= 1) w2 += __shfl_xor(w2, o, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = w2;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += wsum[q];
    b.mrho[2 * m + 1] = t;
  }
}

__device__ __forceinline__ bool merge_range(int n, int s_lvl, int half, int m, int& r0, int& s) {
  r0 = m * s_lvl;
  s = min(s_lvl, n - r0);
  return s > half;
}

__global__ void dc_rotate_kernel(int n, int s_lvl, int half, double* __restrict__ Q, DcBufs b) {
  int m = blockIdx.x, r0, s;
  if (!merge_range(n, s_lvl, half, m, r0, s)) return;
  int nrot = b.meta[3 * m + 2];
  int r = r0 + blockIdx.y * blockDim.x + threadIdx.x;
  if (r >= r0 + s) return;
  for (int t = 0; t < nrot; ++t) {
    int pc = b.rcol[2 * r0 + 2 * t], qc = b.rcol[2 * r0 + 2 * t + 1];
    double c = b.rcs[2 * r0 + 2 * t], sn = b.rcs[2 * r0 + 2 * t + 1];
    double x = Q[(size_t)pc * n + r], yv = Q[(size_t)qc * n + r];
    Q[(size_t)pc * n + r] = c * x + sn * yv;
    Q[(size_t)qc * n + r] = c * yv - sn * x;
  }
}

// One wave per root: lanes split the sums over the k poles (the xor-butterfly
// sums leave every lane the same bits, so the iteration stays wave-uniform).
__global__ __launch_bounds__(256) void dc_secular_kernel(int n, int s_lvl, int half, DcBufs b) {
  int m = blockIdx.x, r0, s;
  if (!merge_range(n, s_lvl, half, m, r0, s)) return;
  int k = b.meta[3 * m];
  int lane = threadIdx.x & 63;
  int i = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= k) return;
  const double* dl = b.dl + r0;
  const double* w = b.wv + r0;
  const double rho = b.mrho[2 * m];
  auto ev = [&](int o, double tau, int side, double* dpo, double* dpf, double* absum) {
    double f = 0.0, po = 0.0, pf = 0.0, ab = 0.0;
    const double dlo = dl[o];
    for (int j = lane; j < k; j += 64) {
      double del = (dl[j] - dlo) - tau;
      double t = w[j] / del;
      double term = rho * w[j] * t;
      f += term;
      ab += fabs(term);
      double dt = rho * t * t;
      bool origin_side = side > 0 ? (j <= o) : (j >= o);
      if (origin_side) po += dt; else pf += dt;
    }
    *dpo = wave_sum(po);
    *dpf = wave_sum(pf);
    *absum = 1.0 + wave_sum(ab);
    return 1.0 + wave_sum(f);
  };
  int o;
  double tv;
  dc_secular_root_with(i, k, dl, rho, b.mrho[2 * m + 1], ev, &o, &tv);
  if (lane == 0) {
    b.org[r0 + i] = o;
    b.tau[r0 + i] = tv;
  }
}

__device__ __forceinline__ double wave_prod(double v) {
  for (int o = 32; o > 0; o >>= 1) v *= __shfl_xor(v, o, 64);
  return v;
}

// zhat_i (Gu-Eisenstat, as dc_zhat) with one wave per i
__global__ __launch_bounds__(256) void dc_zhat_kernel(int n, int s_lvl, int half, DcBufs b) {
  int m = blockIdx.x, r0, s;
  if (!merge_range(n, s_lvl, half, m, r0, s)) return;
  int k = b.meta[3 * m];
  int lane = threadIdx.x & 63;
  int i = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= k) return;
  const double* dl = b.dl + r0;
  const int* org = b.org + r0;
  const double* tau = b.tau + r0;
  const double dli = dl[i];
  double p = 1.0;
  for (int j = lane; j < k; j += 64)
    if (j != i) p *= ((dli - dl[org[j]]) - tau[j]) / (dli - dl[j]);
  p = wave_prod(p) * ((dli - dl[org[i]]) - tau[i]);
  if (lane == 0) {
    double z2 = -p;
    if (z2 < 0.0) z2 = 0.0;
    b.zh[r0 + i] = copysign(sqrt(z2), b.wv[r0 + i]);
  }
}

// column j of the rank-one eigenvector matrix, normalised (one wave per column)
__global__ __launch_bounds__(256) void dc_vec_kernel(int n, int s_lvl, int half, double* __restrict__ Vb,
                                                     DcBufs b) {
  int m = blockIdx.x, r0, s;
  if (!merge_range(n, s_lvl, half, m, r0, s)) return;
  int k = b.meta[3 * m];
  int lane = threadIdx.x & 63;
  int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (j >= k) return;
  const double* dl = b.dl + r0;
  double lamo = dl[b.org[r0 + j]];
  double tj = b.tau[r0 + j];
  double s2 = 0.0;
  for (int i = lane; i < k; i += 64) {
    double v = b.zh[r0 + i] / ((dl[i] - lamo) - tj);
    s2 += v * v;
  }
  for (int o = 32; o > 0; o >>= 1) s2 += __shfl_xor(s2, o, 64);
  double inv = 1.0 / sqrt(s2);
  double* Vm = Vb + (size_t)r0 * s_lvl;
  for (int i = lane; i < k; i += 64)
    Vm[i + (size_t)j * k] = b.zh[r0 + i] / ((dl[i] - lamo) - tj) * inv;
}

// new block columns: nondeflated (gemm part, small merges only), deflated
// copies, and eigenvalues
__global__ void dc_apply_kernel(int n, int s_lvl, int half, const double* __restrict__ Q,
                                double* __restrict__ Qn, const double* __restrict__ Vb,
                                double* __restrict__ Dn, DcBufs b, int do_gemm) {
  int m = blockIdx.x, r0, s;
  if (!merge_range(n, s_lvl, half, m, r0, s)) return;
  int k = b.meta[3 * m];
  int64_t tot = (int64_t)s * s;
  for (int64_t e = (int64_t)blockIdx.y * blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.y * blockDim.x) {
    int r = (int)(e % s), jc = (int)(e / s);
    if (jc < k) {
      if (r == 0) Dn[r0 + jc] = b.dl[r0 + b.org[r0 + jc]] + b.tau[r0 + jc];
      if (!do_gemm) continue;
      const double* Vm = Vb + (size_t)r0 * s_lvl + (size_t)jc * k;
      double acc = 0.0;
      for (int t = 0; t < k; ++t) acc += Q[(size_t)b.gcol[r0 + t] * n + r0 + r] * Vm[t];
      Qn[(size_t)(r0 + jc) * n + r0 + r] = acc;
    } else {
      int t2 = jc - k;
      if (r == 0) Dn[r0 + jc] = b.dval[r0 + t2];
      Qn[(size_t)(r0 + jc) * n + r0 + r] = Q[(size_t)b.dcol[r0 + t2] * n + r0 + r];
    }
  }
}

// a trailing block with no sibling at this level is carried over unchanged
__global__ void dc_carry_kernel(int n, int r0, int s, const double* __restrict__ Q,
                                double* __restrict__ Qn, const double* __restrict__ D,
                                double* __restrict__ Dn) {
  int64_t tot = (int64_t)s * s;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    int r = (int)(e % s), c = (int)(e / s);
    Qn[(size_t)(r0 + c) * n + r0 + r] = Q[(size_t)(r0 + c) * n + r0 + r];
    if (r == 0) Dn[r0 + c] = D[r0 + c];
  }
}

__global__ void dc_table_kernel(int k, const int* __restrict__ gcol, int n, int64_t* __restrict__ tab) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < k) tab[t] = (int64_t)gcol[t] * n;
}

__global__ __launch_bounds__(1024) void dc_final_sort_kernel(int n, const double* __restrict__ D,
                                                             int* __restrict__ perm,
                                                             double* __restrict__ wout) {
  extern __shared__ double lds_d[];
  int np2 = 1;
  while (np2 < n) np2 <<= 1;
  double* key = lds_d;
  int* val = (int*)(key + np2);
  for (int i = threadIdx.x; i < np2; i += blockDim.x) {
    key[i] = i < n ? -D[i] : INFINITY;
    val[i] = i < n ? i : 0;
  }
  __syncthreads();
  bitonic_sort(key, val, np2);
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    perm[i] = val[i];
    wout[i] = -key[i];
  }
}

// out = min over p < k of perm[p] among those < kt (else kt), as a double
__global__ __launch_bounds__(256) void dc_first_kept_kernel(const int* __restrict__ perm, int k,
                                                            int kt, double* __restrict__ out) {
  __shared__ int sh[4];
  int v = kt;
  for (int p = threadIdx.x; p < k; p += 256) v = min(v, perm[p] < kt ? perm[p] : kt);
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = (double)min(min(sh[0], sh[1]), min(sh[2], sh[3]));
}

__global__ void gather_cols_kernel(int n, int k, const double* __restrict__ Q,
                                   const int* __restrict__ perm, double* __restrict__ U) {
  int64_t tot = (int64_t)n * k;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    int r = (int)(e % n), c = (int)(e / n);
    U[e] = Q[(size_t)perm[c] * n + r];
  }
}

struct Carver {
  char* base;
  size_t off = 0;
  template <typename T>
  T* take(size_t count) {
    off = (off + 255) & ~(size_t)255;
    T* p = (T*)(base + off);
    off += count * sizeof(T);
    return p;
  }
};

int set_lds_limit(itg_ctx* ctx) {
  static bool done = false;
  if (done) return ITG_OK;
  ITG_HIP(ctx, hipFuncSetAttribute((const void*)dc_prepare_kernel,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 8192 * 12));
  ITG_HIP(ctx, hipFuncSetAttribute((const void*)dc_final_sort_kernel,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 8192 * 12));
  ITG_HIP(ctx, hipFuncSetAttribute((const void*)trd_tail_kernel,
                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (TRD_TAIL * (TRD_TAIL + 1) + 2 * TRD_TAIL) * 8));
  done = true;
  return ITG_OK;
}

}  // namespace

static int eig_workspace(itg_ctx* ctx, size_t bytes, void** out) {
  if (bytes > ctx->eig_ws_bytes) {
    if (ctx->eig_ws) {
      ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
      ITG_HIP(ctx, hipFree(ctx->eig_ws));
      ctx->eig_ws = nullptr;
      ctx->eig_ws_bytes = 0;
    }
    if (hipMalloc(&ctx->eig_ws, bytes) != hipSuccess) {
      hipGetLastError();
      return itg_fail(ctx, ITG_EOOM, "eigensolver workspace of %zu bytes failed", bytes);
    }
    ctx->eig_ws_bytes = bytes;
  }
  *out = ctx->eig_ws;
  return ITG_OK;
}

// Tridiagonalise A (n x n FP64, full symmetric storage) in place: d, e out;
// reflectors left in A's strictly lower part (v_j(j+1) = 1 explicit), tau out.
static int tridiagonalize(itg_ctx* ctx, int n, double* A, double* d, double* e, double* tau,
                          double* VW, double* Ypart, double* ab, double* pnorm, double* pyv,
                          double* sc) {
  // VW = [V | W | V], n x NB each: the trailing update of a full panel is one
  // K = 2 NB product A22 -= [V W] [W V]^T
  double* Vp = VW;
  double* Wp = VW + (size_t)n * NB;
  double* V2 = VW + 2 * (size_t)n * NB;
  for (int p = 0; p < n; p += NB) {
    if (n - p <= TRD_TAIL) {
      // every earlier panel's update is applied: the rest in one block
      if (n - p >= 2) {
        const size_t lds = ((size_t)(n - p) * (n - p + 1) + 2 * (size_t)(n - p)) * sizeof(double);
        trd_tail_kernel<<<1, 1024, lds, ctx->stream>>>(A, n, p, d, e, tau);
      } else {
        trd_last_kernel<<<1, 1, 0, ctx->stream>>>(A, n, d, tau);
      }
      ITG_LAUNCH_CHECK(ctx);
      return ITG_OK;
    }
    int nbp = std::min(NB, n - p);
    const int npn = (n + TRD_ROWS - 1) / TRD_ROWS;
    int pn_off = 0, pn_cnt = npn;
    if (p < n - 1) {
      trd_colnorm_kernel<<<npn, 256, 0, ctx->stream>>>(A, n, p, pnorm);
      ITG_LAUNCH_CHECK(ctx);
    }
    for (int i = 0; i < nbp; ++i) {
      int j = p + i;
      if (j == n - 1) {
        trd_last_kernel<<<1, 1, 0, ctx->stream>>>(A, n, d, tau);
        ITG_LAUNCH_CHECK(ctx);
        break;
      }
      int m = n - j - 1;
      int T = (m + TS - 1) / TS;
      int ntiles = T * (T + 1) / 2;
      itg_tev tk = itg_tevents(ctx, ITG_K_TRD_SYMV);
      hipExtLaunchKernelGGL(trd_symv_kernel, dim3(ntiles + i), dim3(256), 0, ctx->stream,
                            tk.a, tk.b, 0, (const double*)A, n, j, i,
                            (const double*)(pnorm + pn_off), pn_cnt, Vp, (const double*)Wp, V2,
                            Ypart, ab, pyv, ntiles, d, e, tau, sc);
      ITG_LAUNCH_CHECK(ctx);
      // lower triangle of A22 read once, the panel pairs over m rows
      itg_tcommit(ctx, tk, ITG_K_TRD_SYMV, 2.0 * m * (m + 2.0 * i),
                  8.0 * (0.5 * m * (m + TS) + 2.0 * i * m + 2.0 * T * m));
      int next = (i + 1 < nbp) ? 1 : 0;
      const int b0 = (j + 1) / POST_ROWS;
      const int nblk = (n + POST_ROWS - 1) / POST_ROWS - b0;
      itg_tev tp = itg_tevents(ctx, ITG_K_TRD_POST);
      hipExtLaunchKernelGGL(trd_post_kernel, dim3(nblk), dim3(256), 0, ctx->stream, tp.a, tp.b, 0,
                            A, n, j, i, ntiles, (const double*)Vp, Wp, (const double*)Ypart, T,
                            (const double*)ab, (const double*)pyv, (const double*)tau,
                            (const double*)e, (const double*)sc, next, pnorm, b0);
      ITG_LAUNCH_CHECK(ctx);
      // the next column's reflector norm: partials of these blocks only
      pn_off = b0;
      pn_cnt = nblk;
      itg_tcommit(ctx, tp, ITG_K_TRD_POST, 8.0 * m * (i + 1), 8.0 * m * (2.0 * i + 5 + T));
    }
    int q = p + nbp;
    if (q < n) {
      int m2 = n - q;
      double* A22 = A + (size_t)q * n + q;
      int rc;
      // the tiled symv reads A22's 64 x 64 tiles on and below its diagonal
      // (diagonal tiles whole), the post kernel and the column norms the
      // lower part: the lower triangle plus one 128-tile band suffices
      if (nbp == NB) {
        rc = itg_gemm_internal(ctx, ITG_F64, ITG_MODE_NATIVE, 'N', 'T', m2, m2, 2 * NB, -1.0,
                               Vp + q, n, Wp + q, n, 1.0, A22, n, true);
        if (rc) return rc;
      } else {
        rc = itg_gemm_internal(ctx, ITG_F64, ITG_MODE_NATIVE, 'N', 'T', m2, m2, nbp, -1.0, Vp + q,
                               n, Wp + q, n, 1.0, A22, n, true);
        if (rc) return rc;
        rc = itg_gemm_internal(ctx, ITG_F64, ITG_MODE_NATIVE, 'N', 'T', m2, m2, nbp, -1.0, Wp + q,
                               n, Vp + q, n, 1.0, A22, n, true);
        if (rc) return rc;
      }
    }
  }
  return ITG_OK;
}

// Eigen-decomposition of the tridiagonal (d, e): eigenvalues D (unsorted)
// and eigenvectors as the columns of the returned buffer (Qa or Qb).
// With `top` non-null the last merge's eigenvector product is not launched
// but returned in *top (top->M = 0 if there is none), for the caller to run
// on the columns it keeps.
static int tridiag_dc(itg_ctx* ctx, int n, const double* d, const double* e, double* Qa,
                      double* Qb, double* Da, double* Db, double* Vb, DcBufs bufs, int64_t* tab,
                      double** Qout, double** Dout, GettArgs* top) {
  if (top) top->M = 0;
  dc_init_kernel<<<itg_blocks((int64_t)n * n, 256, 16384), 256, 0, ctx->stream>>>(n, d, e, Da, Qa, Qb);
  ITG_LAUNCH_CHECK(ctx);
  double *Q = Qa, *Qn = Qb, *D = Da, *Dn = Db;
  std::vector<int> meta;
  for (int half = 1; half < n; half *= 2) {
    int s_lvl = 2 * half;
    int nm = (n + s_lvl - 1) / s_lvl;
    size_t lds = (size_t)s_lvl * (sizeof(double) + sizeof(int) +
                                  (s_lvl <= DC_ZLDS ? sizeof(double) + sizeof(int) : 0));
    dc_prepare_kernel<<<nm, 1024, lds, ctx->stream>>>(n, s_lvl, half, D, Q, e, bufs);
    ITG_LAUNCH_CHECK(ctx);
    meta.assign(3 * nm, 0);
    ITG_HIP(ctx, hipMemcpyAsync(meta.data(), bufs.meta, 3 * nm * sizeof(int), hipMemcpyDeviceToHost,
                                ctx->stream));
    ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    dim3 g256(nm, (s_lvl + 255) / 256);
    dc_rotate_kernel<<<g256, 256, 0, ctx->stream>>>(n, s_lvl, half, Q, bufs);
    dim3 g4(nm, (s_lvl + 3) / 4);
    dc_secular_kernel<<<g4, 256, 0, ctx->stream>>>(n, s_lvl, half, bufs);
    dc_zhat_kernel<<<g4, 256, 0, ctx->stream>>>(n, s_lvl, half, bufs);
    dc_vec_kernel<<<dim3(nm, (s_lvl + 3) / 4), 256, 0, ctx->stream>>>(n, s_lvl, half, Vb, bufs);
    ITG_LAUNCH_CHECK(ctx);
    bool small = s_lvl <= SMALL_MERGE;
    if (!small) {
      for (int m = 0; m < nm; ++m) {
        int r0 = m * s_lvl, s = std::min(s_lvl, n - r0), k = meta[3 * m];
        if (s <= half || k == 0) continue;
        dc_table_kernel<<<(k + 255) / 256, 256, 0, ctx->stream>>>(k, bufs.gcol + r0, n, tab);
        // Qn[r0.., r0+j] (j < k) = sum_t Q[r0.., gcol[t]] V[t, j], computed as
        // C^T(j, r) = V^T(j, t) A^T(t, r) so the stores run along rows
        GettArgs g = {};
        g.A = Vb + (size_t)r0 * s_lvl;
        g.B = Q + r0;
        g.C = Qn + (size_t)r0 * n + r0;
        g.M = k; g.N = s; g.K = k;
        g.sAm = k; g.sAk = 1;
        g.sBk = -1; g.tBk = tab; g.sBn = 1;
        g.sCm = n; g.sCn = 1;
        g.alpha = 1.0; g.beta = 0.0;
        if (top && s_lvl >= n) {
          *top = g;   // the last merge (nm == 1, r0 == 0)
          continue;
        }
        int rc = itg_gett_launch(ctx, ITG_F64, ITG_MODE_NATIVE, &g, true, false);
        if (rc) return rc;
      }
    }
    dim3 gap(nm, (unsigned)std::min<int64_t>(((int64_t)s_lvl * s_lvl + 255) / 256, 8192));
    dc_apply_kernel<<<gap, 256, 0, ctx->stream>>>(n, s_lvl, half, Q, Qn, Vb, Dn, bufs, small ? 1 : 0);
    ITG_LAUNCH_CHECK(ctx);
    int r0l = (nm - 1) * s_lvl, sl = n - r0l;
    if (sl <= half) {
      dc_carry_kernel<<<itg_blocks((int64_t)sl * sl, 256, 4096), 256, 0, ctx->stream>>>(n, r0l, sl, Q, Qn, D, Dn);
      ITG_LAUNCH_CHECK(ctx);
    }
    std::swap(Q, Qn);
    std::swap(D, Dn);
  }
  *Qout = Q;
  *Dout = D;
  return ITG_OK;
}

// U (n x k, FP64) <- Q U with Q = H_0 H_1 ... H_{nrefl-1}, reflectors stored
// in A (ld n) as described at bt_build_v_kernel
static int back_transform(itg_ctx* ctx, int n, int nrefl, int off, int k, const double* A,
                          const double* tau, double* U, double* Vb, double* G, double* T,
                          double* X, double* X2) {
  if (nrefl < 1) return ITG_OK;
  int last = ((nrefl - 1) / NB) * NB;
  for (int p = last; p >= 0; p -= NB) {
    int nbp = std::min(NB, nrefl - p);
    int m2 = n - p - off;
    if (nbp <= 0 || m2 <= 0) continue;
    bt_build_v_kernel<<<itg_blocks((int64_t)n * nbp, 256, 4096), 256, 0, ctx->stream>>>(A, n, p, nbp, off, Vb);
    ITG_LAUNCH_CHECK(ctx);
    const double* V1 = Vb + (p + off);
    int rc = itg_gemm_internal(ctx, ITG_F64, ITG_MODE_NATIVE, 'T', 'N', nbp, nbp, m2, 1.0, V1, n,
                               V1, n, 0.0, G, nbp);
    if (rc) return rc;
    bt_larft_kernel<<<1, 64, 0, ctx->stream>>>(nbp, G, tau + p, T);
    ITG_LAUNCH_CHECK(ctx);
    rc = itg_gemm_internal(ctx, ITG_F64, ITG_MODE_NATIVE, 'T', 'N', nbp, k, m2, 1.0, V1, n,
                           U + (p + off), n, 0.0, X, nbp);
    if (rc) return rc;
    rc = itg_gemm_internal(ctx, ITG_F64, ITG_MODE_NATIVE, 'N', 'N', nbp, k, nbp, 1.0, T, nbp, X,
                           nbp, 0.0, X2, nbp);
    if (rc) return rc;
    rc = itg_gemm_internal(ctx, ITG_F64, ITG_MODE_NATIVE, 'N', 'N', m2, k, nbp, -1.0, V1, n, X2,
                           nbp, 1.0, U + (p + off), n);
    if (rc) return rc;
  }
  return ITG_OK;
}

// U (n x k) <- Q U for Q = H_0 ... H_{n-off-1}, reflector j stored in
// column j of A from row j + off on (the tridiagonalisation: off = 1, its
// rows 0..j cleared here and every T formed up front by bt_tall_kernel; the
// two-stage band reduction: off = its bandwidth, A already laid out and
// Tall given). Groups of BTG panels go last-to-first: Y = V_G^T U,
// Gg = V_G^T V_G, the group solve for X2, U -= V_G X2.
static int back_transform_trd(itg_ctx* ctx, int n, int k, double* A, const double* tau,
                              double* U, double* Tall, double* Y, double* X2, double* Gg,
                              int off = 1, bool have_t = false) {
  int nrefl = n - off;
  if (nrefl < 1 || k < 1) return ITG_OK;
  if (!have_t) {
    bt_clear_upper_kernel<<<itg_blocks((int64_t)n * n, 256, 16384), 256, 0, ctx->stream>>>(A, n);
    int npan = (nrefl + NB - 1) / NB;
    bt_tall_kernel<<<npan, 256, 0, ctx->stream>>>(A, n, nrefl, tau, Tall);
    ITG_LAUNCH_CHECK(ctx);
  }
  const int gsz = BTG * NB;
  int ngrp = (nrefl + gsz - 1) / gsz;
  for (int g = ngrp - 1; g >= 0; --g) {
    int p0 = g * gsz;
    int nbg = std::min(gsz, nrefl - p0);
    int m2 = n - p0 - off;
    const double* VG = A + (size_t)p0 * n + (p0 + off);   // rows p0+off.., columns p0..p0+nbg
    int rc = itg_gemm_internal(ctx, ITG_F64, ITG_MODE_NATIVE, 'T', 'N', nbg, k, m2, 1.0, VG, n,
                               U + (p0 + off), n, 0.0, Y, nbg);
    if (rc) return rc;
    rc = itg_gemm_internal(ctx, ITG_F64, ITG_MODE_NATIVE, 'T', 'N', nbg, nbg, m2, 1.0, VG, n, VG,
                           n, 0.0, Gg, nbg);
    if (rc) return rc;
    bt_group_solve_kernel<<<(k + 7) / 8, 256, 0, ctx->stream>>>(
        nbg, k, Y, Gg, Tall + (size_t)(p0 / NB) * NB * NB, X2);
    ITG_LAUNCH_CHECK(ctx);
    rc = itg_gemm_internal(ctx, ITG_F64, ITG_MODE_NATIVE, 'N', 'N', m2, k, nbg, -1.0, VG, n, X2,
                           nbg, 1.0, U + (p0 + off), n);
    if (rc) return rc;
  }
  return ITG_OK;
}

extern "C" int itg_eigh_trunc(itg_ctx* ctx, int dtype, int64_t n64, void* Ain, double* w,
                              void* Uout, int64_t maxdim, double cutoff, int abs_cut, int rel_cut,
                              int64_t* kout, double* truncerr, double* docut) {
  ITG_CHECK_CTX(ctx);
  const bool cplx = dtype == ITG_C64;   // Hermitian (heevd's route)
  if (!cplx) ITG_CHECK_DTYPE(ctx, dtype);
  if (n64 < 0) return itg_fail(ctx, ITG_EDIM, "eigh: negative dimension");
  if (n64 > 8192) return itg_fail(ctx, ITG_EUNSUPPORTED, "eigh: n = %lld > 8192", (long long)n64);
  if (!kout) return itg_fail(ctx, ITG_EARG, "eigh: null k");
  if (n64 == 0) {
    *kout = 0;
    if (truncerr) *truncerr = 0.0;
    if (docut) *docut = 0.0;
    return ITG_OK;
  }
  int rc = set_lds_limit(ctx);
  if (rc) return rc;
  int n = (int)n64;
  size_t nn = (size_t)n * n;
  bool f32 = dtype == ITG_F32;
  // the two-stage reduction (dense -> band -> tridiagonal) on request
  // (ITG_EIGH_TWOSTAGE=1, read per call): correct, but at n = 4096 it takes
  // 113 ms against the one-stage reduction's 77 ms, its bulge chasing and
  // tall-panel QRs being latency-bound
  const char* two_env = getenv("ITG_EIGH_TWOSTAGE");
  const bool two = !cplx && n >= TWO_STAGE_MIN && two_env && two_env[0] == '1';
  const size_t Tw = (size_t)(n + 63) / 64;   // tile rows of the tiled symv
  size_t bytes = (3 * nn + (f32 ? 2 * nn : 0)) * sizeof(double) + 3 * (size_t)n * NB * 8 +
                 (2 * (size_t)BTG * NB * n + (size_t)BTG * NB * BTG * NB) * 8 +
                 (size_t)((n + NB - 1) / NB) * NB * NB * 8 +
                 (size_t)n * 8 * 26 + 4 * (size_t)n * NB * 8 + 4 * NB * NB * 8 + 64 * 1024 +
                 (Tw * n + Tw * (Tw + 1) / 2 + 64) * 8 + 64 * 256 +
                 ((size_t)n / POST_ROWS + 1) * 8 +
                 (two ? eig2::doubles_needed(n) * 8 + 256 : 0) +
                 (cplx ? (nn + itg_zhetrd_scratch_doubles(n) + 2 * (size_t)n) * 8 + 512 : 0);
  void* ws;
  rc = eig_workspace(ctx, bytes, &ws);
  if (rc) return rc;
  Carver cv{(char*)ws};
  double* Qa = cv.take<double>(nn);
  double* Qb = cv.take<double>(nn);
  // D&C merge vectors; afterwards the back-transform's compact-WY factors
  double* Vbuf = cv.take<double>(std::max(nn, (size_t)((n + NB - 1) / NB) * NB * NB));
  // complex: A stays the caller's (complex) matrix, U a real n x k staging
  // buffer for the tridiagonal's eigenvectors
  double* A = f32 ? cv.take<double>(nn) : cplx ? nullptr : (double*)Ain;
  double* U = (f32 || cplx) ? cv.take<double>(nn) : (double*)Uout;
  double* zscr = cplx ? cv.take<double>(itg_zhetrd_scratch_doubles(n)) : nullptr;
  double* ztau = cplx ? cv.take<double>(2 * (size_t)n) : nullptr;
  const size_t Tmax = (size_t)(n + TS - 1) / TS;
  double* VW = cv.take<double>(3 * (size_t)n * NB);
  double* Ypart = cv.take<double>(Tmax * n);
  double* ab = cv.take<double>(2 * NB);
  double* pnorm = cv.take<double>((size_t)n / POST_ROWS + 1);
  double* pyv = cv.take<double>(Tmax * (Tmax + 1) / 2 + 64);
  double* trd_sc = cv.take<double>(2);
  double* d = cv.take<double>(n);
  double* e = cv.take<double>(n);
  double* tau = cv.take<double>(n);
  double* Da = cv.take<double>(n);
  double* Db = cv.take<double>(n);
  DcBufs b;
  b.ds = cv.take<double>(n); b.zs = cv.take<double>(n); b.dl = cv.take<double>(n);
  b.wv = cv.take<double>(n); b.dval = cv.take<double>(n); b.rcs = cv.take<double>(2 * (size_t)n);
  b.tau = cv.take<double>(n); b.zh = cv.take<double>(n); b.mrho = cv.take<double>(2 * (size_t)n);
  b.cs = cv.take<int>(n); b.gcol = cv.take<int>(n); b.dcol = cv.take<int>(n);
  b.rcol = cv.take<int>(2 * (size_t)n); b.org = cv.take<int>(n); b.meta = cv.take<int>(3 * (size_t)n);
  int* perm = cv.take<int>(n);
  int64_t* tab = cv.take<int64_t>(n);
  double* Gm = cv.take<double>(NB * NB);
  double* Tm = cv.take<double>(NB * NB);
  double* Ygrp = cv.take<double>((size_t)BTG * NB * n);
  double* X2grp = cv.take<double>((size_t)BTG * NB * n);
  double* Ggrp = cv.take<double>((size_t)BTG * NB * BTG * NB);
  double* out3 = cv.take<double>(4);
  eig2::Bufs eb{};
  if (two) eig2::carve(n, cv.take<double>(eig2::doubles_needed(n)), &eb);
  if (cv.off > bytes) return itg_fail(ctx, ITG_EOOM, "eigh: workspace layout overflow");
  if (f32) {
    rc = itg_convert(ctx, (int64_t)nn, ITG_F32, Ain, ITG_F64, A);
    if (rc) return rc;
  }
  // 1. tridiagonalise
  if (cplx) rc = itg_zhetrd(ctx, n, Ain, d, e, ztau, zscr);
  else if (two) rc = eig2::reduce(ctx, n, A, d, e, eb);
  else rc = tridiagonalize(ctx, n, A, d, e, tau, VW, Ypart, ab, pnorm, pyv, trd_sc);
  if (rc) return rc;
  // 2. divide and conquer
  double *Q, *D;
  GettArgs top = {};
  rc = tridiag_dc(ctx, n, d, e, Qa, Qb, Da, Db, Vbuf, b, tab, &Q, &D, &top);
  if (rc) return rc;
  // 3. sort descending, truncate
  {
    int np2 = 1;
    while (np2 < n) np2 <<= 1;
    dc_final_sort_kernel<<<1, 1024, (size_t)np2 * 12, ctx->stream>>>(n, D, perm, w);
    ITG_LAUNCH_CHECK(ctx);
  }
  int64_t k = n;
  double te = 0.0, dc = 0.0;
  if (cutoff >= 0.0 || (maxdim > 0 && maxdim < n)) {
    rc = itg_truncate_device_async(ctx, w, n, maxdim, 1, cutoff < 0.0 ? 0.0 : cutoff, abs_cut,
                                   rel_cut, ITG_TRUNC_NDTENSORS, out3);
    if (rc) return rc;
    double h[3];
    ITG_HIP(ctx, hipMemcpyAsync(h, out3, 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    k = (int64_t)h[0];
    te = h[1];
    dc = h[2];
  }
  // the last merge's eigenvector product, for its non-deflated columns from
  // the first one kept on (Q columns [0, top.M) are the non-deflated ones;
  // their eigenvalues increase with the column, so the kept ones are a tail)
  if (top.M > 0) {
    int64_t j0 = 0;
    if (k < n) {
      dc_first_kept_kernel<<<1, 256, 0, ctx->stream>>>(perm, (int)k, (int)top.M, out3 + 3);
      ITG_LAUNCH_CHECK(ctx);
      double h0;
      ITG_HIP(ctx, hipMemcpyAsync(&h0, out3 + 3, sizeof(double), hipMemcpyDeviceToHost,
                                  ctx->stream));
      ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
      j0 = (int64_t)h0;
    }
    if (j0 < top.M) {
      GettArgs g = top;
      g.A = (const double*)top.A + (size_t)j0 * top.sAm;
      g.C = (double*)top.C + (size_t)j0 * top.sCm;
      g.M = top.M - j0;
      rc = itg_gett_launch(ctx, ITG_F64, ITG_MODE_NATIVE, &g, true, false);
      if (rc) return rc;
    }
  }
  // 4. kept eigenvectors: gather, back-transform
  gather_cols_kernel<<<itg_blocks((int64_t)n * k, 256, 16384), 256, 0, ctx->stream>>>(n, (int)k, Q, perm, U);
  ITG_LAUNCH_CHECK(ctx);
  if (cplx) {
    rc = itg_zunmtr(ctx, n, (int)k, Ain, ztau, U, Uout);
  } else if (two) {
    rc = eig2::apply_q2(ctx, n, (int)k, U, eb);
    if (rc) return rc;
    rc = back_transform_trd(ctx, n, (int)k, A, nullptr, U, eb.Tall, Ygrp, X2grp, Ggrp, eig2::SB,
                            true);
  } else {
    rc = back_transform_trd(ctx, n, (int)k, A, tau, U, Vbuf, Ygrp, X2grp, Ggrp);
  }
  if (rc) return rc;
  if (f32) {
    rc = itg_convert(ctx, (int64_t)n * k, ITG_F64, U, ITG_F32, Uout);
    if (rc) return rc;
  }
  *kout = k;
  if (truncerr) *truncerr = te;
  if (docut) *docut = dc;
  return ITG_OK;
}

// ====================================================================== QR
// Householder QR (culinearalgebra.jl:110-121): reflectors of W = A column by
// column, then Q = H_0...H_{r-1} [I; 0] by the same compact-WY
// back-transform, R = triu(W).
namespace {

__global__ __launch_bounds__(1024) void qr_reflect_kernel(double* __restrict__ W, int m, int j,
                                                          double* __restrict__ tau) {
  __shared__ double red[16];
  __shared__ double sc;
  double* a = W + (size_t)j * m;
  double part = 0.0;
  for (int r = j + 1 + threadIdx.x; r < m; r += blockDim.x) part += a[r] * a[r];
  double xn2 = block_sum(part, red);
  if (threadIdx.x == 0) {
    double alpha = a[j], beta = alpha, tv = 0.0, scale = 1.0;
    if (xn2 != 0.0) {
      beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
      tv = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
    a[j] = beta;
    tau[j] = tv;
    sc = scale;
  }
  __syncthreads();
  double scale = sc;
  for (int r = j + 1 + threadIdx.x; r < m; r += blockDim.x) a[r] *= scale;
}

// W[j:m, c] -= tau_j v (v^T W[j:m, c]) for j < c < nend (the rest of the
// panel), one 256-thread block per column
__global__ __launch_bounds__(256) void qr_apply_kernel(double* __restrict__ W, int m, int nend,
                                                       int j, const double* __restrict__ tau) {
  __shared__ double sh4[4];
  const int c = j + 1 + blockIdx.x;
  if (c >= nend) return;
  const double tj = tau[j];
  if (tj == 0.0) return;
  const double* v = W + (size_t)j * m;
  double* col = W + (size_t)c * m;
  double s = 0.0;
  for (int r = j + threadIdx.x; r < m; r += 256) s += (r == j ? 1.0 : v[r]) * col[r];
  s = block_sum256(s, sh4);
  const double f = tj * s;
  for (int r = j + threadIdx.x; r < m; r += 256) col[r] -= f * (r == j ? 1.0 : v[r]);
}

__global__ void qr_init_q_kernel(int m, int r, double* __restrict__ Q) {
  int64_t tot = (int64_t)m * r;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x)
    Q[e] = ((e % m) == (e / m)) ? 1.0 : 0.0;
}

__global__ void qr_extract_r_kernel(int m, int n, int r, const double* __restrict__ W,
                                    double* __restrict__ R) {
  int64_t tot = (int64_t)r * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    int i = (int)(e % r), c = (int)(e / r);
    R[e] = i <= c ? W[i + (size_t)c * m] : 0.0;
  }
}

// ------------------------------------------------------- one-sided Jacobi
__global__ __launch_bounds__(256) void jac_round_kernel(double* __restrict__ W, int m,
                                                        double* __restrict__ V, int nv, int ncols,
                                                        int N, int rnd, double tol,
                                                        double small2,
                                                        double* __restrict__ offs) {
  __shared__ double sh[3][4];
  int i = blockIdx.x;
  int p, q;
  if (i == 0) { p = rnd; q = N - 1; }
  else { p = (rnd + i) % (N - 1); q = (rnd - i + (N - 1)) % (N - 1); }
  if (p >= ncols || q >= ncols) return;
  double* a = W + (size_t)p * m;
  double* b = W + (size_t)q * m;
  double sa = 0.0, sb = 0.0, sg = 0.0;
  for (int r = threadIdx.x; r < m; r += blockDim.x) {
    double x = a[r], y = b[r];
    sa += x * x; sb += y * y; sg += x * y;
  }
  for (int o = 32; o > 0; o >>= 1) {
    sa += __shfl_down(sa, o, 64);
    sb += __shfl_down(sb, o, 64);
    sg += __shfl_down(sg, o, 64);
  }
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) { sh[0][w] = sa; sh[1][w] = sb; sh[2][w] = sg; }
  __syncthreads();
  sa = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
  sb = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
  sg = sh[2][0] + sh[2][1] + sh[2][2] + sh[2][3];
  // a column at the rounding level of ||A||_F has no direction left to fix
  // (its cosines stay O(1) noise): such pairs count as converged
  double rel = (sa > small2 && sb > small2) ? fabs(sg) / sqrt(sa * sb) : 0.0;
  if (threadIdx.x == 0) offs[i] = fmax(offs[i], rel);
  if (rel <= tol) return;
  double zeta = (sb - sa) / (2.0 * sg);
  double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
  double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
  for (int r = threadIdx.x; r < m; r += blockDim.x) {
    double x = a[r], y = b[r];
    a[r] = c * x - s * y;
    b[r] = s * x + c * y;
  }
  double* va = V + (size_t)p * nv;
  double* vb = V + (size_t)q * nv;
  for (int r = threadIdx.x; r < nv; r += blockDim.x) {
    double x = va[r], y = vb[r];
    va[r] = c * x - s * y;
    vb[r] = s * x + c * y;
  }
}

__global__ __launch_bounds__(256) void col_norm_kernel(const double* __restrict__ W, int m, int ncols,
                                                       double* __restrict__ out) {
  int lane = threadIdx.x & 63;
  int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= ncols) return;
  const double* col = W + (size_t)c * m;
  double s = 0.0;
  for (int r = lane; r < m; r += 64) s += col[r] * col[r];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) out[c] = sqrt(s);
}

__global__ void square_kernel(int n, const double* __restrict__ x, double* __restrict__ y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i] * x[i];
}

// Uo[:, i] = W[:, perm[i]] / s_i (zero column if s_i == 0), Vo[:, i] = V[:, perm[i]]
__global__ void svd_gather_kernel(int m, int nv, int k, const double* __restrict__ W,
                                  const double* __restrict__ V, const int* __restrict__ perm,
                                  const double* __restrict__ s, double* __restrict__ Uo,
                                  double* __restrict__ Vo, int ldu, int ldv) {
  int64_t tot = (int64_t)(m > nv ? m : nv) * k;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t rows = m > nv ? m : nv;
    int r = (int)(e % rows), i = (int)(e / rows);
    int c = perm[i];
    if (r < m) Uo[r + (size_t)i * ldu] = s[i] > 0.0 ? W[r + (size_t)c * m] / s[i] : 0.0;
    if (r < nv) Vo[r + (size_t)i * ldv] = V[r + (size_t)c * nv];
  }
}

__global__ void transpose_kernel(int m, int n, const double* __restrict__ A, double* __restrict__ B) {
  int64_t tot = (int64_t)m * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    int i = (int)(e % m), j = (int)(e / m);
    B[j + (size_t)i * n] = A[e];
  }
}

}  // namespace

extern "C" int itg_qr(itg_ctx* ctx, int dtype, int64_t m64, int64_t n64, const void* Ain, void* Qo,
                      void* Ro) {
  ITG_CHECK_CTX(ctx);
  if (dtype != ITG_C64) ITG_CHECK_DTYPE(ctx, dtype);
  if (m64 < 0 || n64 < 0) return itg_fail(ctx, ITG_EDIM, "qr: negative dimension");
  if (m64 * n64 > ((int64_t)1 << 31)) return itg_fail(ctx, ITG_EUNSUPPORTED, "qr: matrix too large");
  if (m64 == 0 || n64 == 0) return ITG_OK;
  if (dtype == ITG_C64) return itg_c64_qr(ctx, m64, n64, Ain, Qo, Ro);
  int m = (int)m64, n = (int)n64, r = std::min(m, n);
  bool f32 = dtype == ITG_F32;
  size_t bytes = ((size_t)m * n + (size_t)m * r + (size_t)r * n + (size_t)m * NB + 2 * (size_t)NB * n +
                  2 * NB * NB + (size_t)r + 64) * sizeof(double) + 16 * 256;
  void* ws;
  int rc = eig_workspace(ctx, bytes, &ws);
  if (rc) return rc;
  Carver cv{(char*)ws};
  double* W = cv.take<double>((size_t)m * n);
  double* Q = cv.take<double>((size_t)m * r);
  double* R = cv.take<double>((size_t)r * n);
  double* Vb = cv.take<double>((size_t)m * NB);
  double* X = cv.take<double>((size_t)NB * n);
  double* X2 = cv.take<double>((size_t)NB * n);
  double* Gm = cv.take<double>(NB * NB);
  double* Tm = cv.take<double>(NB * NB);
  double* tau = cv.take<double>(r);
  rc = itg_convert(ctx, (int64_t)m * n, dtype, Ain, ITG_F64, W);
  if (rc) return rc;
  // blocked Householder QR (dgeqrf): each panel of NB columns is factored
  // column by column within itself, then the trailing columns take the
  // panel's block reflector, W2 -= V T^T (V^T W2), as GEMMs
  for (int p = 0; p < r; p += NB) {
    const int nbp = std::min(NB, r - p), pe = p + nbp;
    for (int j = p; j < pe; ++j) {
      qr_reflect_kernel<<<1, 1024, 0, ctx->stream>>>(W, m, j, tau);
      if (j + 1 < pe)
        qr_apply_kernel<<<pe - j - 1, 256, 0, ctx->stream>>>(W, m, pe, j, tau);
      ITG_LAUNCH_CHECK(ctx);
    }
    const int nt = n - pe, m2 = m - p;
    if (nt <= 0) continue;
    bt_build_v_kernel<<<itg_blocks((int64_t)m * nbp, 256, 4096), 256, 0, ctx->stream>>>(W, m, p, nbp,
                                                                                         0, Vb);
    ITG_LAUNCH_CHECK(ctx);
    const double* V1 = Vb + p;
    double* W2 = W + (size_t)pe * m + p;
    rc = itg_gemm_internal(ctx, ITG_F64, ITG_MODE_NATIVE, 'T', 'N', nbp, nbp, m2, 1.0, V1, m, V1,
                           m, 0.0, Gm, nbp);
    if (rc) return rc;
    bt_larft_kernel<<<1, 64, 0, ctx->stream>>>(nbp, Gm, tau + p, Tm);
    ITG_LAUNCH_CHECK(ctx);
    rc = itg_gemm_internal(ctx, ITG_F64, ITG_MODE_NATIVE, 'T', 'N', nbp, nt, m2, 1.0, V1, m, W2,
                           m, 0.0, X, nbp);
    if (rc) return rc;
    rc = itg_gemm_internal(ctx, ITG_F64, ITG_MODE_NATIVE, 'T', 'N', nbp, nt, nbp, 1.0, Tm, nbp, X,
                           nbp, 0.0, X2, nbp);
    if (rc) return rc;
    rc = itg_gemm_internal(ctx, ITG_F64, ITG_MODE_NATIVE, 'N', 'N', m2, nt, nbp, -1.0, V1, m, X2,
                           nbp, 1.0, W2, m);
    if (rc) return rc;
  }
  qr_init_q_kernel<<<itg_blocks((int64_t)m * r, 256, 8192), 256, 0, ctx->stream>>>(m, r, Q);
  ITG_LAUNCH_CHECK(ctx);
  rc = back_transform(ctx, m, r, 0, r, W, tau, Q, Vb, Gm, Tm, X, X2);
  if (rc) return rc;
  qr_extract_r_kernel<<<itg_blocks((int64_t)r * n, 256, 8192), 256, 0, ctx->stream>>>(m, n, r, W, R);
  ITG_LAUNCH_CHECK(ctx);
  rc = itg_convert(ctx, (int64_t)m * r, ITG_F64, Q, dtype, Qo);
  if (rc) return rc;
  return itg_convert(ctx, (int64_t)r * n, ITG_F64, R, dtype, Ro);
  (void)f32;
}

// smallest column count at which the Jacobi SVD is preconditioned by the
// Gram matrix's eigenvectors
constexpr int SVD_PRECOND_MIN = 64;

extern "C" int itg_svd_trunc(itg_ctx* ctx, int dtype, int64_t m64, int64_t n64, void* Ain, void* Uo,
                             double* So, void* Vo, int64_t maxdim, int64_t mindim, double cutoff,
                             int abs_cut, int rel_cut, int64_t* kout, double* truncerr,
                             double* docut) {
  ITG_CHECK_CTX(ctx);
  if (dtype != ITG_C64) ITG_CHECK_DTYPE(ctx, dtype);
  if (m64 < 0 || n64 < 0) return itg_fail(ctx, ITG_EDIM, "svd: negative dimension");
  if (!kout) return itg_fail(ctx, ITG_EARG, "svd: null k");
  if (m64 == 0 || n64 == 0) {
    *kout = 0;
    if (truncerr) *truncerr = 0.0;
    if (docut) *docut = 0.0;
    return ITG_OK;
  }
  if (dtype == ITG_C64)
    return itg_c64_svd_trunc(ctx, m64, n64, Ain, Uo, So, Vo, maxdim, mindim, cutoff, abs_cut,
                             rel_cut, kout, truncerr, docut);
  if (std::min(m64, n64) > 8192 || m64 * n64 > ((int64_t)1 << 31))
    return itg_fail(ctx, ITG_EUNSUPPORTED, "svd: matrix too large");
  int rc = set_lds_limit(ctx);
  if (rc) return rc;
  int m = (int)m64, n = (int)n64;
  bool tr = m < n;                 // work on A^T so that rows >= cols
  int rows = tr ? n : m, cols = tr ? m : n;
  int N = cols + (cols & 1);
  // preconditioning (below) runs the eigensolver, which owns eig_ws: the
  // SVD's buffers live in their own grow-only workspace
  const bool precond = cols >= SVD_PRECOND_MIN;
  size_t bytes = ((size_t)rows * cols * 2 + (size_t)cols * cols + (size_t)rows * cols +
                  (size_t)cols * cols + 8 * (size_t)N + 64) * sizeof(double) + 16 * 256 +
                 (precond ? ((size_t)rows * cols + (size_t)cols * cols + cols) * 8 + 3 * 256 : 0);
  if (bytes > ctx->svd_ws_bytes) {
    if (ctx->svd_ws) {
      ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
      ITG_HIP(ctx, hipFree(ctx->svd_ws));
      ctx->svd_ws = nullptr;
      ctx->svd_ws_bytes = 0;
    }
    if (hipMalloc(&ctx->svd_ws, bytes) != hipSuccess) {
      hipGetLastError();
      return itg_fail(ctx, ITG_EOOM, "svd workspace of %zu bytes failed", bytes);
    }
    ctx->svd_ws_bytes = bytes;
  }
  void* ws = ctx->svd_ws;
  Carver cv{(char*)ws};
  double* A64 = cv.take<double>((size_t)m * n);
  double* W = cv.take<double>((size_t)rows * cols);
  double* V = cv.take<double>((size_t)cols * cols);
  double* Ub = cv.take<double>((size_t)rows * cols);
  double* Vbo = cv.take<double>((size_t)cols * cols);
  double* offs = cv.take<double>(N);
  double* sig = cv.take<double>(N);
  double* sig_sorted = cv.take<double>(N);
  double* P = cv.take<double>(N);
  int* perm = cv.take<int>(N);
  double* out3 = cv.take<double>(4);
  rc = itg_convert(ctx, (int64_t)m * n, dtype, Ain, ITG_F64, A64);
  if (rc) return rc;
  if (tr) transpose_kernel<<<itg_blocks((int64_t)m * n, 256, 8192), 256, 0, ctx->stream>>>(m, n, A64, W);
  else ITG_HIP(ctx, hipMemcpyAsync(W, A64, (size_t)m * n * 8, hipMemcpyDeviceToDevice, ctx->stream));
  if (precond) {
    // Jacobi preconditioned by the Gram matrix's eigenvectors: W <- W V0
    // with W^T W = V0 diag(w) V0^T, V <- V0. The columns of W V0 are
    // orthogonal up to eps ||A||^2 / (s_i s_j), so the one-sided sweeps
    // below (which keep their relative accuracy) converge in a few sweeps
    double* G = cv.take<double>((size_t)cols * cols);
    double* W2 = cv.take<double>((size_t)rows * cols);
    double* wG = cv.take<double>(cols);
    if (cv.off > bytes) return itg_fail(ctx, ITG_EOOM, "svd: workspace layout overflow");
    // (one operand buffer: the exactly symmetric product)
    rc = itg_gemm(ctx, ITG_F64, ITG_MODE_NATIVE, 'T', 'N', cols, cols, rows, 1.0, W, rows, W, rows,
                  0.0, G, cols);
    if (rc) return rc;
    int64_t kg = 0;
    double teg = 0.0, dcg = 0.0;
    rc = itg_eigh_trunc(ctx, ITG_F64, cols, G, wG, V, 0, -1.0, 0, 1, &kg, &teg, &dcg);
    if (rc) return rc;
    rc = itg_gemm_internal(ctx, ITG_F64, ITG_MODE_NATIVE, 'N', 'N', rows, cols, cols, 1.0, W, rows,
                           V, cols, 0.0, W2, rows);
    if (rc) return rc;
    W = W2;
  } else {
    qr_init_q_kernel<<<itg_blocks((int64_t)cols * cols, 256, 8192), 256, 0, ctx->stream>>>(cols, cols, V);
    ITG_LAUNCH_CHECK(ctx);
  }
  double tol = 4.0 * ITG_DC_EPS * std::sqrt((double)rows);
  double fro = 0.0;
  rc = itg_nrm2(ctx, ITG_F64, (int64_t)m * n, A64, &fro);
  if (rc) return rc;
  double small2 = (ITG_DC_EPS * fro) * (ITG_DC_EPS * fro);
  std::vector<double> hoffs(N / 2);
  bool conv = cols < 2;
  for (int sweep = 0; sweep < 60 && !conv; ++sweep) {
    ITG_HIP(ctx, hipMemsetAsync(offs, 0, (N / 2) * sizeof(double), ctx->stream));
    for (int rnd = 0; rnd < N - 1; ++rnd) {
      jac_round_kernel<<<N / 2, 256, 0, ctx->stream>>>(W, rows, V, cols, cols, N, rnd, tol, small2,
                                                       offs);
    }
    ITG_LAUNCH_CHECK(ctx);
    ITG_HIP(ctx, hipMemcpyAsync(hoffs.data(), offs, (N / 2) * sizeof(double),
                                hipMemcpyDeviceToHost, ctx->stream));
    ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double mx = 0.0;
    for (double x : hoffs) mx = std::max(mx, x);
    conv = mx <= tol;
  }
  if (!conv) return itg_fail(ctx, ITG_ECONV, "svd: Jacobi sweeps did not converge");
  col_norm_kernel<<<(cols + 3) / 4, 256, 0, ctx->stream>>>(W, rows, cols, sig);
  int np2 = 1;
  while (np2 < cols) np2 <<= 1;
  dc_final_sort_kernel<<<1, 1024, (size_t)np2 * 12, ctx->stream>>>(cols, sig, perm, sig_sorted);
  square_kernel<<<(cols + 255) / 256, 256, 0, ctx->stream>>>(cols, sig_sorted, P);
  ITG_LAUNCH_CHECK(ctx);
  rc = itg_truncate_device_async(ctx, P, cols, maxdim, mindim, cutoff, abs_cut, rel_cut,
                                 ITG_TRUNC_NDTENSORS, out3);
  if (rc) return rc;
  double h[3];
  ITG_HIP(ctx, hipMemcpyAsync(h, out3, 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  int k = (int)h[0];
  svd_gather_kernel<<<itg_blocks((int64_t)std::max(rows, cols) * k, 256, 8192), 256, 0, ctx->stream>>>(
      rows, cols, k, W, V, perm, sig_sorted, Ub, Vbo, rows, cols);
  ITG_LAUNCH_CHECK(ctx);
  // A = Ub S Vbo^T (or its transpose)
  double* Ua = tr ? Vbo : Ub;   // m x k, ld m
  double* Va = tr ? Ub : Vbo;   // n x k, ld n
  rc = itg_convert(ctx, (int64_t)m * k, ITG_F64, Ua, dtype, Uo);
  if (rc) return rc;
  rc = itg_convert(ctx, (int64_t)n * k, ITG_F64, Va, dtype, Vo);
  if (rc) return rc;
  ITG_HIP(ctx, hipMemcpyAsync(So, sig_sorted, std::min(m, n) * sizeof(double), hipMemcpyDeviceToDevice,
                              ctx->stream));
  *kout = k;
  if (truncerr) *truncerr = h[1];
  if (docut) *docut = h[2];
  return ITG_OK;
}
