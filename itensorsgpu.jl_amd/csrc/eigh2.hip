// Two-stage reduction of a symmetric matrix to tridiagonal form, for the
// density-matrix eigensolver at large n (the eigen step of factorize,
// src/tensor/culinearalgebra.jl:74-108, reached from replacebond!).
//
// Stage 1, dense -> band of bandwidth SB (the LAPACK dsytrd_sy2sb scheme):
//   per SB-column panel P = A[c+SB:, c:c+SB] (m rows)
//   - Householder QR of P. Panels of fewer than 512 rows are factored by one
//     workgroup with the panel in registers (sb_panel_direct_kernel). Taller
//     ones go through TSQR: leaves of 256-511 rows (sb_tsqr_leaf_kernel),
//     the stacked leaf R factors (sb_tsqr_root_kernel), then the Householder
//     reconstruction of Ballard et al. 2014, "Reconstructing Householder
//     vectors from tall-skinny QR" (sb_hr_kernel): LU without pivoting of
//     Q - S, S the sign choice that keeps every pivot >= 1 in magnitude, gives
//     V = L, T = -U S L1^{-T}, R_H = S R.
//   - two-sided update A22 <- H^T A22 H, H = I - V T V^T, as
//     X = A22 V (MFMA), W = (X - V Z) T with Z = T^T (V^T X) / 2,
//     A22 -= V W^T + W V^T (one MFMA product of inner dimension 2 SB).
// Stage 2, band -> tridiagonal by bulge chasing (Lang 1993, with the task
//   decomposition of Haidar, Ltaief & Dongarra 2011). Sweep s annihilates
//   column s below the subdiagonal; its task 0 reflects rows s+1..s+SB from
//   both sides, task k >= 1 takes the SB x SB block below the previous task's
//   rows, applies the previous reflector from the right (creating the
//   bulge), annihilates the bulge's first column with a new reflector applied
//   from the left, and applies that one from both sides to the next diagonal
//   block. Sweep s+1's task k may run once sweep s has finished task k+1. One
//   persistent kernel: each workgroup chases SB2_W consecutive sweeps in
//   lockstep (one wave per sweep, sweep i two tasks behind sweep i-1), and
//   waits on the previous block's last sweep through a progress counter.
// Back-transformation of the kept eigenvectors U (the caller then applies
//   the stage-1 reflectors as compact-WY groups): the stage-2 reflectors of
//   G2 consecutive sweeps with the same task index k form one block
//   reflector W(g,k) = H(s0,k) ... H(s0+G2-1,k); Q2 = prod_g W(g,Kmax)...W(g,0),
//   so U <- Q2 U applies, for g from last to first, W(g,0) first. Blocks
//   (g,k) on one anti-diagonal (ngroups-1-g) + k touch disjoint rows: one
//   launch per anti-diagonal.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "eigh2.h"

namespace eig2 {
namespace {

constexpr int LDB = 2 * SB;     // band storage rows per column: band and bulge
constexpr int LEAF = 256;       // TSQR leaf rows (the last leaf takes the rest)
constexpr int DIRECT_MAX = 512; // panels with fewer rows: one workgroup
constexpr int ROOT_MAX = 1024;  // stacked R rows the root factorisation holds
constexpr int SB2_W = 4;        // sweeps (waves) per bulge-chasing workgroup
constexpr int G2 = 32;          // sweeps per back-transform block
constexpr int SPIN_LIMIT = 1 << 22;

inline size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// --------------------------------------------------------------- helpers
// DPP lane exchange of a double within a row of 16 lanes
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// sum over each row of 16 lanes, every lane of the row getting the same
// bits: pairwise exchanges only (quad_perm [1,0,3,2] and [2,3,0,1], then
// row_half_mirror and row_mirror swap whole quads and half-rows), so each
// partner pair adds the same two values
__device__ __forceinline__ double row16_sum(double v) {
  v += dpp_d<0xB1>(v);
  v += dpp_d<0x4E>(v);
  v += dpp_d<0x141>(v);
  v += dpp_d<0x140>(v);
  return v;
}

// sum over aligned groups of QC lanes (QC = 16 or 32)
template <int QC>
__device__ __forceinline__ double redq(double v) {
  v = row16_sum(v);
  if (QC == 32) v += __shfl_xor(v, 16, 64);
  return v;
}

__device__ __forceinline__ double wave_sum(double v) {
  v = row16_sum(v);
  v += __shfl_xor(v, 16, 64);
  v += __shfl_xor(v, 32, 64);
  return v;
}

// LDS written by some lanes of a wave, then read by others of the same wave
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// tasks of sweep s (s <= n - 3): k with s + 1 + k SB <= n - 1
__host__ __device__ __forceinline__ int sb_nt(int n, int s) { return (n - 2 - s) / SB + 1; }

// sum_{u=0}^{U} floor(u / SB)
__host__ __device__ __forceinline__ long long sb_flsum(long long U) {
  if (U < 0) return 0;
  const long long q = (U + 1) / SB, r = (U + 1) % SB;
  return (long long)SB * q * (q - 1) / 2 + q * r;
}

// index of task (s, 0) in the reflector store: sum_{s' < s} sb_nt(n, s')
__host__ __device__ __forceinline__ long long sb_off(int n, int s) {
  return (long long)s + sb_flsum(n - 2) - sb_flsum((long long)n - 2 - s);
}

// ------------------------------------------------- panel QR (block level)
// Householder QR of the r x SB matrix held as reg[i] = P(q*32 + i, c) by the
// block's SB*QC threads (thread t: column c = t / QC, row chunk q = t % QC),
// r <= 32*QC. On return reg holds the LAPACK geqrf layout (R on and above
// the diagonal, reflectors below it, unit diagonal implicit) and tau_sh the
// scalar factors (zero for columns without a reflector). One barrier per
// column: the column's owner threads form the reflector (dlarfg) and publish
// it in LDS (two buffers, alternating), then every later column applies it.
// reflector chunk q of 32 rows lives at v[q * VQ + i] (padded: chunks of
// the same column are read together by the QC threads of that column)
constexpr int VQ = 33;

template <int QC>
__device__ void qr32(double (&reg)[32], int r, double* vsh, double* tau_sh) {
  const int t = threadIdx.x, c = t / QC, q = t % QC;
  const int nref = min(r, SB);
  for (int j = 0; j < nref; ++j) {
    double* v = vsh + (j & 1) * (VQ * QC);
    if (c == j) {
      // four partial sums: dependent FMA chains of 8, not 32
      double x4[4] = {0.0, 0.0, 0.0, 0.0}, al = 0.0;
#pragma unroll
      for (int i = 0; i < 32; ++i) {
        const int row = q * 32 + i;
        if (row > j) x4[i & 3] += reg[i] * reg[i];
        if (row == j) al = reg[i];
      }
      double xn2 = redq<QC>((x4[0] + x4[1]) + (x4[2] + x4[3]));
      al = redq<QC>(al);
      double tau = 0.0, scal = 0.0, beta = al;
      if (xn2 > 0.0) {
        beta = -copysign(sqrt(al * al + xn2), al);
        tau = (beta - al) / beta;
        scal = 1.0 / (al - beta);
      }
#pragma unroll
      for (int i = 0; i < 32; ++i) {
        const int row = q * 32 + i;
        double vi = 0.0;
        if (row == j) {
          reg[i] = beta;
          vi = 1.0;
        } else if (row > j) {
          reg[i] *= scal;
          vi = reg[i];
        }
        v[q * VQ + i] = vi;
      }
      if (q == 0) tau_sh[j] = tau;
    }
    __syncthreads();
    if (c > j) {
      double vr[32];
      double w4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int i = 0; i < 32; ++i) {
        vr[i] = v[q * VQ + i];
        w4[i & 3] += vr[i] * reg[i];
      }
      const double w = redq<QC>((w4[0] + w4[1]) + (w4[2] + w4[3])) * tau_sh[j];
#pragma unroll
      for (int i = 0; i < 32; ++i) reg[i] -= w * vr[i];
    }
  }
  if (t >= nref && t < SB) tau_sh[t] = 0.0;
  __syncthreads();
}

// explicit thin Q (r x SB) = H_0 ... H_{nref-1} [I; 0] from qr32's output,
// accumulated backwards: qr[i] = Q(q*32 + i, c). With FROM_REG the owner
// threads take the reflectors from their qr32 registers (reg); otherwise
// they read them back from Vg (geqrf layout, rows from row0, ld ldv), where
// the caller stored qr32's result, so that no registers hold them alongside
// qr (the 1024-thread root).
template <int QC, bool FROM_REG>
__device__ void form_q(const double (&reg)[32], double (&qr)[32], int r, double* vsh,
                       const double* tau_sh, const double* __restrict__ Vg, int row0, int ldv) {
  const int t = threadIdx.x, c = t / QC, q = t % QC;
  const int nref = min(r, SB);
#pragma unroll
  for (int i = 0; i < 32; ++i) qr[i] = (q * 32 + i == c) ? 1.0 : 0.0;
  for (int j = nref - 1; j >= 0; --j) {
    double* v = vsh + (j & 1) * (VQ * QC);
    if (c == j) {
#pragma unroll
      for (int i = 0; i < 32; ++i) {
        const int row = q * 32 + i;
        double vi = 0.0;
        if (row == j) vi = 1.0;
        else if (row > j && row < r) vi = FROM_REG ? reg[i] : Vg[row0 + row + (size_t)c * ldv];
        v[q * VQ + i] = vi;
      }
    }
    __syncthreads();
    if (c >= j) {
      double vr[32];
      double w4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int i = 0; i < 32; ++i) {
        vr[i] = v[q * VQ + i];
        w4[i & 3] += vr[i] * qr[i];
      }
      const double w = redq<QC>((w4[0] + w4[1]) + (w4[2] + w4[3])) * tau_sh[j];
#pragma unroll
      for (int i = 0; i < 32; ++i) qr[i] -= w * vr[i];
    }
  }
  __syncthreads();
}

// ------------------------------------------------------- panel kernels
// Panel of m < DIRECT_MAX rows (ld lda), factored in one block: geqrf layout
// back into P, the explicit V (unit diagonal, zeros above) into VW columns
// [0, SB) and [2 SB, 3 SB) (ld ldv), and T (dlarft forward columnwise,
// SB x SB, zeros below the diagonal) into Tout.
__global__ __launch_bounds__(512) void sb_panel_direct_kernel(double* __restrict__ P, int lda,
                                                             int m, double* __restrict__ VW,
                                                             int ldv, double* __restrict__ Tout) {
  constexpr int QC = 16;
  __shared__ double vsh[2 * VQ * QC];
  __shared__ double tau_sh[SB];
  __shared__ double Vs[64][SB + 1];
  __shared__ double Gs[SB][SB + 1];
  __shared__ double Ts[SB][SB + 1];
  __shared__ double tmp[SB];
  const int t = threadIdx.x, c = t / QC, q = t % QC;
  double reg[32];
#pragma unroll
  for (int i = 0; i < 32; ++i) {
    const int row = q * 32 + i;
    reg[i] = row < m ? P[row + (size_t)c * lda] : 0.0;
  }
  qr32<QC>(reg, m, vsh, tau_sh);
#pragma unroll
  for (int i = 0; i < 32; ++i) {
    const int row = q * 32 + i;
    if (row < m) {
      P[row + (size_t)c * lda] = reg[i];
      const double vx = row < c ? 0.0 : (row == c ? 1.0 : reg[i]);
      VW[row + (size_t)c * ldv] = vx;
      VW[row + (size_t)(2 * SB + c) * ldv] = vx;
    }
  }
  // Gram matrix V^T V through 64-row chunks of LDS
  double g0 = 0.0, g1 = 0.0;   // entries t and t + 512 (a = e % SB, b = e / SB)
  for (int r0 = 0; r0 < m; r0 += 64) {
    __syncthreads();
    if (q * 32 >= r0 && q * 32 < r0 + 64) {
#pragma unroll
      for (int i = 0; i < 32; ++i) {
        const int row = q * 32 + i;
        Vs[row - r0][c] = row < m ? (row < c ? 0.0 : (row == c ? 1.0 : reg[i])) : 0.0;
      }
    }
    __syncthreads();
    {
      int a = t % SB, b = t / SB;
      double s = 0.0;
      for (int rr = 0; rr < 64; ++rr) s += Vs[rr][a] * Vs[rr][b];
      g0 += s;
      a = (t + 512) % SB;
      b = (t + 512) / SB;
      s = 0.0;
      for (int rr = 0; rr < 64; ++rr) s += Vs[rr][a] * Vs[rr][b];
      g1 += s;
    }
  }
  Gs[t % SB][t / SB] = g0;
  Gs[(t + 512) % SB][(t + 512) / SB] = g1;
  Ts[t % SB][t / SB] = 0.0;
  Ts[(t + 512) % SB][(t + 512) / SB] = 0.0;
  __syncthreads();
  for (int i = 0; i < SB; ++i) {
    if (t < i) tmp[t] = -tau_sh[i] * Gs[t][i];
    __syncthreads();
    if (t < i) {
      double acc = 0.0;
      for (int l = t; l < i; ++l) acc += Ts[t][l] * tmp[l];
      Ts[t][i] = acc;
    }
    if (t == i) Ts[i][i] = tau_sh[i];
    __syncthreads();
  }
  Tout[t] = Ts[t % SB][t / SB];
  Tout[t + 512] = Ts[(t + 512) % SB][(t + 512) / SB];
}

// TSQR leaves: block i factors rows [i LEAF, i LEAF + r) of the panel
// (r = LEAF, the last block the rest, < 2 LEAF), writes its explicit Q into
// Qleaf (same rows, ld ldq) and its R into Rstack rows [SB i, SB i + SB)
// (ld ROOT_MAX).
__global__ __launch_bounds__(512) void sb_tsqr_leaf_kernel(const double* __restrict__ P, int lda,
                                                          int m, int nleaf,
                                                          double* __restrict__ Qleaf, int ldq,
                                                          double* __restrict__ Rstack) {
  constexpr int QC = 16;
  __shared__ double vsh[2 * VQ * QC];
  __shared__ double tau_sh[SB];
  const int t = threadIdx.x, c = t / QC, q = t % QC;
  const int row0 = blockIdx.x * LEAF;
  const int r = blockIdx.x == nleaf - 1 ? m - row0 : LEAF;
  double reg[32], qreg[32];
#pragma unroll
  for (int i = 0; i < 32; ++i) {
    const int row = q * 32 + i;
    reg[i] = row < r ? P[row0 + row + (size_t)c * lda] : 0.0;
  }
  qr32<QC>(reg, r, vsh, tau_sh);
  if (q == 0) {
#pragma unroll
    for (int i = 0; i < 32; ++i)
      Rstack[(SB * blockIdx.x + i) + (size_t)c * ROOT_MAX] = i <= c ? reg[i] : 0.0;
  }
  form_q<QC, true>(reg, qreg, r, vsh, tau_sh, P, row0, lda);
#pragma unroll
  for (int i = 0; i < 32; ++i) {
    const int row = q * 32 + i;
    if (row < r) Qleaf[row0 + row + (size_t)c * ldq] = qreg[i];
  }
}

// QR of the rs x SB stacked R factors (rs <= 32 QC, factored in place):
// explicit Q into Qtop (ld ROOT_MAX), R into Rsm (SB x SB).
template <int QC>
__global__ __launch_bounds__(SB * QC) void sb_tsqr_root_kernel(double* __restrict__ Rstack,
                                                              int rs, double* __restrict__ Qtop,
                                                              double* __restrict__ Rsm) {
  __shared__ double vsh[2 * VQ * QC];
  __shared__ double tau_sh[SB];
  const int t = threadIdx.x, c = t / QC, q = t % QC;
  double reg[32], qreg[32];
#pragma unroll
  for (int i = 0; i < 32; ++i) {
    const int row = q * 32 + i;
    reg[i] = row < rs ? Rstack[row + (size_t)c * ROOT_MAX] : 0.0;
  }
  qr32<QC>(reg, rs, vsh, tau_sh);
  if (q == 0) {
#pragma unroll
    for (int i = 0; i < 32; ++i) Rsm[i + SB * c] = i <= c ? reg[i] : 0.0;
  }
  if (QC > 16) {
    // the reflectors go back in place and are re-read by form_q
#pragma unroll
    for (int i = 0; i < 32; ++i) {
      const int row = q * 32 + i;
      if (row < rs) Rstack[row + (size_t)c * ROOT_MAX] = reg[i];
    }
  }
  form_q<QC, (QC <= 16)>(reg, qreg, rs, vsh, tau_sh, Rstack, 0, ROOT_MAX);
#pragma unroll
  for (int i = 0; i < 32; ++i) {
    const int row = q * 32 + i;
    if (row < rs) Qtop[row + (size_t)c * ROOT_MAX] = qreg[i];
  }
}

// Householder reconstruction. The panel's thin Q is Q_i Qtop_i on leaf i's
// rows. Every block forms the top SB x SB of Q (leaf 0's first rows times
// Qtop's first SB rows) and its LU with the sign choice, then writes its
// leaf's rows of V = (Q - [S; 0]) U^{-1} (rows >= SB: Qleaf row times
// M_i = Qtop_i U^{-1}; leaf 0's first SB rows: L1) in geqrf layout into the
// panel (with R_H = S R above the diagonal) and explicitly into VW. Block 0
// also writes T = -U S L1^{-T}.
__global__ __launch_bounds__(512) void sb_hr_kernel(double* __restrict__ P, int lda, int m,
                                                   int nleaf, const double* __restrict__ Qleaf,
                                                   int ldq, const double* __restrict__ Qtop,
                                                   const double* __restrict__ Rsm,
                                                   double* __restrict__ VW, int ldv,
                                                   double* __restrict__ Tout) {
  __shared__ double F[SB][SB + 1];
  __shared__ double Qt0[SB][SB + 1];
  __shared__ double Qti[SB][SB + 1];
  __shared__ double Mi[SB][SB + 1];
  __shared__ double Rs[SB][SB + 1];
  __shared__ double Tw[SB][SB + 1];
  __shared__ double sgn[SB];
  const int t = threadIdx.x;
  const int blk = blockIdx.x;
  const int row0 = blk * LEAF;
  const int r = blk == nleaf - 1 ? m - row0 : LEAF;
  for (int e = t; e < SB * SB; e += 512) {
    const int a = e % SB, b = e / SB;
    Qt0[a][b] = Qtop[a + (size_t)b * ROOT_MAX];
    Qti[a][b] = Qtop[(SB * blk + a) + (size_t)b * ROOT_MAX];
    Mi[a][b] = Qleaf[a + (size_t)b * ldq];   // leaf 0's first rows, staged
    Rs[a][b] = Rsm[a + SB * b];
  }
  __syncthreads();
  for (int e = t; e < SB * SB; e += 512) {
    const int a = e % SB, b = e / SB;
    double s = 0.0;
    for (int l = 0; l < SB; ++l) s += Mi[a][l] * Qt0[l][b];
    F[a][b] = s;
  }
  __syncthreads();
  // LU of F - S without pivoting, S_jj = -sign(current pivot)
  for (int j = 0; j < SB; ++j) {
    if (t == 0) {
      const double s = F[j][j] >= 0.0 ? -1.0 : 1.0;
      sgn[j] = s;
      F[j][j] -= s;
    }
    __syncthreads();
    if (t > j && t < SB) F[t][j] /= F[j][j];
    __syncthreads();
    const int w = SB - 1 - j;
    for (int e = t; e < w * w; e += 512) {
      const int a = j + 1 + e % w, b = j + 1 + e / w;
      F[a][b] -= F[a][j] * F[j][b];
    }
    __syncthreads();
  }
  // M = Qtop_i U^{-1}: row a solves x U = Qtop_i(a, :)
  if (t < SB) {
    for (int b = 0; b < SB; ++b) {
      double x = Qti[t][b];
      for (int l = 0; l < b; ++l) x -= Mi[t][l] * F[l][b];
      Mi[t][b] = x / F[b][b];
    }
  }
  __syncthreads();
  if (t < r) {
    const int row = row0 + t;
    if (row < SB) {
      for (int cc = 0; cc < SB; ++cc) {
        const double vx = cc < row ? F[row][cc] : (cc == row ? 1.0 : 0.0);
        P[row + (size_t)cc * lda] = cc < row ? F[row][cc] : sgn[row] * Rs[row][cc];
        VW[row + (size_t)cc * ldv] = vx;
        VW[row + (size_t)(2 * SB + cc) * ldv] = vx;
      }
    } else {
      double qrow[SB];
#pragma unroll
      for (int l = 0; l < SB; ++l) qrow[l] = Qleaf[row + (size_t)l * ldq];
#pragma unroll 4
      for (int cc = 0; cc < SB; ++cc) {
        double s = 0.0;
#pragma unroll
        for (int l = 0; l < SB; ++l) s += qrow[l] * Mi[l][cc];
        P[row + (size_t)cc * lda] = s;
        VW[row + (size_t)cc * ldv] = s;
        VW[row + (size_t)(2 * SB + cc) * ldv] = s;
      }
    }
  }
  if (blk == 0) {
    // T L1^T = -U S, row by row: forward substitution with the unit lower L1
    if (t < SB) {
      for (int b = 0; b < SB; ++b) {
        double z = b >= t ? -F[t][b] * sgn[b] : 0.0;
        for (int l = 0; l < b; ++l) z -= F[b][l] * Tw[t][l];
        Tw[t][b] = z;
      }
    }
    __syncthreads();
    for (int e = t; e < SB * SB; e += 512) {
      const int a = e % SB, b = e / SB;
      Tout[e] = b >= a ? Tw[a][b] : 0.0;
    }
  }
}

// partial V^T X over 256-row chunks (rows [256 p, 256 p + 256) of m)
__global__ __launch_bounds__(256) void sb_m0part_kernel(const double* __restrict__ X, int ldx,
                                                       const double* __restrict__ VW, int ldv,
                                                       int m, double* __restrict__ Mpart) {
  __shared__ double Vs[64][SB + 1];
  __shared__ double Xs[64][SB + 1];
  const int t = threadIdx.x;
  const int r0 = blockIdx.x * 256;
  const int rows = min(256, m - r0);
  double g[4] = {0.0, 0.0, 0.0, 0.0};
  for (int ch = 0; ch < rows; ch += 64) {
    __syncthreads();
    for (int e = t; e < 64 * SB; e += 256) {
      const int rr = e % 64, cc = e / 64;
      const bool ok = ch + rr < rows;
      Vs[rr][cc] = ok ? VW[(r0 + ch + rr) + (size_t)cc * ldv] : 0.0;
      Xs[rr][cc] = ok ? X[(r0 + ch + rr) + (size_t)cc * ldx] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = t + 256 * u, a = e % SB, b = e / SB;
      double s = 0.0;
      for (int rr = 0; rr < 64; ++rr) s += Vs[rr][a] * Xs[rr][b];
      g[u] += s;
    }
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) Mpart[(size_t)blockIdx.x * SB * SB + t + 256 * u] = g[u];
}

// W = (X - V Z) T, Z = T^T (V^T X) / 2, for this block's 256 rows, into VW
// columns [SB, 2 SB)
__global__ __launch_bounds__(256) void sb_wfinal_kernel(const double* __restrict__ X, int ldx,
                                                       double* __restrict__ VW, int ldv, int m,
                                                       const double* __restrict__ Mpart,
                                                       int nparts, const double* __restrict__ T) {
  __shared__ double Ms[SB][SB + 1];
  __shared__ double Ts[SB][SB + 1];
  __shared__ double Zs[SB][SB + 1];
  const int t = threadIdx.x;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int e = t + 256 * u;
    double s = 0.0;
    for (int p = 0; p < nparts; ++p) s += Mpart[(size_t)p * SB * SB + e];
    Ms[e % SB][e / SB] = s;
    Ts[e % SB][e / SB] = T[e];
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int e = t + 256 * u, a = e % SB, b = e / SB;
    double s = 0.0;
    for (int l = 0; l <= a; ++l) s += Ts[l][a] * Ms[l][b];
    Zs[a][b] = 0.5 * s;
  }
  __syncthreads();
  const int row = blockIdx.x * 256 + t;
  if (row >= m) return;
  double xu[SB], vx[SB];
#pragma unroll
  for (int c = 0; c < SB; ++c) {
    xu[c] = X[row + (size_t)c * ldx];
    vx[c] = VW[row + (size_t)c * ldv];
  }
#pragma unroll
  for (int c = 0; c < SB; ++c) {
    double s = xu[c];
#pragma unroll
    for (int a = 0; a < SB; ++a) s -= vx[a] * Zs[a][c];
    xu[c] = s;
  }
  // w(c) = sum_{a <= c} u(a) T(a, c), c descending so u(a < c) is intact
#pragma unroll
  for (int c = SB - 1; c >= 0; --c) {
    double s = 0.0;
#pragma unroll
    for (int a = 0; a <= c; ++a) s += xu[a] * Ts[a][c];
    xu[c] = s;
  }
#pragma unroll
  for (int c = 0; c < SB; ++c) VW[row + (size_t)(SB + c) * ldv] = xu[c];
}

// band of A (lower: A(j + r, j), r <= SB) into AB, bulge rows zero
__global__ void sb_band_kernel(const double* __restrict__ A, int n, double* __restrict__ AB) {
  const int j = blockIdx.x, r = threadIdx.x;
  AB[r + (size_t)LDB * j] = (r <= SB && j + r < n) ? A[(j + r) + (size_t)j * n] : 0.0;
}

// reflector j of the band reduction (j < n - SB) as column j of A: zeros
// above row j + SB, one at j + SB, the stored vector below
__global__ void sb_v1_kernel(double* __restrict__ A, int n) {
  const int64_t tot = (int64_t)n * (n - SB);
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(e % n), j = (int)(e / n);
    if (r < j + SB) A[e] = 0.0;
    else if (r == j + SB) A[e] = 1.0;
  }
}

// the last panel's T (fewer than SB reflectors) from leading dimension SB
// to leading dimension nb, as the back-transform reads it (in place; the
// destination never passes the source)
__global__ void sb_t_repack_kernel(double* __restrict__ T, int nb) {
  for (int b = 0; b < nb; ++b)
    for (int a = 0; a < nb; ++a) T[a + nb * b] = T[a + SB * b];
}

__global__ void sb_flags_kernel(int nsw, int* __restrict__ prog, int* __restrict__ err) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nsw; i += gridDim.x * blockDim.x)
    prog[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) *err = 0;
}

// ------------------------------------------------------------ bulge chasing
// One wave's reflector from x (lane l holds x(l & 31), rows < lc):
// (I - tau v v^T) x = beta e_0 (dlarfg); returns v(l & 31), sets tau, beta.
__device__ __forceinline__ double lane0(double v) {
  return __hiloint2double(__builtin_amdgcn_readfirstlane(__double2hiint(v)),
                          __builtin_amdgcn_readfirstlane(__double2loint(v)));
}

// (valid on the lanes of the first half-wave, which hold x(0..31))
__device__ __forceinline__ double wave_house(double x, int lc, double& tau, double& beta) {
  const int lane = threadIdx.x & 63, li = lane & 31, lh = lane >> 5;
  const double al = lane0(x);
  const double xn2 = wave_sum((lh == 0 && li >= 1 && li < lc) ? x * x : 0.0);
  tau = 0.0;
  beta = al;
  double scal = 0.0;
  if (xn2 > 0.0) {
    beta = -copysign(sqrt(al * al + xn2), al);
    tau = (beta - al) / beta;
    scal = 1.0 / (al - beta);
  }
  return li == 0 ? 1.0 : (li < lc ? x * scal : 0.0);
}

// D (rows/cols [c0, c0 + lc)) <- H D H, H = I - tau v v^T; lane (li, lh)
// holds dreg[q] = D(li, 2q + lh); vv = v in LDS, vli = v(li)
__device__ __forceinline__ void wave_two_sided(double (&dreg)[16], double tau, double vli,
                                               const double* vv, double* ww) {
  const int lane = threadIdx.x & 63, li = lane & 31, lh = lane >> 5;
  double p = 0.0;
#pragma unroll
  for (int q = 0; q < 16; ++q) p += dreg[q] * vv[2 * q + lh];
  p += __shfl_xor(p, 32, 64);
  p *= tau;
  const double K = 0.5 * tau * wave_sum(lh == 0 ? p * vli : 0.0);
  const double wli = p - K * vli;
  if (lh == 0) ww[li] = wli;
  wave_sync();
#pragma unroll
  for (int q = 0; q < 16; ++q) dreg[q] -= vli * ww[2 * q + lh] + wli * vv[2 * q + lh];
}

template <int W>
__global__ __launch_bounds__(64 * W) void sb2st_kernel(double* __restrict__ AB, int n, int nsw,
                                                      double* __restrict__ V2,
                                                      double* __restrict__ tau2,
                                                      int* __restrict__ prog,
                                                      int* __restrict__ err) {
  __shared__ double Bs[W][SB][SB + 1];   // [wave][column][row]
  __shared__ double vv_s[W][SB], ww_s[W][SB], zz_s[W][SB], vp_s[W][SB];
  __shared__ int stop;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 31, lh = lane >> 5;
  double* vv = vv_s[wave];
  double* ww = ww_s[wave];
  double* zz = zz_s[wave];
  double* vp = vp_s[wave];
  if (threadIdx.x == 0) stop = 0;
  __syncthreads();
  for (int blk = blockIdx.x; blk * W < nsw; blk += gridDim.x) {
    const int s = blk * W + wave;
    const int nt = s < nsw ? sb_nt(n, s) : 0;
    const int steps = 2 * (W - 1) + sb_nt(n, blk * W);
    const int ntp = blk > 0 ? sb_nt(n, blk * W - 1) : 0;
    const long long off = s < nsw ? sb_off(n, s) : 0;
    const int sl = blk * W + W - 1;   // the block's last sweep, published for the next block
    const int ntl = sl < nsw ? sb_nt(n, sl) : 0;
    double taup = 0.0;
    int seen = 0;
    for (int t = 0; t <= steps; ++t) {
      // wave 0's task at step t needs the previous block's last sweep to
      // have finished its task t + 1 (or all of its tasks)
      if (wave == 0 && blk > 0 && t < steps && t < nt) {
        const int need = min(t + 2, ntp);
        int it = 0;
        while (seen < need) {
          seen = __hip_atomic_load(prog + (s - 1), __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
          if (seen >= need) break;
          if (++it > SPIN_LIMIT) {
            if (lane == 0) {
              stop = 1;
              __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      __syncthreads();
      // every wave's writes of step t - 1 are complete: publish the last
      // sweep's progress to the next block
      if (wave == W - 1 && lane == 0 && t > 0) {
        const int kl = t - 1 - 2 * (W - 1);
        if (kl >= 0 && kl < ntl)
          __hip_atomic_store(prog + sl, kl + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      }
      if (stop || t == steps) break;
      const int k = t - 2 * wave;
      if (k < 0 || k >= nt) continue;
      const int c0 = s + 1 + k * SB;
      const int lc = min(SB, n - c0);
      double tau, beta, vli;
      // the diagonal block D = A(c0.., c0..), which the first half of the
      // task does not touch: its loads go out first
      double dreg[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int j = 2 * q + lh;
        double dv = 0.0;
        if (li < lc && j < lc)
          dv = li >= j ? AB[(li - j) + (size_t)LDB * (c0 + j)] : AB[(j - li) + (size_t)LDB * (c0 + li)];
        dreg[q] = dv;
      }
      if (k == 0) {
        const double x = li < lc ? AB[(1 + li) + (size_t)LDB * s] : 0.0;
        vli = wave_house(x, lc, tau, beta);
        if (lh == 0 && li < lc) AB[(1 + li) + (size_t)LDB * s] = li == 0 ? beta : 0.0;
      } else {
        const int p0 = c0 - SB;
        double breg[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const int j = 2 * q + lh;
          breg[q] = li < lc ? AB[(SB + li - j) + (size_t)LDB * (p0 + j)] : 0.0;
        }
        // right application of the previous reflector: B <- B - taup (B vp) vp^T
        double y = 0.0;
#pragma unroll
        for (int q = 0; q < 16; ++q) y += breg[q] * vp[2 * q + lh];
        y += __shfl_xor(y, 32, 64);
#pragma unroll
        for (int q = 0; q < 16; ++q) breg[q] -= taup * y * vp[2 * q + lh];
        // new reflector from the bulge's first column, B(li, 0) = breg[0] on
        // the first half-wave
        vli = wave_house(breg[0], lc, tau, beta);
        if (lh == 0) {
          breg[0] = li == 0 ? beta : 0.0;
          vv[li] = vli;
        }
        // left application to the other columns: z = v^T B (transpose in LDS)
#pragma unroll
        for (int q = 0; q < 16; ++q) Bs[wave][2 * q + lh][li] = breg[q];
        wave_sync();
        vli = vv[li];
        {
          double zp = 0.0;
#pragma unroll
          for (int i = 0; i < 16; ++i) zp += vv[16 * lh + i] * Bs[wave][li][16 * lh + i];
          zp += __shfl_xor(zp, 32, 64);
          if (lh == 0) zz[li] = zp;
        }
        wave_sync();
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const int j = 2 * q + lh;
          if (j >= 1) breg[q] -= tau * vli * zz[j];
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const int j = 2 * q + lh;
          if (li < lc) AB[(SB + li - j) + (size_t)LDB * (p0 + j)] = breg[q];
        }
      }
      if (lh == 0) vv[li] = vli;
      wave_sync();
      // two-sided application to the diagonal block D = A(c0.., c0..)
      wave_two_sided(dreg, tau, vli, vv, ww);
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int j = 2 * q + lh;
        if (li < lc && j <= li) AB[(li - j) + (size_t)LDB * (c0 + j)] = dreg[q];
      }
      if (lh == 0) V2[(size_t)(off + k) * SB + li] = vli;
      if (lane == 0) tau2[off + k] = tau;
      // this reflector is the next task's right factor
      if (lh == 0) vp[li] = vli;
      taup = tau;
      wave_sync();
    }
    if (stop) break;   // a wait ran out: leave every remaining block alone
  }
}

__global__ void sb_de_kernel(const double* __restrict__ AB, int n, double* __restrict__ d,
                             double* __restrict__ e) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    d[j] = AB[(size_t)LDB * j];
    if (j < n - 1) e[j] = AB[1 + (size_t)LDB * j];
  }
}

// ------------------------------------------------------ back-transform (Q2)
// sweeps of group g with task k: a prefix of the group (task counts fall
// with s)
__device__ __forceinline__ int q2_width(int n, int nsw, int s0, int k) {
  int gb = 0;
  while (gb < G2 && s0 + gb < nsw && sb_nt(n, s0 + gb) > k) ++gb;
  return gb;
}

// V of block (g, k) into LDS: column jj (sweep s0 + jj) holds its reflector
// from row jj on (rows relative to top = s0 + 1 + k SB)
__device__ __forceinline__ void q2_load_v(int n, int s0, int k, int gb, int nr,
                                          const double* __restrict__ V2,
                                          double (*Vs)[G2 + 1]) {
  for (int e = threadIdx.x; e < (G2 + SB) * G2; e += blockDim.x) {
    const int r = e % (G2 + SB), jj = e / (G2 + SB);
    double v = 0.0;
    if (jj < gb && r < nr && r >= jj && r - jj < SB)
      v = V2[(size_t)(sb_off(n, s0 + jj) + k) * SB + (r - jj)];
    Vs[r][jj] = v;
  }
}

// compact-WY factor of every block (dlarft forward columnwise)
__global__ __launch_bounds__(256) void q2_t_kernel(int n, int nsw, int kmax,
                                                  const double* __restrict__ V2,
                                                  const double* __restrict__ tau2,
                                                  double* __restrict__ Tq2) {
  __shared__ double Vs[G2 + SB][G2 + 1];
  __shared__ double Gs[G2][G2 + 1];
  __shared__ double Ts[G2][G2 + 1];
  __shared__ double ts[G2];
  __shared__ double tmp[G2];
  const int g = blockIdx.x / kmax, k = blockIdx.x % kmax;
  const int s0 = g * G2;
  if (s0 >= nsw || k >= sb_nt(n, s0)) return;
  const int gb = q2_width(n, nsw, s0, k);
  const int top = s0 + 1 + k * SB;
  const int nr = min(gb - 1 + SB, n - top);
  const int t = threadIdx.x;
  q2_load_v(n, s0, k, gb, nr, V2, Vs);
  if (t < G2) ts[t] = t < gb ? tau2[sb_off(n, s0 + t) + k] : 0.0;
  __syncthreads();
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int e = t + 256 * u, a = e % G2, b = e / G2;
    double s = 0.0;
    for (int r = 0; r < G2 + SB; ++r) s += Vs[r][a] * Vs[r][b];
    Gs[a][b] = s;
    Ts[a][b] = 0.0;
  }
  __syncthreads();
  for (int i = 0; i < G2; ++i) {
    if (t < i) tmp[t] = -ts[i] * Gs[t][i];
    __syncthreads();
    if (t < i) {
      double acc = 0.0;
      for (int l = t; l < i; ++l) acc += Ts[t][l] * tmp[l];
      Ts[t][i] = acc;
    }
    if (t == i) Ts[i][i] = ts[i];
    __syncthreads();
  }
  double* T = Tq2 + ((size_t)g * kmax + k) * G2 * G2;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int e = t + 256 * u;
    T[e] = Ts[e % G2][e / G2];
  }
}

// U(rows of block (g, k), this block's 64 columns) <- W(g,k) U for the
// blocks on anti-diagonal `step`, as three small products on FP64 MFMA
// (v_mfma_f64_16x16x4: lane (lr, lg) supplies A(lr, lg) and B(lg, lr) of
// each k-step and holds D(4 r + lg, lr), the layout of the next product's B
// operand): X = V^T E (32 x 64 x 16 per wave), Y = T X, E -= V Y. Wave w
// takes columns 16 w .. 16 w + 15 of the block's 64; its E fragments are
// loaded once and updated in registers. Only the k-steps where V or T can
// be nonzero are issued (V(r, jj) != 0 for 0 <= r - jj < SB; T upper).
typedef double q2d4_t __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void q2_apply_kernel(int n, int nsw, int ngroups, int kmax,
                                                      int step, int kmin,
                                                      const double* __restrict__ V2,
                                                      const double* __restrict__ Tq2,
                                                      double* __restrict__ U, int ldu, int kc) {
  __shared__ double Vs[G2 + SB][G2 + 1];
  __shared__ double Ts[G2][G2 + 1];
  const int k = kmin + blockIdx.x;
  const int g = ngroups - 1 - step + k;
  const int s0 = g * G2;
  if (g < 0 || s0 >= nsw || k >= kmax || k >= sb_nt(n, s0)) return;
  const int gb = q2_width(n, nsw, s0, k);
  const int top = s0 + 1 + k * SB;
  const int nr = min(gb - 1 + SB, n - top);
  q2_load_v(n, s0, k, gb, nr, V2, Vs);
  const double* T = Tq2 + ((size_t)g * kmax + k) * G2 * G2;
  for (int e = threadIdx.x; e < G2 * G2; e += 256) Ts[e % G2][e / G2] = T[e];
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lr = lane & 15, lg = lane >> 4;
  const int col = blockIdx.y * 64 + wave * 16 + lr;
  const bool colok = col < kc;
  double* Uc = U + (size_t)(colok ? col : 0) * ldu + top;
  // E(4 kk + lg, col), kk = 0..15 (rows 0..63 of the block)
  double ef[16];
#pragma unroll
  for (int kk = 0; kk < 16; ++kk) {
    const int r = 4 * kk + lg;
    ef[kk] = (colok && r < nr) ? Uc[r] : 0.0;
  }
  // X = V^T E: reflectors 0..15 use rows 0..47, reflectors 16..31 rows 16..63
  q2d4_t x0 = {0.0, 0.0, 0.0, 0.0}, x1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int kk = 0; kk < 12; ++kk)
    x0 = __builtin_amdgcn_mfma_f64_16x16x4f64(Vs[4 * kk + lg][lr], ef[kk], x0, 0, 0, 0);
#pragma unroll
  for (int kk = 4; kk < 16; ++kk)
    x1 = __builtin_amdgcn_mfma_f64_16x16x4f64(Vs[4 * kk + lg][16 + lr], ef[kk], x1, 0, 0, 0);
  // Y = T X (T upper triangular: rows 16..31 need columns 16..31 only)
  q2d4_t y0 = {0.0, 0.0, 0.0, 0.0}, y1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int kk = 0; kk < 4; ++kk)
    y0 = __builtin_amdgcn_mfma_f64_16x16x4f64(Ts[lr][4 * kk + lg], x0[kk], y0, 0, 0, 0);
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
    y0 = __builtin_amdgcn_mfma_f64_16x16x4f64(Ts[lr][16 + 4 * kk + lg], x1[kk], y0, 0, 0, 0);
    y1 = __builtin_amdgcn_mfma_f64_16x16x4f64(Ts[16 + lr][16 + 4 * kk + lg], x1[kk], y1, 0, 0, 0);
  }
  // D = V Y, 16-row tiles mt: rows 0..15 meet reflectors 0..15, rows 48..63
  // reflectors 16..31, the middle tiles all 32
  q2d4_t u0 = {0.0, 0.0, 0.0, 0.0}, u1 = u0, u2 = u0, u3 = u0;
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
    u0 = __builtin_amdgcn_mfma_f64_16x16x4f64(Vs[lr][4 * kk + lg], y0[kk], u0, 0, 0, 0);
    u1 = __builtin_amdgcn_mfma_f64_16x16x4f64(Vs[16 + lr][4 * kk + lg], y0[kk], u1, 0, 0, 0);
    u2 = __builtin_amdgcn_mfma_f64_16x16x4f64(Vs[32 + lr][4 * kk + lg], y0[kk], u2, 0, 0, 0);
  }
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
    u1 = __builtin_amdgcn_mfma_f64_16x16x4f64(Vs[16 + lr][16 + 4 * kk + lg], y1[kk], u1, 0, 0, 0);
    u2 = __builtin_amdgcn_mfma_f64_16x16x4f64(Vs[32 + lr][16 + 4 * kk + lg], y1[kk], u2, 0, 0, 0);
    u3 = __builtin_amdgcn_mfma_f64_16x16x4f64(Vs[48 + lr][16 + 4 * kk + lg], y1[kk], u3, 0, 0, 0);
  }
  // E(16 mt + 4 r + lg) -= D, held as ef[4 mt + r]
  if (!colok) return;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int b0 = 4 * r + lg;
    if (b0 < nr) Uc[b0] = ef[r] - u0[r];
    if (16 + b0 < nr) Uc[16 + b0] = ef[4 + r] - u1[r];
    if (32 + b0 < nr) Uc[32 + b0] = ef[8 + r] - u2[r];
    if (48 + b0 < nr) Uc[48 + b0] = ef[12 + r] - u3[r];
  }
}

}  // namespace

size_t doubles_needed(int n) {
  Bufs b;
  return (carve(n, nullptr, &b) + 7) / 8 + 32;
}

size_t carve(int n, void* base, Bufs* b) {
  const int nsw = std::max(n - 2, 0);
  const long long ntask = nsw > 0 ? sb_off(n, nsw) : 0;
  const int ngroups = (nsw + G2 - 1) / G2;
  const int kmax = nsw > 0 ? sb_nt(n, 0) : 0;
  const int npanel = (n + SB - 1) / SB;
  char* p = (char*)base;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    void* r = p + off;
    off += al256(bytes);
    return r;
  };
  b->AB = (double*)take(((size_t)LDB * n + 64) * 8);   // + trash slots
  b->V2 = (double*)take((size_t)ntask * SB * 8);
  b->tau2 = (double*)take((size_t)ntask * 8);
  b->Tq2 = (double*)take((size_t)ngroups * kmax * G2 * G2 * 8);
  b->Tall = (double*)take((size_t)npanel * SB * SB * 8);
  b->VW = (double*)take(3 * (size_t)n * SB * 8);
  b->X = (double*)take((size_t)n * SB * 8);
  b->Mpart = (double*)take((size_t)((n + 255) / 256) * SB * SB * 8);
  b->Qleaf = (double*)take((size_t)n * SB * 8);
  b->Rstack = (double*)take((size_t)ROOT_MAX * SB * 8);
  b->Qtop = (double*)take((size_t)ROOT_MAX * SB * 8);
  b->Rsm = (double*)take((size_t)SB * SB * 8);
  b->prog = (int*)take((size_t)(nsw + 1) * 4);
  b->err = (int*)take(256);
  return off;
}

int reduce(itg_ctx* ctx, int n, double* A, double* d, double* e, const Bufs& b) {
  if (n < 2 * SB + 2) return itg_fail(ctx, ITG_EUNSUPPORTED, "two-stage reduction: n = %d too small", n);
  if ((n - SB) / LEAF > ROOT_MAX / SB)
    return itg_fail(ctx, ITG_EUNSUPPORTED, "two-stage reduction: n = %d too large", n);
  // 1. dense -> band
  for (int c = 0; c + SB < n; c += SB) {
    const int m = n - c - SB;
    double* P = A + (c + SB) + (size_t)c * n;
    double* T = b.Tall + (size_t)(c / SB) * SB * SB;
    if (m < DIRECT_MAX) {
      sb_panel_direct_kernel<<<1, 512, 0, ctx->stream>>>(P, n, m, b.VW, n, T);
      ITG_LAUNCH_CHECK(ctx);
    } else {
      const int nleaf = m / LEAF;
      sb_tsqr_leaf_kernel<<<nleaf, 512, 0, ctx->stream>>>(P, n, m, nleaf, b.Qleaf, n, b.Rstack);
      ITG_LAUNCH_CHECK(ctx);
      const int rs = SB * nleaf;
      if (rs <= 512)
        sb_tsqr_root_kernel<16><<<1, 512, 0, ctx->stream>>>(b.Rstack, rs, b.Qtop, b.Rsm);
      else
        sb_tsqr_root_kernel<32><<<1, 1024, 0, ctx->stream>>>(b.Rstack, rs, b.Qtop, b.Rsm);
      ITG_LAUNCH_CHECK(ctx);
      sb_hr_kernel<<<nleaf, 512, 0, ctx->stream>>>(P, n, m, nleaf, b.Qleaf, n, b.Qtop, b.Rsm,
                                                   b.VW, n, T);
      ITG_LAUNCH_CHECK(ctx);
    }
    // A22 <- H^T A22 H
    double* A22 = A + (c + SB) + (size_t)(c + SB) * n;
    int rc = itg_gemm_internal(ctx, ITG_F64, ITG_MODE_NATIVE, 'N', 'N', m, SB, m, 1.0, A22, n,
                               b.VW, n, 0.0, b.X, n);
    if (rc) return rc;
    const int nparts = (m + 255) / 256;
    sb_m0part_kernel<<<nparts, 256, 0, ctx->stream>>>(b.X, n, b.VW, n, m, b.Mpart);
    sb_wfinal_kernel<<<nparts, 256, 0, ctx->stream>>>(b.X, n, b.VW, n, m, b.Mpart, nparts, T);
    ITG_LAUNCH_CHECK(ctx);
    rc = itg_gemm_internal(ctx, ITG_F64, ITG_MODE_NATIVE, 'N', 'T', m, m, 2 * SB, -1.0, b.VW, n,
                           b.VW + (size_t)SB * n, n, 1.0, A22, n);
    if (rc) return rc;
  }
  {
    // the last panel: n - SB - c_last reflectors
    const int c_last = ((n - SB - 1) / SB) * SB;
    const int nb = n - SB - c_last;
    if (nb < SB) {
      sb_t_repack_kernel<<<1, 1, 0, ctx->stream>>>(b.Tall + (size_t)(c_last / SB) * SB * SB, nb);
      ITG_LAUNCH_CHECK(ctx);
    }
  }
  // 2. band storage; the reflectors laid out for the back-transform
  sb_band_kernel<<<n, LDB, 0, ctx->stream>>>(A, n, b.AB);
  sb_v1_kernel<<<itg_blocks((int64_t)n * (n - SB), 256, 16384), 256, 0, ctx->stream>>>(A, n);
  ITG_LAUNCH_CHECK(ctx);
  // 3. band -> tridiagonal
  const int nsw = n - 2;
  sb_flags_kernel<<<itg_blocks(nsw, 256, 64), 256, 0, ctx->stream>>>(nsw, b.prog, b.err);
  const int nblk = (nsw + SB2_W - 1) / SB2_W;
  const int nt0 = sb_nt(n, 0);
  const int nwg = std::min(nblk, (nt0 + 2 * SB2_W) / (2 * SB2_W) + 2);
  sb2st_kernel<SB2_W><<<nwg, 64 * SB2_W, 0, ctx->stream>>>(b.AB, n, nsw, b.V2, b.tau2, b.prog,
                                                          b.err);
  ITG_LAUNCH_CHECK(ctx);
  sb_de_kernel<<<itg_blocks(n, 256, 64), 256, 0, ctx->stream>>>(b.AB, n, d, e);
  ITG_LAUNCH_CHECK(ctx);
  int herr = 0;
  ITG_HIP(ctx, hipMemcpyAsync(&herr, b.err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  ITG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (herr)
    return itg_fail(ctx, ITG_EHIP, "bulge chasing: a workgroup's wait on the previous sweep "
                                   "exceeded its bound (not all workgroups resident?)");
  return ITG_OK;
}

int apply_q2(itg_ctx* ctx, int n, int k, double* U, const Bufs& b) {
  const int nsw = n - 2;
  if (nsw <= 0 || k <= 0) return ITG_OK;
  const int ngroups = (nsw + G2 - 1) / G2;
  const int kmax = sb_nt(n, 0);
  q2_t_kernel<<<ngroups * kmax, 256, 0, ctx->stream>>>(n, nsw, kmax, b.V2, b.tau2, b.Tq2);
  ITG_LAUNCH_CHECK(ctx);
  const int ntile = (k + 63) / 64;
  for (int step = 0; step < ngroups + kmax - 1; ++step) {
    const int kmin = std::max(0, step - (ngroups - 1));
    const int kmx = std::min(step, kmax - 1);
    dim3 grid((unsigned)(kmx - kmin + 1), (unsigned)ntile);
    q2_apply_kernel<<<grid, 256, 0, ctx->stream>>>(n, nsw, ngroups, kmax, step, kmin, b.V2,
                                                    b.Tq2, U, n, k);
    ITG_LAUNCH_CHECK(ctx);
  }
  return ITG_OK;
}

}  // namespace eig2
