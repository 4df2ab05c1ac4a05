// Two-stage reduction of a symmetric matrix to tridiagonal form (eigh2.hip),
// used by itg_eigh_trunc for large n.
#pragma once

#include <cstddef>

#include "itg_internal.h"

namespace eig2 {

constexpr int SB = 32;   // bandwidth of the intermediate band matrix

// device buffers of the two-stage path (carved from the eigensolver
// workspace by the caller; sizes from `doubles_needed`)
struct Bufs {
  double* AB;      // band + bulge storage, 2 SB rows per column
  double* V2;      // bulge-chasing reflectors, SB per task
  double* tau2;    // their scalar factors
  double* Tq2;     // compact-WY factors of the back-transform blocks
  double* Tall;    // compact-WY factor of every band-reduction panel
  double* VW;      // [V | W | V], n x SB each, of the current panel
  double* X;       // A22 V of the current panel (n x SB)
  double* Mpart;   // partial V^T X per 256-row chunk
  double* Qleaf;   // explicit Q of the TSQR leaves (n x SB)
  double* Rstack;  // stacked leaf R factors (1024 x SB)
  double* Qtop;    // explicit Q of the stacked-R factorisation (1024 x SB)
  double* Rsm;     // its R (SB x SB)
  int* prog;       // tasks completed per sweep (bulge chasing)
  int* err;        // set when a bulge-chasing wait exceeds its bound
};

// doubles of workspace the buffers above take for an n x n matrix
size_t doubles_needed(int n);
// carve them from `base` (256-byte aligned pieces); returns bytes used
size_t carve(int n, void* base, Bufs* b);

// A (n x n, both triangles) -> tridiagonal (d, e). A keeps the band
// reduction's reflectors in the layout back_transform expects (column j:
// zero above row j + SB, one at j + SB, the reflector below), and b.Tall
// their compact-WY factors per SB-column panel; b.V2/tau2 keep the bulge
// chasing's reflectors for apply_q2.
int reduce(itg_ctx* ctx, int n, double* A, double* d, double* e, const Bufs& b);

// U (n x k, leading dimension n) <- Q2 U, Q2 the bulge chasing's orthogonal
// factor (the band matrix is Q2 T Q2^T).
int apply_q2(itg_ctx* ctx, int n, int k, double* U, const Bufs& b);

}  // namespace eig2
