/*
 * itg.h — C ABI of libitg, an MI355X (gfx950) backend for the dense tensor
 * storage of ITensors.jl.
 *
 * Conventions shared by every entry point
 *  - Tensors are flat device buffers, column-major over their index order
 *    (first index fastest), exactly the layout of NDTensors `Dense` storage
 *    (src/tensor/cudense.jl:19 reshapes `data` to `dims(inds)...`), so a Julia
 *    shim passes `data` pointers with no copy.
 *  - dtype: ITG_F64 or ITG_F32 (storage type of every tensor argument).
 *  - mode : arithmetic of a contraction. ITG_MODE_NATIVE computes in the
 *    storage type (FP64 MFMA for F64); ITG_MODE_F32 rounds operands to FP32
 *    and uses FP32 MFMA; ITG_MODE_BF16X3 splits each operand into hi+lo BF16
 *    parts and sums the three leading products on BF16 MFMA with FP32
 *    accumulation. Results are written in the storage type.
 *  - Every call returns 0 (ITG_OK) or a negative ITG_E* code; the message is
 *    available from itg_last_error(ctx).
 *  - One context owns one HIP stream; calls are asynchronous on it except
 *    those returning host scalars (dot, nrm2, truncate, eigh/svd `k`), which
 *    synchronise, as the reference does at cudense.jl:25,27 and
 *    cutruncate.jl:10,38,56,82. A context is used by one host thread.
 *  - Outputs whose size is data dependent (truncated factorisations) are
 *    written into caller-allocated buffers of the untruncated size; the
 *    library reports the kept dimension k and the caller uses the leading k
 *    columns (column-major, so no compaction is needed).
 */
#ifndef ITG_H
#define ITG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct itg_ctx itg_ctx;

enum itg_status {
  ITG_OK = 0,
  ITG_EDIM = -1,         /* shape/label mismatch  (Julia: DimensionMismatch) */
  ITG_EDTYPE = -2,       /* unsupported dtype     (Julia: ArgumentError)     */
  ITG_EOOM = -3,         /* device allocation failed                         */
  ITG_EHIP = -4,         /* HIP runtime error                                */
  ITG_EUNSUPPORTED = -5, /* valid request this build does not implement      */
  ITG_EARG = -6,         /* invalid argument      (Julia: ArgumentError)     */
  ITG_ECONV = -7         /* iterative method did not converge                */
};

/* ITG_C64: ComplexF64, interleaved (re, im) pairs (Julia's ComplexF64 and
 * Dense{ComplexF64} layout; the reference parameterises its tensor, contraction,
 * eigen and davidson tests over it: test/test_cucontract.jl:9,
 * test/test_cuitensor.jl:10, test/test_cuiterativesolvers.jl:13,
 * src/tensor/culinearalgebra.jl:84-85). Accepted by itg_fill (imaginary part
 * 0), itg_convert (to and from ITG_F64: real part / zero imaginary part),
 * itg_randn, itg_permute, itg_add_permuted, itg_scal, itg_axpby (real
 * coefficients), itg_dot (conj(x).y into out[0], out[1]), itg_nrm2,
 * itg_contract (ITG_MODE_NATIVE, ranks up to ITG_MAX_RANK - 2),
 * itg_eigh_trunc (Hermitian), itg_svd_trunc (A = U diag(S) V^H) and itg_qr
 * (Q unitary); counts n are complex elements. */
enum itg_dtype { ITG_F64 = 0, ITG_F32 = 1, ITG_C64 = 2 };
enum itg_mode { ITG_MODE_NATIVE = 0, ITG_MODE_F32 = 1, ITG_MODE_BF16X3 = 2 };

/* truncate semantics: CPU NDTensors (default) or the reference GPU kernel
 * (src/tensor/cutruncate.jl:1-93), kept for parity with its golden vectors. */
enum itg_trunc_semantics { ITG_TRUNC_NDTENSORS = 0, ITG_TRUNC_REFGPU = 1 };

/* ---------------------------------------------------------------- context */
/* Replaces CUDA.jl device/stream setup (src/ITensorsGPU.jl:3-6). */
int itg_init(int device, itg_ctx** ctx);
int itg_destroy(itg_ctx* ctx);
int itg_sync(itg_ctx* ctx);
const char* itg_last_error(itg_ctx* ctx);
/* Opaque hipStream_t of the context (for callers that record events). */
void* itg_stream(itg_ctx* ctx);
/* Run the context's work on a caller-owned hipStream_t (e.g. the stream of
 * the host framework that owns the device buffers); NULL restores the
 * context's own stream. The caller keeps the stream alive. */
int itg_set_stream(itg_ctx* ctx, void* stream);
/* Device properties: number of CUs and the gfx arch number (e.g. 950). */
int itg_device_info(itg_ctx* ctx, int* num_cu, int* gfx_arch, double* hbm_gb);
/* Per-kernel event timing on the context's stream (no counterpart in the
 * reference; CUDA.jl's @elapsed in its benchmarks plays this role). While
 * enabled, every launch of the kernel classes in `mask` (bit ITG_K_*) is
 * bracketed by HIP events and credited with its algorithmic flops and
 * bytes. A non-zero mask clears and starts, 0 stops. itg_timing_report
 * synchronises and sums one class. */
enum {
  ITG_K_GEMM_SPLIT = 0,   /* split-BF16 GEMM (bf16x3 mode) */
  ITG_K_GETT = 1,         /* contraction / GEMM kernel, every mode */
  ITG_K_WPASS = 2,        /* Heff W1.W2 pass */
  ITG_K_TRD_SYMV = 3,     /* tridiagonalisation: reflector + symmetric product */
  ITG_K_TRD_POST = 4,     /* tridiagonalisation: w and next-column update */
  ITG_K_COUNT = 5
};
int itg_timing(itg_ctx* ctx, int mask);
/* Change the set of timed kernel classes without clearing what was recorded
 * (to time some classes over a sub-range of a run). */
int itg_timing_set(itg_ctx* ctx, int mask);
int itg_timing_report(itg_ctx* ctx, int kind, int64_t* launches, double* ms,
                      double* flops, double* bytes);

/* ----------------------------------------------------------------- memory */
/* Replace CuVector allocation / CuArray(x) / collect(x)
 * (src/tensor/cudense.jl:4-12,62-72; src/cuitensor.jl:18-33). */
int itg_malloc(itg_ctx* ctx, size_t bytes, void** ptr);
int itg_free(itg_ctx* ctx, void* ptr);
int itg_memcpy_h2d(itg_ctx* ctx, void* dst, const void* src, size_t bytes);
int itg_memcpy_d2h(itg_ctx* ctx, void* dst, const void* src, size_t bytes);
int itg_memcpy_d2d(itg_ctx* ctx, void* dst, const void* src, size_t bytes);
/* x[i] = value, i < n (cudense.jl:6-10 `fill!`). */
int itg_fill(itg_ctx* ctx, int dtype, int64_t n, double value, void* x);
/* Cast between storage types (same length). */
int itg_convert(itg_ctx* ctx, int64_t n, int src_dtype, const void* src,
                int dst_dtype, void* dst);

/* -------------------------------------------------------------------- rng */
/* Gaussian fill replacing CURAND randn! (src/cuitensor.jl:37,43-44;
 * src/mps/cumps.jl:47): Philox4x32-10 (Salmon et al., SC'11) keyed by `seed`,
 * counter = offset/4 + i/4, Box-Muller on each 32-bit pair. Element i of the
 * stream is the same on the host oracle and on the device. */
int itg_randn(itg_ctx* ctx, int dtype, int64_t n, uint64_t seed,
              uint64_t offset, void* x);

/* ------------------------------------------------------- data movement */
/* B = alpha * permutedims(A, perm): B has rank `rank`, dims dimsA[perm[j]],
 * and B[j0..] = A[...] with B's mode j being A's mode perm[j] (0-based).
 * Bit-exact for alpha == 1. Replaces Base.permute!/permutedims!!
 * (src/tensor/cudense.jl:38-60,447-500). */
int itg_permute(itg_ctx* ctx, int dtype, int rank, const int64_t* dimsA,
                const int32_t* perm, double alpha, const void* A, void* B);
/* B = alpha * permutedims(A, perm) + beta * B, B dims = dimsA[perm].
 * Replaces the permuted +/- (src/tensor/cudense.jl:333-445). */
int itg_add_permuted(itg_ctx* ctx, int dtype, int rank, const int64_t* dimsA,
                     const int32_t* perm, double alpha, const void* A,
                     double beta, void* B);

/* ------------------------------------------------------------------ BLAS-1 */
/* x *= alpha (cudense.jl:22,502). */
int itg_scal(itg_ctx* ctx, int dtype, int64_t n, double alpha, void* x);
/* y = alpha*x + beta*y. */
int itg_axpby(itg_ctx* ctx, int dtype, int64_t n, double alpha, const void* x,
              double beta, void* y);
/* Davidson diagonal preconditioner t = q ./ (diag - lambda), where
 * |diag - lambda| below `floor` is replaced by floor (sign kept). No
 * counterpart in the reference, whose davidson is unpreconditioned. */
int itg_precond_diag(itg_ctx* ctx, int dtype, int64_t n, const void* q,
                     const void* diag, double lambda, double floor, void* t);
/* x *= (re + i im), x ComplexF64 (n complex elements). */
int itg_zscal(itg_ctx* ctx, int64_t n, double re, double im, void* x);
/* y = (are + i aim) x + (bre + i bim) y, ComplexF64. */
int itg_zaxpby(itg_ctx* ctx, int64_t n, double are, double aim, const void* x,
               double bre, double bim, void* y);
/* x = conj(x) in place (dag of the storage, cudense.jl's conj!); a no-op for
 * real dtypes. */
int itg_conj(itg_ctx* ctx, int dtype, int64_t n, void* x);
/* FP64-accumulated dot and 2-norm returned to the host (cudense.jl:25-27).
 * For ITG_C64 the dot is sum conj(x) y and `out` takes two doubles. */
int itg_dot(itg_ctx* ctx, int dtype, int64_t n, const void* x, const void* y,
            double* out);
int itg_nrm2(itg_ctx* ctx, int dtype, int64_t n, const void* x, double* out);

/* ------------------------------------------------------------ contraction */
/* C = alpha * sum_{contracted} A * B + beta * C.
 * Labels follow the einsum convention: a label present in A and B and absent
 * from C is summed over; every label of C appears in exactly one of A or B.
 * Ranks 0..ITG_MAX_RANK. Replaces contract!!/_contract!/_gemm_contract!/
 * outer! (src/tensor/cudense.jl:74-110,129-331; src/tensor/dense.jl:1-48):
 * index permutations are fused into the operand loads and the epilogue of one
 * MFMA kernel. */
#define ITG_MAX_RANK 16
int itg_contract(itg_ctx* ctx, int dtype, int mode,
                 int rA, const int64_t* dimsA, const int32_t* labelsA, const void* A,
                 int rB, const int64_t* dimsB, const int32_t* labelsB, const void* B,
                 int rC, const int64_t* dimsC, const int32_t* labelsC, void* C,
                 double alpha, double beta);
/* Plain column-major GEMM C(m,n) = alpha*op(A)*op(B) + beta*C, op = 'N'|'T'
 * (order-2 `*`, src/tensor/culinearalgebra.jl:9-17). With A == B (same
 * pointer and leading dimension), 'N','T' or 'T','N', m == n >= 256 and
 * beta == 0 (a density matrix) only the lower triangle is multiplied and the
 * upper one mirrored: C comes out exactly symmetric. */
int itg_gemm(itg_ctx* ctx, int dtype, int mode, char transA, char transB,
             int64_t m, int64_t n, int64_t k, double alpha, const void* A,
             int64_t lda, const void* B, int64_t ldb, double beta, void* C,
             int64_t ldc);

/* ---------------------------------------------------------- Diag storage */
/* C = alpha * sum_{contracted} D * B + beta * C for a Diag tensor D of rank
 * rD (dims dD, labels lD): D(i1..irD) = ddata[k] when every index equals k
 * (nd = min(dD) entries), else 0; with ddata NULL, `uniform` (one value,
 * two doubles re/im for ITG_C64) is on the whole diagonal (UniformDiag).
 * Labels as in itg_contract (no batch modes); a label of D kept in C takes
 * the diagonal position from the output coordinates, and when every label of
 * D is contracted the diagonal is summed. Replaces the Diag x Dense
 * contract! methods of src/tensor/cudiag.jl:118-162 (the uniform one, which
 * the reference runs as A.*B ignoring index order (test_cudiag.jl:49,95
 * @test_broken), follows the labels here). */
int itg_diag_contract(itg_ctx* ctx, int dtype, int rD, const int64_t* dD,
                      const int32_t* lD, const void* ddata, int64_t nd,
                      const double* uniform, int rB, const int64_t* dB,
                      const int32_t* lB, const void* B, int rC, const int64_t* dC,
                      const int32_t* lC, void* C, double alpha, double beta);
/* z = x .* y element-wise: Diag x Diag storage products (cudiag.jl:118-140,
 * `copyto!(Cstore, Astore.*Bstore)`). */
int itg_hadamard(itg_ctx* ctx, int dtype, int64_t n, const void* x,
                 const void* y, void* z);
/* Copy S (rank `rank`, dims dimsS, column-major) into the block of D (dims
 * dimsD) that starts at `offsets`; D's other entries are left as they are.
 * The direct sum of link spaces in MPS/MPO `add` (ITensors' MPS +,
 * test/test_cumpo.jl "add"). */
int itg_embed(itg_ctx* ctx, int dtype, int rank, const int64_t* dimsS,
              const void* S, const int64_t* dimsD, const int64_t* offsets,
              void* D);

/* ------------------------------------------------------------- truncation */
/* P: host array of n non-increasing weights (squared singular values or
 * density-matrix eigenvalues). Returns the kept count, the truncation error
 * and `docut` exactly as truncate! does (src/tensor/cutruncate.jl:1-93 for
 * ITG_TRUNC_REFGPU; NDTensors' CPU truncate! for ITG_TRUNC_NDTENSORS). */
int itg_truncate(itg_ctx* ctx, const double* P, int64_t n, int64_t maxdim,
                 int64_t mindim, double cutoff, int absolute_cutoff,
                 int relative_cutoff, int semantics, int64_t* k,
                 double* truncerr, double* docut);
/* Same decision for weights already resident on the device (length n),
 * computed by one kernel; one 3-value read-back. */
int itg_truncate_device(itg_ctx* ctx, const double* P_dev, int64_t n,
                        int64_t maxdim, int64_t mindim, double cutoff,
                        int absolute_cutoff, int relative_cutoff,
                        int semantics, int64_t* k, double* truncerr,
                        double* docut);

/* -------------------------------------------------------- factorisations */
/* Symmetric eigendecomposition A = U diag(w) U^T of a column-major n x n
 * symmetric matrix (both triangles stored); for ITG_C64 the Hermitian
 * A = U diag(w) U^H (heevd!'s branch, culinearalgebra.jl:84-85), U complex. Eigenvalues are returned in DESCENDING order
 * into `w` (device, FP64, length n) and the matching eigenvectors into the
 * columns of U (device, n x n, dtype). Then truncate! decides k
 * (src/tensor/culinearalgebra.jl:74-108: syevd, reverse, truncate!; mindim
 * is not applied, as there). A is destroyed. Pass maxdim <= 0 for "no
 * maxdim" and cutoff < 0 for no truncation (k = n). */
int itg_eigh_trunc(itg_ctx* ctx, int dtype, int64_t n, void* A, double* w,
                   void* U, int64_t maxdim, double cutoff, int absolute_cutoff,
                   int relative_cutoff, int64_t* k, double* truncerr,
                   double* docut);
/* Truncated SVD of a column-major m x n matrix A = U diag(S) V^T
 * (src/tensor/culinearalgebra.jl:33-72). r = min(m,n). U: m x r, S: r
 * (FP64, descending), V: n x r (device). P = S.^2 feeds truncate! with
 * maxdim/mindim/cutoff; k is the kept count. A is destroyed. */
int itg_svd_trunc(itg_ctx* ctx, int dtype, int64_t m, int64_t n, void* A,
                  void* U, double* S, void* V, int64_t maxdim, int64_t mindim,
                  double cutoff, int absolute_cutoff, int relative_cutoff,
                  int64_t* k, double* truncerr, double* docut);
/* Thin QR of a column-major m x n matrix: Q m x r, R r x n, r = min(m,n)
 * (src/tensor/culinearalgebra.jl:110-121). A is not modified. */
int itg_qr(itg_ctx* ctx, int dtype, int64_t m, int64_t n, const void* A,
           void* Q, void* R);

/* ------------------------------------------- effective Hamiltonian (2-site) */
/* out(a',s1',s2',b') = sum L(a',wl,a) W1(wl,wm,s1',s1) W2(wm,wr,s2',s2)
 *                          R(b',wr,b) psi(a,s1,s2,b)
 * with L (chiL,wl,chiL), W1 (wl,wm,d,d), W2 (wm,wr,d,d), R (chiR,wr,chiR),
 * psi/out (chiL,d,d,chiR). The ProjMPO product of ITensors (reached through
 * dmrg, test/dmrg.jl:27): contraction order L.psi -> .W1 -> .W2 -> .R, with
 * W1.W2 fused into one (wl d d) x (d d wr) operator applied in a single HBM
 * pass; the two large steps run on MFMA in `mode`, the W pass in FP64.
 * `work` must hold itg_heff2_work_elems() 4-byte units (the same count serves
 * FP64 and FP32 applies). */
int64_t itg_heff2_work_elems(int64_t chiL, int64_t wl, int64_t wm, int64_t wr,
                             int64_t d, int64_t chiR);
int itg_heff2_apply(itg_ctx* ctx, int dtype, int mode, int64_t chiL, int64_t wl,
                    int64_t wm, int64_t wr, int64_t d, int64_t chiR,
                    const void* L, const void* W1, const void* W2, const void* R,
                    const void* psi, void* out, void* work);
/* The same product split in two for the Krylov loop at one bond: prepare
 * once per (L, W1, W2, R) into `prep` (itg_heff2_prep_elems() 4-byte units:
 * the fused W1.W2 operator and, in ITG_MODE_BF16X3, L and R split into bf16
 * planes), then apply per vector with the same dtype, mode and dims. The
 * ProjMPO caches its environments per bond the same way (ITensors' ProjMPO
 * holds LR for the current position, reached through dmrg). */
int64_t itg_heff2_prep_elems(int64_t chiL, int64_t wl, int64_t wm, int64_t wr,
                             int64_t d, int64_t chiR);
int itg_heff2_prepare(itg_ctx* ctx, int dtype, int mode, int64_t chiL, int64_t wl,
                      int64_t wm, int64_t wr, int64_t d, int64_t chiR,
                      const void* L, const void* W1, const void* W2, const void* R,
                      void* prep);
int itg_heff2_apply_prepared(itg_ctx* ctx, int dtype, int mode, int64_t chiL,
                             int64_t wl, int64_t wm, int64_t wr, int64_t d,
                             int64_t chiR, const void* L, const void* R,
                             const void* prep, const void* psi, void* out,
                             void* work);
/* Davidson for the lowest eigenpair of the prepared operator (the eigsolve
 * dmrg runs per bond, test/dmrg.jl:27 through ITensors' davidson), on the
 * device: FP64 Krylov basis of up to maxiter+1 (<= 17) vectors, two
 * Gram-Schmidt passes with batched dot products, the projected eigenproblem
 * on the host, ITensors' stopping rule on the residual norm and errgoal.
 * psi (FP64, chiL x d x d x chiR) is the start vector and receives the
 * normalised Ritz vector; diag (the operator's diagonal) selects the
 * diagonal preconditioner, null none. work: itg_heff2_work_elems() 4-byte
 * units; kwork:
 * itg_davidson2_work_elems() doubles. In ITG_MODE_BF16X3 the Heff apply is
 * captured once per (shape, buffers) into a HIP graph and replayed. */
int64_t itg_davidson2_work_elems(int64_t chiL, int64_t d, int64_t chiR, int maxiter);
int itg_davidson2(itg_ctx* ctx, int dtype, int mode, int64_t chiL, int64_t wl,
                  int64_t wm, int64_t wr, int64_t d, int64_t chiR, const void* L,
                  const void* R, const void* prep, void* psi, int maxiter,
                  double errgoal, const void* diag, void* work, void* kwork,
                  double* energy, int* matvecs);
/* itg_davidson2 with the learned preconditioner (no counterpart in the
 * reference, whose davidson has no preconditioner): each correction is
 * first the network's (itg_mlp_features -> itg_mlp_infer -> itg_mlp_scatter
 * on the residual, in = 3 d^2, out = d^2 per bond pair, plus the skip step
 * at rms 1: skip 0 none, 1 the residual, 2 the diagonal step with `floor`)
 * and is kept only if it lowers the residual norm; otherwise the basis
 * vector is dropped, the previous Ritz pair restored and the diagonal step
 * taken. Only the first `net_steps` steps of a call try the network (0:
 * every step). `diag` (the Heff diagonal) is required. mwork:
 * itg_davidson2_mlp_work_elems() floats; the rest as itg_davidson2.
 * `reverted` counts the learned steps taken back. */
int64_t itg_davidson2_mlp_work_elems(int64_t chiL, int64_t d, int64_t chiR, int hidden);
int itg_davidson2_mlp(itg_ctx* ctx, int dtype, int mode, int64_t chiL, int64_t wl,
                      int64_t wm, int64_t wr, int64_t d, int64_t chiR, const void* L,
                      const void* R, const void* prep, void* psi, int maxiter,
                      double errgoal, const void* diag, const float* params,
                      int hidden, int skip, int net_steps, double floor, void* mwork,
                      void* work, void* kwork, double* energy, int* matvecs,
                      int* reverted);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). The same product as itg_heff2_apply, computed by the
 * same kernels, with no `work` argument: the work buffer and any GEMM
 * split-K partials live in a fixed library-owned device array of
 * itg_heff2_noalloc_max_bytes() bytes (64 MiB; one copy per GPU, held until
 * the process ends), so the call makes no hipMalloc/hipFree. Shapes whose work
 * buffer exceeds it (about chi > 380 at w = 5, d = 2) fail with
 * ITG_EUNSUPPORTED. Every call on one GPU shares that array: calls that
 * overlap on one GPU (two streams, two host threads) race. Per-kernel event
 * timing is off during the call. */
int64_t itg_heff2_noalloc_max_bytes(void);
int itg_heff2_apply_noalloc(itg_ctx* ctx, int dtype, int mode, int64_t chiL,
                            int64_t wl, int64_t wm, int64_t wr, int64_t d,
                            int64_t chiR, const void* L, const void* W1,
                            const void* W2, const void* R, const void* psi,
                            void* out);
/* Diagonal of the same operator, shape (chiL,d,d,chiR). */
int itg_heff2_diag(itg_ctx* ctx, int dtype, int64_t chiL, int64_t wl,
                   int64_t wm, int64_t wr, int64_t d, int64_t chiR,
                   const void* L, const void* W1, const void* W2, const void* R,
                   void* diag);

/* ------------------------------------ learned Davidson preconditioner (MLP) */
/* No counterpart in the reference. A 3-layer MLP (in -> h -> h -> out, GELU)
 * over feature rows; FP32 weights, row-major activations X[rows][in].
 * Parameters are packed [W1(h x in), b1(h), W2(h x h), b2(h), W3(out x h),
 * b3(out)]; `acts` holds 4*rows*h floats of saved activations for backward. */
int64_t itg_mlp_param_count(int in, int hidden, int out);
int itg_mlp_forward(itg_ctx* ctx, int rows, int in, int hidden, int out,
                    const float* params, const float* X, float* Y, float* acts);
/* Inference only: Y = MLP(X) with no saved activations, for the
 * preconditioner's many-row evaluations. hidden == 256 with in <= 16 and
 * out <= 4, or in <= 32 and out <= 9, run one fused kernel (activations stay on chip; the hidden x
 * hidden product in bf16 with FP32 accumulation, ~3e-3 relative to the FP32
 * forward; params 16-byte aligned) and leave `work` unused; other shapes
 * use the forward's FP32 products with `work` (2*rows*hidden floats) for the
 * activations. */
int itg_mlp_infer(itg_ctx* ctx, int rows, int in, int hidden, int out,
                  const float* params, const float* X, float* Y, float* work);
/* Mean-squared-error loss against T; accumulates parameter gradients into
 * grad (overwritten) and returns the loss to the host. */
int itg_mlp_backward(itg_ctx* ctx, int rows, int in, int hidden, int out,
                     const float* params, const float* X, const float* T,
                     const float* acts, const float* Y, float* grad,
                     double* loss);
/* One Adam step (Kingma & Ba 2015) on packed parameters. */
int itg_mlp_adam(itg_ctx* ctx, int64_t nparam, float* params, const float* grad,
                 float* m, float* v, int step, float lr, float beta1,
                 float beta2, float eps);
/* Feature rows of a two-site FP64 residual r (chiL,d,d,chiR): one row per
 * bond pair (a,b), row index a + chiL*b; rows row0 + k*row_step for
 * k < nrows. Each row holds 3*d*d floats: r/rms(r), delta/rms(delta) and
 * (r/delta)/rms(r/delta) over the d*d physical block, delta = diag - lambda
 * clipped to |delta| >= floor. The rms scales are taken over all of r. */
int itg_mlp_features(itg_ctx* ctx, int64_t chiL, int64_t d, int64_t chiR,
                     const double* r, const double* diag, double lambda,
                     double floor, int64_t row0, int64_t row_step,
                     int64_t nrows, float* X);
/* The same rows of an FP64 vector v, times scale, as d*d floats per row. */
int itg_mlp_gather(itg_ctx* ctx, int64_t chiL, int64_t d, int64_t chiR,
                   const double* v, double scale, int64_t row0,
                   int64_t row_step, int64_t nrows, float* T);
/* Inverse of itg_mlp_gather for the contiguous rows row0..row0+nrows-1. */
int itg_mlp_scatter(itg_ctx* ctx, int64_t chiL, int64_t d, int64_t chiR,
                    const float* Y, int64_t row0, int64_t nrows, double* v);

#ifdef __cplusplus
}
#endif
#endif /* ITG_H */
