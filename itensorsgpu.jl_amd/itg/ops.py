"""Storage-level operations on device buffers, one call into libitg each.

A device tensor is a flat torch tensor on the GPU holding column-major data
(first index fastest), the layout of NDTensors `Dense` storage
(src/tensor/cudense.jl:19). Shapes travel separately as tuples of ints.
These functions are what a Julia `RocDense` shim would `ccall`; the ITensor
layer in `itensor.py` builds on them.
"""
import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import ctx, i32arr, i64arr, lib, ptr, dtype_code


def _torch():
    import torch
    return torch


def empty(n, dtype="float64"):
    torch = _torch()
    return torch.empty(max(int(n), 0), dtype=getattr(torch, dtype), device="cuda")


def zeros(n, dtype="float64"):
    torch = _torch()
    return torch.zeros(max(int(n), 0), dtype=getattr(torch, dtype), device="cuda")


def prod(dims):
    p = 1
    for d in dims:
        p *= int(d)
    return p


def to_device(host_flat, dtype="float64"):
    """roc(.) data movement: host column-major flat array -> device buffer."""
    torch = _torch()
    a = np.ascontiguousarray(np.asarray(host_flat, dtype=np.dtype(dtype)).ravel())
    t = empty(a.size, dtype)
    if a.size:
        c = ctx()
        c.check(lib().itg_memcpy_h2d(c.ptr, ptr(t), a.ctypes.data_as(C.c_void_p), a.nbytes))
    return t


_NP = {"float64": np.float64, "float32": np.float32, "complex128": np.complex128}


def dtype_name(t):
    return str(t.dtype).split(".")[-1]


def to_host(t):
    """cpu(.) data movement: device buffer -> host numpy array (flat)."""
    out = np.empty(t.numel(), dtype=_NP[dtype_name(t)])
    if out.size:
        c = ctx()
        c.check(lib().itg_memcpy_d2h(c.ptr, out.ctypes.data_as(C.c_void_p), ptr(t), out.nbytes))
    return out


def fill(t, value):
    c = ctx()
    c.check(lib().itg_fill(c.ptr, dtype_code(t), t.numel(), float(value), ptr(t)))
    return t


def randn(n, seed, offset=0, dtype="float64"):
    t = empty(n, dtype)
    c = ctx()
    c.check(lib().itg_randn(c.ptr, dtype_code(t), int(n), int(seed), int(offset), ptr(t)))
    return t


def convert(t, dtype):
    out = empty(t.numel(), dtype)
    c = ctx()
    c.check(lib().itg_convert(c.ptr, t.numel(), dtype_code(t), ptr(t), dtype_code(out), ptr(out)))
    return out


# ------------------------------------------------------------------ BLAS-1
def _is_cplx(a):
    return isinstance(a, complex) or np.iscomplexobj(a)


def scal(t, alpha):
    """t *= alpha; a complex alpha needs complex storage."""
    c = ctx()
    if _is_cplx(alpha):
        if not t.is_complex():
            raise _lib.ArgumentError("complex scale of real storage")
        a = complex(alpha)
        c.check(lib().itg_zscal(c.ptr, t.numel(), a.real, a.imag, ptr(t)))
        return t
    c.check(lib().itg_scal(c.ptr, dtype_code(t), t.numel(), float(alpha), ptr(t)))
    return t


def axpby(alpha, x, beta, y):
    """y = alpha*x + beta*y (in place on y); complex coefficients need
    complex storage."""
    c = ctx()
    if _is_cplx(alpha) or _is_cplx(beta):
        if not (x.is_complex() and y.is_complex()):
            raise _lib.ArgumentError("complex coefficients on real storage")
        a, b = complex(alpha), complex(beta)
        c.check(lib().itg_zaxpby(c.ptr, y.numel(), a.real, a.imag, ptr(x), b.real, b.imag,
                                 ptr(y)))
        return y
    c.check(lib().itg_axpby(c.ptr, dtype_code(y), y.numel(), float(alpha), ptr(x), float(beta),
                            ptr(y)))
    return y


def conj(t):
    """t = conj(t) in place (no-op on real storage)."""
    c = ctx()
    c.check(lib().itg_conj(c.ptr, dtype_code(t), t.numel(), ptr(t)))
    return t


def precond_diag(q, diag, lam, floor=1e-8):
    """t = q ./ (diag - lam) (floored), the Davidson diagonal correction."""
    t = empty(q.numel(), str(q.dtype).split(".")[-1])
    c = ctx()
    c.check(lib().itg_precond_diag(c.ptr, dtype_code(q), q.numel(), ptr(q), ptr(diag), float(lam),
                                   float(floor), ptr(t)))
    return t


def dot(x, y):
    """sum conj(x) y (FP64-accumulated); complex for complex storage."""
    c = ctx()
    out = (C.c_double * 2)()
    c.check(lib().itg_dot(c.ptr, dtype_code(x), x.numel(), ptr(x), ptr(y), out))
    if x.is_complex():
        return complex(out[0], out[1])
    return out[0]


def nrm2(x):
    c = ctx()
    out = C.c_double()
    c.check(lib().itg_nrm2(c.ptr, dtype_code(x), x.numel(), ptr(x), C.byref(out)))
    return out.value


# ----------------------------------------------------------- data movement
def permute(A, dimsA, perm, alpha=1.0, out=None):
    """B = alpha*permutedims(A, perm) (0-based perm); returns (B, dimsB)."""
    dimsB = tuple(int(dimsA[p]) for p in perm)
    B = out if out is not None else empty(prod(dimsB), str(A.dtype).split(".")[-1])
    c = ctx()
    c.check(lib().itg_permute(c.ptr, dtype_code(A), len(dimsA), i64arr(dimsA), i32arr(perm),
                              float(alpha), ptr(A), ptr(B)))
    return B, dimsB


def add_permuted(A, dimsA, perm, alpha, B, beta):
    """B = alpha*permutedims(A, perm) + beta*B, in place on B."""
    c = ctx()
    c.check(lib().itg_add_permuted(c.ptr, dtype_code(A), len(dimsA), i64arr(dimsA), i32arr(perm),
                                   float(alpha), ptr(A), float(beta), ptr(B)))
    return B


# ------------------------------------------------------------- contraction
def contract(A, dimsA, labelsA, B, dimsB, labelsB, labelsC, mode="native", alpha=1.0,
             beta=0.0, out=None):
    """C = alpha*A.B + beta*C with einsum-style integer labels; returns (C, dimsC)."""
    dmap = {}
    for l, d in zip(labelsA, dimsA):
        dmap[l] = int(d)
    for l, d in zip(labelsB, dimsB):
        dmap.setdefault(l, int(d))
    try:
        dimsC = tuple(dmap[l] for l in labelsC)
    except KeyError as e:
        raise _lib.DimensionMismatch("output label %r missing from inputs" % (e.args[0],))
    if out is None:
        out = empty(prod(dimsC), str(A.dtype).split(".")[-1])
    c = ctx()
    c.check(lib().itg_contract(c.ptr, dtype_code(A), _lib.MODES[mode],
                               len(dimsA), i64arr(dimsA), i32arr(labelsA), ptr(A),
                               len(dimsB), i64arr(dimsB), i32arr(labelsB), ptr(B),
                               len(labelsC), i64arr(dimsC), i32arr(labelsC), ptr(out),
                               float(alpha), float(beta)))
    return out, dimsC


def diag_contract(D, dimsD, labelsD, B, dimsB, labelsB, labelsC, uniform=None, alpha=1.0,
                  beta=0.0, out=None, dtype=None):
    """C = alpha * Diag(D) . B + beta * C; D is the device diagonal (length
    min(dimsD)) or None with `uniform` a number (UniformDiag)."""
    dmap = dict(zip(labelsB, dimsB))
    dmap.update(zip(labelsD, dimsD))
    dimsC = tuple(int(dmap[l]) for l in labelsC)
    dt = dtype or dtype_name(B)
    if out is None:
        out = empty(prod(dimsC), dt)
    u = None
    if D is None:
        z = complex(uniform)
        u = (C.c_double * 2)(z.real, z.imag)
    c = ctx()
    c.check(lib().itg_diag_contract(c.ptr, dtype_code(B), len(dimsD), i64arr(dimsD),
                                    i32arr(labelsD), ptr(D) if D is not None else None,
                                    min(dimsD) if D is not None else 0, u,
                                    len(dimsB), i64arr(dimsB), i32arr(labelsB), ptr(B),
                                    len(labelsC), i64arr(dimsC), i32arr(labelsC), ptr(out),
                                    float(alpha), float(beta)))
    return out, dimsC


def hadamard(x, y, out=None):
    """out = x .* y element-wise."""
    out = out if out is not None else empty(x.numel(), dtype_name(x))
    c = ctx()
    c.check(lib().itg_hadamard(c.ptr, dtype_code(x), x.numel(), ptr(x), ptr(y), ptr(out)))
    return out


def embed(S, dimsS, D, dimsD, offsets):
    """Copy S into the block of D starting at `offsets` (in place on D)."""
    c = ctx()
    c.check(lib().itg_embed(c.ptr, dtype_code(S), len(dimsS), i64arr(dimsS), ptr(S),
                            i64arr(dimsD), i64arr(offsets), ptr(D)))
    return D


def gemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, Cm, ldc, mode="native"):
    c = ctx()
    c.check(lib().itg_gemm(c.ptr, dtype_code(A), _lib.MODES[mode], ta.encode(), tb.encode(),
                           m, n, k, float(alpha), ptr(A), lda, ptr(B), ldb, float(beta),
                           ptr(Cm), ldc))
    return Cm


# -------------------------------------------------------------- truncation
def truncate(P, maxdim=None, mindim=1, cutoff=0.0, absolute_cutoff=False,
             relative_cutoff=True, semantics="ndtensors"):
    """truncate!(P; ...) -> (n_kept, truncerr, docut) for host weights."""
    a, pa = _lib.host_f64(P)
    k, te, dc = C.c_int64(), C.c_double(), C.c_double()
    c = ctx()
    c.check(lib().itg_truncate(c.ptr, pa, a.size, int(maxdim or 0), int(mindim), float(cutoff),
                               int(absolute_cutoff), int(relative_cutoff),
                               _lib.TRUNC_REFGPU if semantics == "refgpu" else _lib.TRUNC_NDTENSORS,
                               C.byref(k), C.byref(te), C.byref(dc)))
    return k.value, te.value, dc.value


def truncate_device(P, maxdim=None, mindim=1, cutoff=0.0, absolute_cutoff=False,
                    relative_cutoff=True, semantics="ndtensors"):
    k, te, dc = C.c_int64(), C.c_double(), C.c_double()
    c = ctx()
    c.check(lib().itg_truncate_device(c.ptr, ptr(P), P.numel(), int(maxdim or 0), int(mindim),
                                      float(cutoff), int(absolute_cutoff), int(relative_cutoff),
                                      _lib.TRUNC_REFGPU if semantics == "refgpu"
                                      else _lib.TRUNC_NDTENSORS,
                                      C.byref(k), C.byref(te), C.byref(dc)))
    return k.value, te.value, dc.value


# ----------------------------------------------------------- factorisations
def eigh_trunc(A, n, maxdim=None, cutoff=-1.0, absolute_cutoff=False, relative_cutoff=True):
    """Eigen-decomposition of a symmetric n x n device matrix (destroyed).
    Returns (w_desc[:k] device, U device n x n buffer (first k columns), k,
    truncerr, docut)."""
    dt = str(A.dtype).split(".")[-1]
    w = empty(n, "float64")
    U = empty(n * n, dt)
    k, te, dc = C.c_int64(), C.c_double(), C.c_double()
    c = ctx()
    c.check(lib().itg_eigh_trunc(c.ptr, dtype_code(A), n, ptr(A), ptr(w), ptr(U),
                                 int(maxdim or 0), float(-1.0 if cutoff is None else cutoff),
                                 int(absolute_cutoff), int(relative_cutoff),
                                 C.byref(k), C.byref(te), C.byref(dc)))
    return w, U, k.value, te.value, dc.value


def svd_trunc(A, m, n, maxdim=None, mindim=1, cutoff=0.0, absolute_cutoff=False,
              relative_cutoff=True):
    dt = str(A.dtype).split(".")[-1]
    r = min(m, n)
    U = empty(m * r, dt)
    S = empty(r, "float64")
    V = empty(n * r, dt)
    k, te, dc = C.c_int64(), C.c_double(), C.c_double()
    c = ctx()
    c.check(lib().itg_svd_trunc(c.ptr, dtype_code(A), m, n, ptr(A), ptr(U), ptr(S), ptr(V),
                                int(maxdim or 0), int(mindim), float(cutoff),
                                int(absolute_cutoff), int(relative_cutoff),
                                C.byref(k), C.byref(te), C.byref(dc)))
    return U, S, V, k.value, te.value, dc.value


def qr(A, m, n):
    dt = str(A.dtype).split(".")[-1]
    r = min(m, n)
    Q = empty(m * r, dt)
    R = empty(r * n, dt)
    c = ctx()
    c.check(lib().itg_qr(c.ptr, dtype_code(A), m, n, ptr(A), ptr(Q), ptr(R)))
    return Q, R


# ------------------------------------------------------------------- Heff
def heff2_work_bytes(chiL, wl, wm, wr, d, chiR):
    """Bytes of the Heff work buffer (itg_heff2_work_elems counts 4-byte
    units, enough for FP64 and FP32 applies alike)."""
    return 4 * int(lib().itg_heff2_work_elems(chiL, wl, wm, wr, d, chiR))


def heff2_work(chiL, wl, wm, wr, d, chiR, dtype=None):
    """Work buffer of one Heff apply, allocated in the 4-byte units the
    library counts it in (`dtype` is accepted for the FP64/FP32 callers and
    does not change the size)."""
    return empty(lib().itg_heff2_work_elems(chiL, wl, wm, wr, d, chiR), "float32")


def heff2_apply(L, W1, W2, R, psi, chiL, wl, wm, wr, d, chiR, mode="native", out=None,
                work=None):
    """ProjMPO two-site product: out = Heff psi (psi, out: chiL x d x d x chiR)."""
    dt = str(psi.dtype).split(".")[-1]
    if out is None:
        out = empty(chiL * d * d * chiR, dt)
    if work is None:
        work = heff2_work(chiL, wl, wm, wr, d, chiR, dt)
    c = ctx()
    c.check(lib().itg_heff2_apply(c.ptr, dtype_code(psi), _lib.MODES[mode], chiL, wl, wm, wr, d,
                                  chiR, ptr(L), ptr(W1), ptr(W2), ptr(R), ptr(psi), ptr(out),
                                  ptr(work)))
    return out


def heff2_apply_noalloc(L, W1, W2, R, psi, chiL, wl, wm, wr, d, chiR, mode="native", out=None):
    """heff2_apply with its work in the library's fixed device scratch (no
    device allocation by the library; `out` still comes from the caller)."""
    dt = str(psi.dtype).split(".")[-1]
    if out is None:
        out = empty(chiL * d * d * chiR, dt)
    c = ctx()
    c.check(lib().itg_heff2_apply_noalloc(c.ptr, dtype_code(psi), _lib.MODES[mode], chiL, wl, wm,
                                          wr, d, chiR, ptr(L), ptr(W1), ptr(W2), ptr(R), ptr(psi),
                                          ptr(out)))
    return out


def heff2_prep(chiL, wl, wm, wr, d, chiR):
    """Buffer for itg_heff2_prepare (counted in 4-byte units)."""
    return empty(lib().itg_heff2_prep_elems(chiL, wl, wm, wr, d, chiR), "float32")


def heff2_prepare(L, W1, W2, R, chiL, wl, wm, wr, d, chiR, mode="native", prep=None):
    if prep is None:
        prep = heff2_prep(chiL, wl, wm, wr, d, chiR)
    c = ctx()
    c.check(lib().itg_heff2_prepare(c.ptr, dtype_code(L), _lib.MODES[mode], chiL, wl, wm, wr, d,
                                    chiR, ptr(L), ptr(W1), ptr(W2), ptr(R), ptr(prep)))
    return prep


def heff2_apply_prepared(L, R, prep, psi, chiL, wl, wm, wr, d, chiR, mode="native", out=None,
                         work=None):
    dt = str(psi.dtype).split(".")[-1]
    if out is None:
        out = empty(chiL * d * d * chiR, dt)
    if work is None:
        work = heff2_work(chiL, wl, wm, wr, d, chiR, dt)
    c = ctx()
    c.check(lib().itg_heff2_apply_prepared(c.ptr, dtype_code(psi), _lib.MODES[mode], chiL, wl, wm,
                                           wr, d, chiR, ptr(L), ptr(R), ptr(prep), ptr(psi),
                                           ptr(out), ptr(work)))
    return out


def davidson2_work(chiL, d, chiR, maxiter):
    return empty(lib().itg_davidson2_work_elems(chiL, d, chiR, int(maxiter)), "float64")


def davidson2(L, R, prep, psi, chiL, wl, wm, wr, d, chiR, maxiter=2, errgoal=1e-14, diag=None,
              mode="native", work=None, kwork=None):
    """Device Davidson on the prepared Heff; psi (FP64) is the start vector
    and is overwritten with the normalised Ritz vector. Returns
    (energy, matvecs)."""
    if work is None:
        work = heff2_work(chiL, wl, wm, wr, d, chiR, "float64")
    if kwork is None:
        kwork = davidson2_work(chiL, d, chiR, maxiter)
    e = C.c_double()
    nmv = C.c_int()
    c = ctx()
    c.check(lib().itg_davidson2(c.ptr, dtype_code(psi), _lib.MODES[mode], chiL, wl, wm, wr, d,
                                chiR, ptr(L), ptr(R), ptr(prep), ptr(psi), int(maxiter),
                                float(errgoal), ptr(diag) if diag is not None else None,
                                ptr(work), ptr(kwork), C.byref(e), C.byref(nmv)))
    return e.value, nmv.value


SKIPS = {"none": 0, "plain": 1, "diag": 2}


def davidson2_mlp_work(chiL, d, chiR, hidden):
    return empty(lib().itg_davidson2_mlp_work_elems(chiL, d, chiR, int(hidden)), "float32")


def davidson2_mlp(L, R, prep, psi, chiL, wl, wm, wr, d, chiR, diag, params, hidden, skip="diag",
                  floor=1e-8, maxiter=2, errgoal=1e-14, mode="native", work=None, kwork=None,
                  mwork=None, net_steps=0):
    """Device Davidson with the guarded learned correction (itg_davidson2_mlp);
    the first `net_steps` steps try the network (0: all). Returns (energy,
    matvecs, reverted)."""
    if work is None:
        work = heff2_work(chiL, wl, wm, wr, d, chiR, "float64")
    if kwork is None:
        kwork = davidson2_work(chiL, d, chiR, maxiter)
    if mwork is None:
        mwork = davidson2_mlp_work(chiL, d, chiR, hidden)
    e = C.c_double()
    nmv, nrev = C.c_int(), C.c_int()
    c = ctx()
    c.check(lib().itg_davidson2_mlp(c.ptr, dtype_code(psi), _lib.MODES[mode], chiL, wl, wm, wr,
                                    d, chiR, ptr(L), ptr(R), ptr(prep), ptr(psi), int(maxiter),
                                    float(errgoal), ptr(diag), ptr(params), int(hidden),
                                    SKIPS[skip], int(net_steps), float(floor), ptr(mwork),
                                    ptr(work), ptr(kwork),
                                    C.byref(e), C.byref(nmv), C.byref(nrev)))
    return e.value, nmv.value, nrev.value


def heff2_diag(L, W1, W2, R, chiL, wl, wm, wr, d, chiR):
    dt = str(L.dtype).split(".")[-1]
    out = empty(chiL * d * d * chiR, dt)
    c = ctx()
    c.check(lib().itg_heff2_diag(c.ptr, dtype_code(L), chiL, wl, wm, wr, d, chiR, ptr(L),
                                 ptr(W1), ptr(W2), ptr(R), ptr(out)))
    return out


def copy(t):
    """Device-to-device copy of a buffer (stream-ordered)."""
    out = empty(t.numel(), str(t.dtype).split(".")[-1])
    if t.numel():
        c = ctx()
        c.check(lib().itg_memcpy_d2d(c.ptr, ptr(out), ptr(t), t.numel() * t.element_size()))
    return out


# ------------------------------------------------- learned preconditioner
def mlp_param_count(i, h, o):
    return int(lib().itg_mlp_param_count(int(i), int(h), int(o)))


def mlp_forward(P, X, rows, i, h, o, Y=None, acts=None):
    """Y = MLP(X) for `rows` rows of `i` features; returns (Y, acts)."""
    if Y is None:
        Y = empty(rows * o, "float32")
    if acts is None:
        acts = empty(4 * rows * h, "float32")
    c = ctx()
    c.check(lib().itg_mlp_forward(c.ptr, int(rows), int(i), int(h), int(o), ptr(P), ptr(X), ptr(Y),
                                  ptr(acts)))
    return Y, acts


def mlp_infer(P, X, rows, i, h, o, Y=None, work=None):
    """Y = MLP(X) for inference (no saved activations, split-BF16 hidden
    product); work: 2*rows*h floats."""
    if Y is None:
        Y = empty(rows * o, "float32")
    if work is None:
        work = empty(2 * rows * h, "float32")
    c = ctx()
    c.check(lib().itg_mlp_infer(c.ptr, int(rows), int(i), int(h), int(o), ptr(P), ptr(X), ptr(Y),
                                ptr(work)))
    return Y


def mlp_backward(P, X, T, acts, Y, rows, i, h, o, G=None):
    """Gradient of mean((Y - T)^2) into G; returns (G, loss)."""
    if G is None:
        G = empty(P.numel(), "float32")
    loss = C.c_double()
    c = ctx()
    c.check(lib().itg_mlp_backward(c.ptr, int(rows), int(i), int(h), int(o), ptr(P), ptr(X), ptr(T),
                                   ptr(acts), ptr(Y), ptr(G), C.byref(loss)))
    return G, loss.value


def mlp_adam(P, G, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    c = ctx()
    c.check(lib().itg_mlp_adam(c.ptr, P.numel(), ptr(P), ptr(G), ptr(m), ptr(v), int(step),
                               float(lr), float(b1), float(b2), float(eps)))
    return P


def mlp_features(r, diag, lam, chiL, d, chiR, floor=1e-8, row0=0, step=1, nrows=None, X=None):
    """Feature rows (3*d*d floats each) of the residual r at Ritz value lam."""
    if nrows is None:
        nrows = (chiL * chiR - row0 + step - 1) // step
    if X is None:
        X = empty(nrows * 3 * d * d, "float32")
    c = ctx()
    c.check(lib().itg_mlp_features(c.ptr, chiL, d, chiR, ptr(r), ptr(diag), float(lam), float(floor),
                                   int(row0), int(step), int(nrows), ptr(X)))
    return X


def mlp_gather(v, scale, chiL, d, chiR, row0=0, step=1, nrows=None, T=None):
    if nrows is None:
        nrows = (chiL * chiR - row0 + step - 1) // step
    if T is None:
        T = empty(nrows * d * d, "float32")
    c = ctx()
    c.check(lib().itg_mlp_gather(c.ptr, chiL, d, chiR, ptr(v), float(scale), int(row0), int(step),
                                 int(nrows), ptr(T)))
    return T


def mlp_scatter(Y, chiL, d, chiR, row0, nrows, v):
    c = ctx()
    c.check(lib().itg_mlp_scatter(c.ptr, chiL, d, chiR, ptr(Y), int(row0), int(nrows), ptr(v)))
    return v
