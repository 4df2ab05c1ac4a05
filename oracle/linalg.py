"""Order-2 factorisations with truncation, restating
src/tensor/culinearalgebra.jl:33-121 on LAPACK (numpy) in place of cuSOLVER."""
import numpy as np

from .truncate import truncate


def eigh_trunc(A, maxdim=None, cutoff=-1.0, absolute_cutoff=False,
               relative_cutoff=True, semantics="ndtensors"):
    """syevd -> reverse to descending -> truncate! (mindim not passed, as at
    culinearalgebra.jl:92) -> keep leading columns; heevd (zheevd) for a
    complex Hermitian A (culinearalgebra.jl:84-85). Returns
    (w_desc_kept, U_kept, truncerr, docut)."""
    A = np.asarray(A)
    A = A.astype(np.complex128 if np.iscomplexobj(A) else np.float64)
    n = A.shape[0]
    w, U = np.linalg.eigh(A, UPLO="L")
    w = w[::-1].copy()
    U = U[:, ::-1].copy()
    if cutoff is None or cutoff < 0:
        k, truncerr, docut = (n if maxdim is None or maxdim <= 0 else min(maxdim, n)), 0.0, 0.0
        if k < n:
            k, truncerr, docut = truncate(w, semantics, maxdim=maxdim, cutoff=0.0,
                                          absolute_cutoff=absolute_cutoff,
                                          relative_cutoff=relative_cutoff)
    else:
        k, truncerr, docut = truncate(w, semantics, maxdim=maxdim, cutoff=cutoff,
                                      absolute_cutoff=absolute_cutoff,
                                      relative_cutoff=relative_cutoff)
    return w[:k], U[:, :k], truncerr, docut


def svd_trunc(A, maxdim=None, mindim=1, cutoff=0.0, absolute_cutoff=False,
              relative_cutoff=True, semantics="ndtensors"):
    """svd! -> P = S.^2 -> truncate! -> slice U, S, V (culinearalgebra.jl:42-57)."""
    A = np.asarray(A, dtype=np.float64)
    U, S, Vt = np.linalg.svd(A, full_matrices=False)
    k, truncerr, docut = truncate(S ** 2, semantics, maxdim=maxdim, mindim=mindim,
                                  cutoff=cutoff, absolute_cutoff=absolute_cutoff,
                                  relative_cutoff=relative_cutoff)
    return U[:, :k], S[:k], Vt[:k, :].T, truncerr, docut


def qr(A):
    """Thin QR (culinearalgebra.jl:110-121)."""
    return np.linalg.qr(np.asarray(A, dtype=np.float64), mode="reduced")
