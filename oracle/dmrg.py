"""Two-site DMRG in numpy FP64: random MPS, ProjMPO environments and product,
davidson, replacebond!/factorize, sweep driver.

Follows the ITensors 0.2 algorithms the reference reaches through
`dmrg(H, psi0, sweeps)` (test/dmrg.jl:27,75; examples/dmrg.jl:25):
  - sweepnext: bonds 1..N-1 left to right, then N-1..1 (2(N-1) updates/sweep)
  - ProjMPO product noprime(((psi*L)*W_b)*W_{b+1})*R (fixed order)
  - davidson: Rayleigh-Ritz on an orthogonalised Krylov basis, host-side
    projected eigenproblem, optional diagonal preconditioner
  - factorize: which_decomp "eigen" when cutoff >= 1e-12 (density matrix
    rho = phi phi^dag, eigen + truncate!), else "svd"; then normalise.
Shapes: MPS A[l, s, r]; MPO W[wl, wr, s_out, s_in]; L[a', w, a]; R[b', w, b].
"""
import numpy as np

from . import philox
from .linalg import eigh_trunc, svd_trunc


# ----------------------------------------------------------------- MPS init
def link_dims(N, d, chi):
    return [1] + [int(min(chi, d ** min(j, N - j))) for j in range(1, N)] + [1]


def random_mps(N, d, chi, seed):
    """Gaussian site tensors drawn in site order from one Philox stream,
    then right-orthogonalised (centre at site 0) and normalised."""
    dims = link_dims(N, d, chi)
    off = 0
    mps = []
    for j in range(N):
        n = dims[j] * d * dims[j + 1]
        x = philox.randn(n, seed, off)
        x = x / np.linalg.norm(x)           # normalize!(M[i]), cumps.jl:48
        off += n
        mps.append(np.reshape(x, (dims[j], d, dims[j + 1]), order="F"))
    right_orthogonalize(mps)
    mps[0] = mps[0] / np.linalg.norm(mps[0])
    return mps


def right_orthogonalize(mps):
    """LQ sweeps from the right end (orthogonalize!(psi, 1))."""
    for j in range(len(mps) - 1, 0, -1):
        A = mps[j]
        dl, d, dr = A.shape
        M = np.reshape(A, (dl, d * dr), order="F")
        Q, R = np.linalg.qr(M.T)                # M = R^T Q^T
        k = Q.shape[1]
        mps[j] = np.reshape(Q.T, (k, d, dr), order="F")
        mps[j - 1] = np.einsum("lsr,rk->lsk", mps[j - 1], R.T)
        mps[j - 1] = mps[j - 1] / np.linalg.norm(mps[j - 1])


def left_orthonormal_error(A):
    dl, d, dr = A.shape
    M = np.reshape(A, (dl * d, dr), order="F")
    return np.abs(M.T @ M - np.eye(dr)).max()


def right_orthonormal_error(A):
    dl, d, dr = A.shape
    M = np.reshape(A, (dl, d * dr), order="F")
    return np.abs(M @ M.T - np.eye(dl)).max()


# --------------------------------------------------------------- ProjMPO
def update_left(L, A, W):
    """L_new[r', wr, r] = sum L[l', wl, l] A[l,s,r] W[wl,wr,s',s] A[l',s',r']."""
    T = np.einsum("pwl,lsr->pwsr", L, A)
    T = np.einsum("pwsr,wvts->ptrv", T, W)
    return np.einsum("ptrv,ptq->qvr", T, A)


def update_right(R, A, W):
    """R_new[l', wl, l] = sum R[r', wr, r] A[l,s,r] W[wl,wr,s',s] A[l',s',r']."""
    T = np.einsum("qvr,lsr->qvls", R, A)
    T = np.einsum("qvls,wvts->qltw", T, W)
    return np.einsum("qltw,ptq->pwl", T, A)


def heff_apply(L, W1, W2, R, psi):
    """Fixed order L.psi -> .W1 -> .W2 -> .R (ProjMPO product)."""
    T = np.einsum("pwa,asub->pwsub", L, psi)
    T = np.einsum("pwsub,wmts->ptumb", T, W1)
    T = np.einsum("ptumb,mvzu->ptzvb", T, W2)
    return np.einsum("ptzvb,qvb->ptzq", T, R)


def heff_apply_blas(L, W1, W2, R, psi):
    """heff_apply in the same contraction order with every step a BLAS
    matrix product on reshaped column-major views (the CPU baseline's
    form; equal to heff_apply to rounding)."""
    chiL, w, _ = L.shape
    _, d1, d2, chiR = psi.shape
    wr = W2.shape[1]
    T1 = np.reshape(L, (chiL * w, chiL), order="F") @ np.reshape(psi, (chiL, -1), order="F")
    # X[(wl s1 s2), (s1' s2' wr)] = sum_wm W1[wl, wm, s1', s1] W2[wm, wr, s2', s2]
    X = np.einsum("wmts,mvzu->wsutzv", W1, W2).reshape(w * d1 * d2, d1 * d2 * wr, order="F")
    T1 = np.reshape(T1, (chiL, w * d1 * d2, chiR), order="F")
    T3 = np.matmul(X.T[None, :, :], np.moveaxis(T1, 2, 0).transpose(0, 2, 1))  # (b, o, a')
    T3 = np.transpose(T3, (2, 1, 0))                                            # (a', o, b)
    T3 = np.reshape(np.reshape(T3, (chiL, d1 * d2, wr, chiR), order="F"),
                    (chiL * d1 * d2, wr * chiR), order="F")
    out = T3 @ np.reshape(R, (chiR, wr * chiR), order="F").T
    return np.reshape(out, (chiL, d1, d2, chiR), order="F")


def update_left_blas(L, A, W):
    """update_left with BLAS matrix products (same order)."""
    chiL, w, _ = L.shape
    _, d, c = A.shape
    wr = W.shape[1]
    T = np.reshape(L, (chiL * w, chiL), order="F") @ np.reshape(A, (chiL, d * c), order="F")
    T = np.reshape(T, (chiL, w, d, c), order="F")                  # (l', wl, s, r)
    T = np.einsum("pwsr,wvts->ptrv", T, W, optimize=True)           # (l', s', r, wr)
    T = np.reshape(T, (chiL * d, c * wr), order="F")
    out = np.reshape(A, (chiL * d, c), order="F").T @ T            # (r', (r wr))
    return np.transpose(np.reshape(out, (c, c, wr), order="F"), (0, 2, 1))


def heff_diag(L, W1, W2, R):
    Ld = np.einsum("awa->aw", L)
    Rd = np.einsum("bwb->bw", R)
    W1d = np.einsum("wmss->wms", W1)
    W2d = np.einsum("mvuu->mvu", W2)
    return np.einsum("aw,wms,mvu,bv->asub", Ld, W1d, W2d, Rd)


# -------------------------------------------------------------- davidson
def davidson(matvec, phi0, maxiter=2, miniter=1, errgoal=1e-14, northo=2,
             precond=None):
    """Returns (lambda, phi, n_matvec). precond(q, lam) -> correction."""
    phi = phi0 / np.linalg.norm(phi0)
    V = [phi]
    AV = [matvec(phi)]
    nmv = 1
    lam = float(np.vdot(V[0], AV[0]))
    q = AV[0] - lam * V[0]
    M = np.array([[lam]])
    last = np.nan
    maxiter = min(maxiter, phi.size - 1)
    for ni in range(1, maxiter + 1):
        qn = np.linalg.norm(q)
        conv = (qn < errgoal and abs(lam - last) < errgoal) or qn < max(1e-12, errgoal * 1e-3)
        if qn < 1e-20 or (conv and ni > miniter):
            break
        last = lam
        t = precond(q, lam) if precond is not None else q
        for _ in range(northo):
            for v in V:
                t = t - np.vdot(v, t) * v
        tn = np.linalg.norm(t)
        if tn < 1e-14 * max(1.0, qn):
            break
        V.append(t / tn)
        AV.append(matvec(V[-1]))
        nmv += 1
        m = len(V)
        Mn = np.zeros((m, m))
        Mn[: m - 1, : m - 1] = M
        for k in range(m):
            Mn[k, m - 1] = Mn[m - 1, k] = float(np.vdot(V[k], AV[m - 1]))
        M = Mn
        ev, U = np.linalg.eigh(M)
        lam = float(ev[0])
        u = U[:, 0]
        if u[0] < 0:
            u = -u
        phi = sum(u[k] * V[k] for k in range(m))
        q = sum(u[k] * AV[k] for k in range(m)) - lam * phi
    return lam, phi / np.linalg.norm(phi), nmv


# ---------------------------------------------------------- noise term
def noiseterm(L, W1, W2, R, phi, ortho):
    """Density-matrix perturbation of two-site dmrg with noise (ITensors
    0.2 `noiseterm(PH, phi, ortho)`, scaled by `noise!` of
    examples/dmrg.jl:24 / test/dmrg.jl:24): the kept side's half of the
    projected operator applied to phi, MPO link left open, traced over the
    other side. left: indices (a, s1); right: (s2, b), s2 fastest."""
    dl, d1, d2, dr = phi.shape
    if ortho == "left":
        nt = np.einsum("pwa,wvts,asub->ptvub", L, W1, phi)
        N = np.reshape(nt, (dl * d1, -1), order="F")
        return N @ N.T
    nt = np.einsum("asub,mvzu,qvb->asmzq", phi, W2, R)
    N = np.reshape(nt, (-1, d2 * dr), order="F")
    return N.T @ N


# ---------------------------------------------------------- factorisation
def factorize(phi, ortho, maxdim, mindim, cutoff, drho=None):
    """replacebond! core: returns (A_b, A_{b+1}, truncerr, k). A density-
    matrix perturbation `drho` (already scaled by the noise) forces the eigen
    route on rho + drho (factorize's eigen_perturbation)."""
    dl, d1, d2, dr = phi.shape
    P = np.reshape(phi, (dl * d1, d2 * dr), order="F")
    use_eigen = (cutoff is not None and cutoff >= 1e-12) or drho is not None
    if use_eigen:
        if ortho == "left":
            rho = P @ P.T if drho is None else P @ P.T + drho
            w, U, terr, _ = eigh_trunc(rho, maxdim=maxdim, cutoff=cutoff)
            k = U.shape[1]
            A = np.reshape(U, (dl, d1, k), order="F")
            B = np.reshape(U.T @ P, (k, d2, dr), order="F")
            B = B / np.linalg.norm(B)
        else:
            rho = P.T @ P if drho is None else P.T @ P + drho
            w, U, terr, _ = eigh_trunc(rho, maxdim=maxdim, cutoff=cutoff)
            k = U.shape[1]
            B = np.reshape(U.T, (k, d2, dr), order="F")
            A = np.reshape(P @ U, (dl, d1, k), order="F")
            A = A / np.linalg.norm(A)
    else:
        U, S, V, terr, _ = svd_trunc(P, maxdim=maxdim, mindim=mindim,
                                     cutoff=0.0 if cutoff is None else cutoff)
        k = S.size
        if ortho == "left":
            A = np.reshape(U, (dl, d1, k), order="F")
            B = np.reshape((S[:, None] * V.T), (k, d2, dr), order="F")
            B = B / np.linalg.norm(B)
        else:
            B = np.reshape(V.T, (k, d2, dr), order="F")
            A = np.reshape(U * S[None, :], (dl, d1, k), order="F")
            A = A / np.linalg.norm(A)
    return A, B, terr, k


# ------------------------------------------------------------------ sweeps
def sweepnext(N):
    for b in range(N - 1):
        yield b, 1
    for b in range(N - 2, -1, -1):
        yield b, 2


def dmrg(H, psi, maxdims, cutoffs, mindims=None, krylovdim=3, precond=False,
         errgoal=1e-14, record=None, noises=None):
    """Two-site DMRG. H: list of MPO tensors, psi: right-orthogonal MPS with
    its norm on site 0 (modified in place). Returns the final energy.
    noises: per-sweep noise of the density-matrix perturbation (0 = none)."""
    N = len(psi)
    nsw = len(maxdims)
    mindims = mindims or [1] * nsw
    noises = noises or [0.0] * nsw
    one = np.ones((1, 1, 1))
    Ls = [one] + [None] * N
    Rs = [None] * N + [one]
    for j in range(N - 1, 1, -1):            # R[j] covers sites j..N-1
        Rs[j] = update_right(Rs[j + 1], psi[j], H[j])
    energy = 0.0
    for sw in range(nsw):
        maxerr = 0.0
        for b, ha in sweepnext(N):
            L, R = Ls[b], Rs[b + 2]
            W1, W2 = H[b], H[b + 1]
            phi = np.einsum("lsr,rtq->lstq", psi[b], psi[b + 1])

            def mv(x):
                return heff_apply(L, W1, W2, R, x)

            pre = None
            if precond:
                dg = heff_diag(L, W1, W2, R)

                def pre(q, lam):
                    den = dg - lam
                    den = np.where(np.abs(den) < 1e-8, 1e-8, den)
                    return q / den
            energy, phi, _ = davidson(mv, phi, maxiter=krylovdim - 1, errgoal=errgoal,
                                      precond=pre)
            ortho = "left" if ha == 1 else "right"
            drho = None
            if noises[sw] > 0.0:
                drho = noises[sw] * noiseterm(L, W1, W2, R, phi, ortho)
            A, B, terr, k = factorize(phi, ortho, maxdims[sw], mindims[sw], cutoffs[sw], drho)
            psi[b], psi[b + 1] = A, B
            maxerr = max(maxerr, terr)
            if ha == 1:
                Ls[b + 1] = update_left(Ls[b], psi[b], H[b])
            else:
                Rs[b + 1] = update_right(Rs[b + 2], psi[b + 1], H[b + 1])
        if record is not None:
            record.append({"sweep": sw + 1, "energy": energy, "maxerr": maxerr,
                           "maxlinkdim": max(a.shape[2] for a in psi)})
    return energy


def mps_energy(H, psi):
    """<psi|H|psi>/<psi|psi> by a full left-to-right transfer."""
    L = np.ones((1, 1, 1))
    Nn = np.ones((1, 1))
    for A, W in zip(psi, H):
        L = update_left(L, A, W)
        Nn = np.einsum("pl,lsr,psq->qr", Nn, A, A)
    return float(L[0, 0, 0] / Nn[0, 0])
