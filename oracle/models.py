"""Closed-form MPOs for the Hamiltonians on the path and exact-diagonalisation
Hamiltonians that pin them.

MPO convention: one array per site W[wl, wr, s_out, s_in]; the first site has
wl = 1 and the last wr = 1, so contracting the chain with 1x1 boundary
environments gives H. Operator matrices are <s_out| O |s_in>.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def spin_ops(S):
    """Sz, S+, S-, Id in the basis m = S, S-1, ..., -S (ITensors "S=1/2"
    Up,Dn and "S=1" Up,Z0,Dn orderings)."""
    d = int(round(2 * S + 1))
    m = S - np.arange(d)
    Sz = np.diag(m)
    Sp = np.zeros((d, d))
    for i in range(1, d):
        # <m+1| S+ |m> = sqrt(S(S+1) - m(m+1)); state i has m = S - i
        mm = m[i]
        Sp[i - 1, i] = np.sqrt(S * (S + 1) - mm * (mm + 1))
    return Sz, Sp, Sp.T.copy(), np.eye(d)


def heisenberg_mpo(N, J=1.0, S=0.5):
    """H = J sum_j Sz_j Sz_{j+1} + J/2 (S+_j S-_{j+1} + S-_j S+_{j+1})
    (the AutoMPO sums of test/dmrg.jl:9-14, examples/dmrg.jl:10-15), w = 5."""
    Sz, Sp, Sm, I = spin_ops(S)
    d = I.shape[0]
    W = np.zeros((5, 5, d, d))
    W[0, 0] = I
    W[1, 0] = Sp
    W[2, 0] = Sm
    W[3, 0] = Sz
    W[4, 1] = 0.5 * J * Sm
    W[4, 2] = 0.5 * J * Sp
    W[4, 3] = J * Sz
    W[4, 4] = I
    return _ends(W, N)


def tfim_mpo(N, J=1.0, h=0.5):
    """H = -J sum Sz_j Sz_{j+1} - h sum Sx_j with Sx = (S+ + S-)/2
    (test/dmrg.jl:64-68 with J=1, h=0.5), w = 3."""
    Sz, Sp, Sm, I = spin_ops(0.5)
    Sx = 0.5 * (Sp + Sm)
    W = np.zeros((3, 3, 2, 2))
    W[0, 0] = I
    W[1, 0] = Sz
    W[2, 0] = -h * Sx
    W[2, 1] = -J * Sz
    W[2, 2] = I
    return _ends(W, N)


def _ends(W, N):
    w = W.shape[0]
    if N == 1:
        return [W[w - 1:w, 0:1].copy()]
    out = [W[w - 1:w, :].copy()]
    out += [W.copy() for _ in range(N - 2)]
    out.append(W[:, 0:1].copy())
    return out


def tfim_exact_energy(N):
    """Closed form at criticality, open chain (test/dmrg.jl:79)."""
    return 0.25 - 0.25 / np.sin(np.pi / (2 * (2 * N + 1)))


def bond_hamiltonian(N, bonds, S=0.5, J=1.0):
    """Sparse Heisenberg Hamiltonian over an arbitrary bond list, Kronecker
    order with site 0 most significant."""
    Sz, Sp, Sm, I = spin_ops(S)
    d = I.shape[0]

    def op_at(op, j):
        return sp.kron(sp.kron(sp.identity(d ** j), sp.csr_matrix(op)),
                       sp.identity(d ** (N - j - 1)), format="csr")

    H = sp.csr_matrix((d ** N, d ** N))
    for (i, j) in bonds:
        H = H + J * (op_at(Sz, i) @ op_at(Sz, j)
                     + 0.5 * (op_at(Sp, i) @ op_at(Sm, j) + op_at(Sm, i) @ op_at(Sp, j)))
    return H


def tfim_hamiltonian(N, J=1.0, h=0.5):
    Sz, Sp, Sm, I = spin_ops(0.5)
    Sx = 0.5 * (Sp + Sm)

    def op_at(op, j):
        return sp.kron(sp.kron(sp.identity(2 ** j), sp.csr_matrix(op)),
                       sp.identity(2 ** (N - j - 1)), format="csr")

    H = sp.csr_matrix((2 ** N, 2 ** N))
    for j in range(N):
        if j < N - 1:
            H = H - J * (op_at(Sz, j) @ op_at(Sz, j + 1))
        H = H - h * op_at(Sx, j)
    return H


def ground_energy(H):
    if H.shape[0] <= 64:
        return float(np.linalg.eigvalsh(H.toarray())[0])
    return float(spla.eigsh(H, k=1, which="SA", tol=1e-13)[0][0])


def mpo_to_matrix(Ws):
    """Contract an MPO chain to the full operator (Kronecker order, site 0
    most significant). Small N only."""
    T = Ws[0][0]  # (wr, s_out, s_in) with wl = 1
    # accumulate M[wr, out, in] with out/in multi-indices, site 0 major
    M = T
    for W in Ws[1:]:
        # M[a, O, I] * W[a, b, o, i] -> [b, O o, I i]
        R = np.einsum("aOI,aboi->bOoIi", M, W)
        b = R.shape[0]
        M = R.reshape(b, R.shape[1] * R.shape[2], R.shape[3] * R.shape[4])
    return M[0]


def cylinder_bonds(Lx, Ly, snake=True):
    """Nearest-neighbour bonds of an Lx x Ly cylinder (periodic in y) with
    sites ordered column by column, alternating direction when `snake`."""
    def site(x, y):
        if snake and x % 2 == 1:
            y = Ly - 1 - y
        return x * Ly + y

    bonds = set()
    for x in range(Lx):
        for y in range(Ly):
            s = site(x, y)
            if x + 1 < Lx:
                bonds.add(tuple(sorted((s, site(x + 1, y)))))
            if Ly > 2 or y + 1 < Ly:
                t = site(x, (y + 1) % Ly)
                if t != s:
                    bonds.add(tuple(sorted((s, t))))
    return sorted(bonds)
