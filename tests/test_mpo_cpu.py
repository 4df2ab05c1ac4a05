"""CPU tier: the host-side OpSum -> MPO compiler of the package against the
oracle's exact-diagonalisation Hamiltonians (no GPU: MPO construction is
host work, as AutoMPO is in ITensors)."""
import numpy as np
import pytest

from oracle import models


def mpo_arrays(os_, N, S=0.5):
    from itg.mps import OpSum  # noqa: F401  (host code only, no device use)
    return os_.to_mpo_arrays(N, S)


@pytest.mark.parametrize("S,N", [(0.5, 6), (1.0, 4)])
def test_heisenberg_chain(S, N):
    from itg.mps import heisenberg_opsum
    Ws = heisenberg_opsum([(j, j + 1) for j in range(N - 1)]).to_mpo_arrays(N, S)
    assert max(W.shape[1] for W in Ws) == 5
    H = models.mpo_to_matrix(Ws)
    Hed = models.bond_hamiltonian(N, [(j, j + 1) for j in range(N - 1)], S=S).toarray()
    np.testing.assert_allclose(H, Hed, atol=1e-13)


def test_tfim():
    from itg.mps import tfim_opsum
    N = 6
    H = models.mpo_to_matrix(tfim_opsum(N).to_mpo_arrays(N))
    np.testing.assert_allclose(H, models.tfim_hamiltonian(N).toarray(), atol=1e-13)


def test_cylinder_small_equals_ed():
    from itg.mps import cylinder_bonds, heisenberg_opsum
    bonds = cylinder_bonds(2, 3)
    N = 6
    H = models.mpo_to_matrix(heisenberg_opsum(bonds).to_mpo_arrays(N))
    np.testing.assert_allclose(H, models.bond_hamiltonian(N, bonds).toarray(), atol=1e-13)


def test_cylinder_4x16_bond_dimension():
    """4x16 cylinder, snake order: 4 left sites couple across every cut, so
    w = 2 + 3*4 = 14."""
    from itg.mps import cylinder_bonds, heisenberg_opsum
    bonds = cylinder_bonds(16, 4)
    assert sorted(bonds) == sorted(models.cylinder_bonds(16, 4))
    Ws = heisenberg_opsum(bonds).to_mpo_arrays(64)
    assert max(W.shape[1] for W in Ws) == 14


def test_long_range_terms():
    """Arbitrary-range two-site terms (channels pass through sites)."""
    from itg.mps import OpSum
    N = 5
    os_ = OpSum().add(0.7, "Sz", 0, "Sz", 4).add(-0.3, "S+", 1, "S-", 3).add(0.2, "Sx", 2)
    H = models.mpo_to_matrix(os_.to_mpo_arrays(N))
    Sz, Sp, Sm, I = models.spin_ops(0.5)

    def op(o, j):
        mats = [I] * N
        mats[j] = o
        out = mats[0]
        for m in mats[1:]:
            out = np.kron(out, m)
        return out

    ref = 0.7 * op(Sz, 0) @ op(Sz, 4) - 0.3 * op(Sp, 1) @ op(Sm, 3) + 0.2 * op(0.5 * (Sp + Sm), 2)
    np.testing.assert_allclose(H, ref, atol=1e-14)


def test_complex_dm_term_against_kronecker_products():
    """A complex OpSum (a Dzyaloshinskii-Moriya term written with S+/S- and
    imaginary coefficients, and the same term written with Sx/Sy) compiles to
    a ComplexF64 MPO whose matrix is the Kronecker-product Hamiltonian
    D sum_j (Sx_j Sy_j+1 - Sy_j Sx_j+1), Hermitian and not real."""
    from itg.mps import OpSum
    N, D = 4, 0.7
    sx = np.array([[0, 0.5], [0.5, 0]])
    sy = np.array([[0, -0.5j], [0.5j, 0]])
    I2 = np.eye(2)

    def site_op(op, j):
        out = np.eye(1)
        for k in range(N):
            out = np.kron(out, op if k == j else I2)
        return out

    Hk = sum(D * (site_op(sx, j) @ site_op(sy, j + 1) - site_op(sy, j) @ site_op(sx, j + 1))
             for j in range(N - 1))
    a, b = OpSum(), OpSum()
    for j in range(N - 1):
        a.add(0.5j * D, "S+", j, "S-", j + 1)
        a.add(-0.5j * D, "S-", j, "S+", j + 1)
        b.add(D, "Sx", j, "Sy", j + 1)
        b.add(-D, "Sy", j, "Sx", j + 1)
    for os_ in (a, b):
        Ws = os_.to_mpo_arrays(N)
        assert all(W.dtype == np.complex128 for W in Ws)
        H = models.mpo_to_matrix(Ws)
        np.testing.assert_allclose(H, Hk, atol=1e-14)
    assert np.abs(Hk - Hk.conj().T).max() < 1e-15 and np.abs(Hk.imag).max() > 0.1
