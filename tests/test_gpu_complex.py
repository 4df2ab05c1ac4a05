"""GPU parity of ComplexF64 storage through the C ABI: the Philox stream,
data movement, BLAS-1, contraction, Hermitian eigendecomposition and
Davidson, against numpy/LAPACK (zheevd) on the same seeded inputs. The
reference parameterises the same tests over ComplexF64
(test/test_cuitensor.jl:10, test/test_cucontract.jl:9,
test/test_cuiterativesolvers.jl:13-27)."""
import math

import numpy as np
import pytest

from oracle import linalg as olin
from oracle import philox
from oracle import tensor_ops as ops

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import itg
    from itg import ops as d
    itg._lib.ctx()
    return d


def up(dev, a):
    return dev.to_device(ops.to_flat(np.asarray(a, dtype=np.complex128)), "complex128")


def crandn(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


# ------------------------------------------------------------ stream, moves
def test_randn_complex_matches_host_stream(dev):
    for seed, off, n in [(0, 0, 1000), (7, 1001, 4097), (2 ** 40 + 3, 5, 333)]:
        x = dev.to_host(dev.randn(n, seed, off, "complex128"))
        y = philox.randn_complex(n, seed, off)
        assert x.dtype == np.complex128
        np.testing.assert_allclose(x, y, rtol=1e-13, atol=1e-13)


def test_fill_convert_roundtrip(dev):
    rng = np.random.default_rng(1)
    z = crandn(rng, 777)
    d = dev.to_device(z, "complex128")
    assert np.array_equal(dev.to_host(d), z)
    dev.fill(d, 2.5)
    assert np.array_equal(dev.to_host(d), np.full(777, 2.5 + 0j))
    x = rng.standard_normal(777)
    c = dev.convert(dev.to_device(x), "complex128")
    assert np.array_equal(dev.to_host(c), x.astype(np.complex128))
    r = dev.convert(dev.to_device(z, "complex128"), "float64")
    assert np.array_equal(dev.to_host(r), z.real)


@pytest.mark.parametrize("dims,perm", [((2, 3, 4), (2, 0, 1)), ((17, 33, 5), (1, 0, 2)),
                                       ((64, 48), (1, 0)), ((3, 1, 5, 7), (3, 2, 1, 0)),
                                       ((2,) * 14, tuple(np.random.default_rng(5).permutation(14)))])
def test_permute_bit_exact(dev, dims, perm):
    """permute(cuITensor, ...) on complex storage, bit-exact (the large case is
    test_cuitensor.jl's 14 indices of dimension 2)."""
    rng = np.random.default_rng(len(dims))
    A = crandn(rng, *dims)
    B, dB = dev.permute(up(dev, A), dims, perm)
    assert dB == tuple(dims[p] for p in perm)
    assert np.array_equal(dev.to_host(B), ops.to_flat(np.transpose(A, perm)))


def test_add_permuted_and_blas1(dev):
    rng = np.random.default_rng(3)
    A, B = crandn(rng, 6, 7, 8), crandn(rng, 8, 6, 7)
    dB = up(dev, B)
    dev.add_permuted(up(dev, A), (6, 7, 8), (2, 0, 1), -1.5, dB, 2.0)
    np.testing.assert_allclose(ops.from_flat(dev.to_host(dB), (8, 6, 7)),
                               2.0 * B - 1.5 * np.transpose(A, (2, 0, 1)), rtol=1e-15, atol=1e-14)
    for n in [1, 255, 257, 100003]:
        x, y = crandn(rng, n), crandn(rng, n)
        dx, dy = dev.to_device(x, "complex128"), dev.to_device(y, "complex128")
        got = dev.dot(dx, dy)
        want = np.vdot(x, y)
        assert isinstance(got, complex)
        assert abs(got - want) <= 1e-12 * max(1.0, abs(want)) * math.sqrt(n)
        assert math.isclose(dev.nrm2(dx), float(np.linalg.norm(x)), rel_tol=1e-13)
        a, b = 0.3 - 1.2j, -0.7 + 0.25j
        dev.axpby(a, dx, b, dy)
        np.testing.assert_allclose(dev.to_host(dy), a * x + b * y, rtol=1e-14, atol=1e-14)
        dev.axpby(0.5, dx, -2.0, dy)   # real coefficients
        np.testing.assert_allclose(dev.to_host(dy), 0.5 * x - 2.0 * (a * x + b * y),
                                   rtol=1e-14, atol=1e-13)
        dev.scal(dx, 2.0 - 3.0j)
        np.testing.assert_allclose(dev.to_host(dx), (2.0 - 3.0j) * x, rtol=1e-15)
        dev.conj(dx)
        np.testing.assert_allclose(dev.to_host(dx), np.conj((2.0 - 3.0j) * x), rtol=1e-15)


# -------------------------------------------------------------- contraction
CASES = [
    # (dims A, labels A, dims B, labels B, labels C): matrix product, ITensor
    # style 3/4-index contractions with permuted outputs, outer product,
    # full contraction to a scalar, a contraction leaving one side empty
    ((30, 40), (1, 2), (40, 50), (2, 3), (1, 3)),
    ((8, 9, 10), (1, 2, 3), (10, 9, 11), (3, 2, 4), (4, 1)),
    ((5, 6, 7, 8), (1, 2, 3, 4), (8, 7, 9), (4, 3, 5), (5, 2, 1)),
    ((12, 13), (1, 2), (14,), (3,), (3, 1, 2)),
    ((6, 7, 8), (1, 2, 3), (8, 6, 7), (3, 1, 2), ()),
    ((33,), (1,), (33, 65), (1, 2), (2,)),
    ((64, 5, 64), (1, 2, 3), (64, 2, 2, 48), (3, 4, 5, 6), (1, 2, 4, 5, 6)),
]


@pytest.mark.parametrize("case", CASES)
def test_contract_vs_oracle(dev, case):
    dA, lA, dB, lB, lC = case
    rng = np.random.default_rng(sum(dA) + sum(dB))
    A, B = crandn(rng, *dA), crandn(rng, *dB)
    want = ops.contract(A, list(lA), B, list(lB), list(lC))
    C, dC = dev.contract(up(dev, A), dA, lA, up(dev, B), dB, lB, lC)
    got = ops.from_flat(dev.to_host(C), dC) if lC else dev.to_host(C)[0]
    assert rel(np.asarray(got), np.asarray(want)) < 1e-13
    # alpha, beta on a complex C
    C0 = crandn(rng, *dC) if lC else crandn(rng, 1)[0]
    dC0 = up(dev, C0) if lC else dev.to_device(np.array([C0]), "complex128")
    dev.contract(up(dev, A), dA, lA, up(dev, B), dB, lB, lC, alpha=-0.5, beta=2.0, out=dC0)
    got = ops.from_flat(dev.to_host(dC0), dC) if lC else dev.to_host(dC0)[0]
    assert rel(np.asarray(got), -0.5 * np.asarray(want) + 2.0 * C0) < 1e-13


def test_contract_real_modes_rejected(dev):
    from itg._lib import ITGError
    A = up(dev, np.ones((4, 4)))
    with pytest.raises(ITGError):
        dev.contract(A, (4, 4), (1, 2), A, (4, 4), (2, 3), (1, 3), mode="bf16x3")


# ---------------------------------------------------------- Hermitian eigen
def hermitian(rng, n):
    X = crandn(rng, n, n)
    return (X + X.conj().T) / 2


def check_eigh(dev, H, maxdim=None, cutoff=-1.0, tol=1e-12):
    n = H.shape[0]
    w, U, k, te, dc = dev.eigh_trunc(up(dev, H), n, maxdim=maxdim, cutoff=cutoff)
    wr, Ur, ter, dcr = olin.eigh_trunc(H, maxdim=maxdim, cutoff=cutoff)
    assert k == Ur.shape[1]
    scale = max(abs(wr).max(), 1e-300) if k else 1.0
    wh = dev.to_host(w)[:k]
    assert np.abs(wh - wr).max() <= tol * scale
    assert abs(te - ter) <= n * 1e-15 * scale + 1e-12 * ter
    Uh = ops.from_flat(dev.to_host(U)[: n * k], (n, k))
    assert Uh.dtype == np.complex128
    assert np.abs(Uh.conj().T @ Uh - np.eye(k)).max() < 1e-12
    assert np.abs(H @ Uh - Uh * wh).max() < tol * scale * max(1.0, math.sqrt(n) / 8)
    return wh, Uh


@pytest.mark.parametrize("n", [1, 2, 3, 31, 64, 100, 257, 600])
def test_eigh_hermitian_full_spectrum(dev, n):
    check_eigh(dev, hermitian(np.random.default_rng(n), n))


@pytest.mark.parametrize("n,cols,cutoff", [(200, 80, 1e-10), (512, 512, 1e-12)])
def test_eigh_hermitian_density_matrix(dev, n, cols, cutoff):
    """rho = M M^H / |M|^2 (a complex density matrix, a (n - cols)-fold
    cluster at zero for cols < n), truncated as heevd + truncate! decide."""
    rng = np.random.default_rng(n + cols)
    M = crandn(rng, n, cols)
    M /= np.linalg.norm(M)
    check_eigh(dev, M @ M.conj().T, maxdim=n // 2, cutoff=cutoff, tol=1e-13)


def test_eigh_hermitian_degenerate(dev):
    """Exact multiplets (multiplicities 1..5) under a random unitary: every
    eigenvalue, and each returned vector an eigenvector of its value."""
    rng = np.random.default_rng(9)
    vals = np.repeat(np.arange(1.0, 6.0), np.arange(1, 6))   # 15 values
    vals = np.concatenate([vals, rng.uniform(-3, 0, 85)])
    n = vals.size
    Q, _ = np.linalg.qr(crandn(rng, n, n))
    H = (Q * vals) @ Q.conj().T
    H = (H + H.conj().T) / 2
    check_eigh(dev, H)


# ----------------------------------------------------------------- ITensors
def test_itensor_norm_dag_scalar(dev):
    """norm(A) == sqrt(real(scalar(dag(A) * A))) (test_cuitensor.jl) and the
    complex(T) conversion."""
    from itg.itensor import Index, random_roc_itensor
    i, j, k = Index(2, "i"), Index(3, "j"), Index(4, "k")
    A = random_roc_itensor(i, j, k, seed=11, dtype="complex128")
    s = (A.dag() * A).scalar()
    assert isinstance(s, complex) and abs(s.imag) < 1e-12 * abs(s.real)
    assert math.isclose(A.norm(), math.sqrt(s.real), rel_tol=1e-13)
    R = random_roc_itensor(i, j, seed=12)
    Rc = R.complex()
    assert np.array_equal(Rc.array(), R.array().astype(np.complex128))
    B = A * (1.0 - 2.0j)
    np.testing.assert_allclose(B.array(), (1.0 - 2.0j) * A.array(), rtol=1e-15)


def test_complex_davidson(dev):
    """test/test_cuiterativesolvers.jl:13-27: a Hermitian map M(v) = noprime(H v)
    with H = X X^H on (i', i), X random complex; davidson (maxiter 10) from a
    real start on complex storage and from a complex start: M(v) = lambda v
    and lambda the lowest eigenvalue (LAPACK)."""
    from itg.itensor import Index, davidson, itensor
    d = 10
    i = Index(d, "i")
    ip = i.prime()
    rng = np.random.default_rng(13)
    X = crandn(rng, d, d)
    Hm = X @ X.conj().T
    H = itensor(Hm, ip, i, dtype="complex128").roc()

    def M(v):
        return (H * v).noprime()

    w0 = np.linalg.eigvalsh(Hm)[0]
    for x0 in (rng.standard_normal(d), crandn(rng, d)):
        v0 = itensor(x0, i, dtype="complex128").roc()
        lam, v = davidson(M, v0, maxiter=10)
        np.testing.assert_allclose(M(v).array(i), lam * v.array(i), rtol=1e-8,
                                   atol=1e-8 * abs(lam))
        assert abs(v.norm() - 1.0) < 1e-12
        assert abs(lam - w0) < 1e-8 * max(1.0, abs(w0))


def test_factorize_eigen_route_complex(dev):
    """factorize's density-matrix route on complex storage: L orthonormal
    (L^H L = I), L * R reproduces A when nothing is truncated, and the kept
    weights are LAPACK's eigenvalues of A A^H."""
    from itg.itensor import Index, factorize, itensor
    a, b, c = Index(6, "a"), Index(5, "b"), Index(7, "c")
    rng = np.random.default_rng(21)
    X = crandn(rng, 6, 5, 7)
    A = itensor(X, a, b, c, dtype="complex128").roc()
    L, R, spec = factorize(A, [a, b], ortho="left", which_decomp="eigen", cutoff=1e-14)
    Lm = L.array()
    Lm = Lm.reshape(30, -1, order="F")
    assert np.abs(Lm.conj().T @ Lm - np.eye(Lm.shape[1])).max() < 1e-12
    Xm = X.reshape(30, 7, order="F")
    LR = (L * R).array(a, b, c)
    assert rel(LR, X) < 1e-12
    w = np.sort(np.linalg.eigvalsh(Xm @ Xm.conj().T))[::-1][: Lm.shape[1]]
    np.testing.assert_allclose(spec.eigs, w, rtol=1e-10, atol=1e-13 * w[0])


# --------------------------------------------------------------- complex dmrg
def test_complex_dmrg_heisenberg_matches_real_and_ed(dev):
    """The Heisenberg chain (N = 10) on ComplexF64 storage from the complexified
    random MPS: ends at the real run's energy and the exact ground energy."""
    from oracle import models
    from itg.mps import MPO, Sweeps, dmrg, expect_energy, heisenberg_opsum, random_mps
    N = 10
    bonds = [(j, j + 1) for j in range(N - 1)]
    H = MPO.from_opsum(heisenberg_opsum(bonds), N)
    sw = Sweeps(6, maxdim=[8, 16, 32, 64], cutoff=[1e-12])
    Er = dmrg(H, random_mps(N, 2, 4, seed=3), sw)
    psi = random_mps(N, 2, 4, seed=3).complex()
    Ec = dmrg(H, psi, sw)
    E0 = models.ground_energy(models.bond_hamiltonian(N, bonds))
    assert psi.is_complex()
    assert abs(Ec - Er) < 1e-10
    assert E0 - 1e-10 <= Ec < E0 + 1e-9
    assert abs(expect_energy(H, psi) - Ec) < 1e-10


def test_complex_dmrg_dm_interaction_vs_ed(dev):
    """A genuinely complex Hamiltonian: Heisenberg plus a Dzyaloshinskii-Moriya
    term D (Sx_i Sy_j - Sy_i Sx_j) = (iD/2)(S+_i S-_j - S-_i S+_j) on the
    N = 8 chain, D = 0.4, with the density-matrix noise term in the first
    sweeps: the energy of the exact diagonalisation of the same MPO."""
    from oracle import models
    from itg.mps import MPO, OpSum, Sweeps, dmrg, expect_energy, random_mps
    N, D = 8, 0.4
    os_ = OpSum()
    for j in range(N - 1):
        os_.add(1.0, "Sz", j, "Sz", j + 1)
        os_.add(0.5, "S+", j, "S-", j + 1)
        os_.add(0.5, "S-", j, "S+", j + 1)
        os_.add(0.5j * D, "S+", j, "S-", j + 1)
        os_.add(-0.5j * D, "S-", j, "S+", j + 1)
    Ws = os_.to_mpo_arrays(N)
    assert np.iscomplexobj(Ws[1])
    Hm = models.mpo_to_matrix(Ws)
    assert np.abs(Hm - Hm.conj().T).max() < 1e-14
    E0 = float(np.linalg.eigvalsh(Hm)[0])
    H = MPO(Ws)
    assert H.is_complex()
    sw = Sweeps(8, maxdim=[8, 16, 32], cutoff=[1e-12], noise=[1e-6, 1e-7, 1e-8, 0.0])
    psi = random_mps(N, 2, 4, seed=5)
    E = dmrg(H, psi, sw)
    assert psi.is_complex()
    assert abs(E - E0) < 1e-9
    assert abs(expect_energy(H, psi) - E) < 1e-10


# ----------------------------------------------------------------- QR, SVD
@pytest.mark.parametrize("m,n", [(20, 12), (12, 20), (64, 64), (300, 130), (1, 1)])
def test_qr_complex(dev, m, n):
    """Complex thin QR (the reference's qr over ComplexF64 storage): A = QR,
    Q^H Q = I, R upper triangular with LAPACK's |diag R|."""
    rng = np.random.default_rng(m * 7 + n)
    A = crandn(rng, m, n)
    Q, R = dev.qr(up(dev, A), m, n)
    r = min(m, n)
    Qh = ops.from_flat(dev.to_host(Q), (m, r))
    Rh = ops.from_flat(dev.to_host(R), (r, n))
    np.testing.assert_allclose(Qh @ Rh, A, atol=1e-13 * max(1.0, np.abs(A).max()))
    assert np.abs(Qh.conj().T @ Qh - np.eye(r)).max() < 1e-13
    assert np.abs(np.tril(Rh, -1)).max() == 0.0
    np.testing.assert_allclose(np.abs(np.diag(Rh)), np.abs(np.diag(np.linalg.qr(A)[1])),
                               rtol=1e-12)


@pytest.mark.parametrize("m,n", [(50, 30), (30, 50), (200, 200), (300, 120)])
def test_svd_complex(dev, m, n):
    """Complex truncated SVD (A = U S V^H): singular values against LAPACK's
    zgesdd, unitary factors, reconstruction; then a graded spectrum
    truncated by cutoff as truncate! decides on LAPACK's values."""
    from oracle.truncate import ndtensors_truncate
    rng = np.random.default_rng(m + 3 * n)
    A = crandn(rng, m, n)
    U, S, V, k, te, dc = dev.svd_trunc(up(dev, A), m, n)
    r = min(m, n)
    assert k == r
    Uh = ops.from_flat(dev.to_host(U)[: m * k], (m, k))
    Vh = ops.from_flat(dev.to_host(V)[: n * k], (n, k))
    Sh = dev.to_host(S)[:k]
    np.testing.assert_allclose(Sh, np.linalg.svd(A, compute_uv=False), rtol=1e-12)
    assert np.abs(Uh.conj().T @ Uh - np.eye(k)).max() < 1e-12
    assert np.abs(Vh.conj().T @ Vh - np.eye(k)).max() < 1e-12
    np.testing.assert_allclose((Uh * Sh) @ Vh.conj().T, A, atol=1e-12 * np.abs(A).max())
    U0, _ = np.linalg.qr(crandn(rng, m, r))
    V0, _ = np.linalg.qr(crandn(rng, n, r))
    s = 10.0 ** -np.linspace(0, 8, r)
    B = (U0 * s) @ V0.conj().T
    U, S, V, k, te, dc = dev.svd_trunc(up(dev, B), m, n, maxdim=r - 3, cutoff=1e-12)
    kr, ter, _ = ndtensors_truncate(s ** 2, maxdim=r - 3, cutoff=1e-12)
    assert k == kr
    assert abs(te - ter) <= 1e-6 * ter + 1e-24
    np.testing.assert_allclose(dev.to_host(S)[:k], s[:k], rtol=1e-9)


def test_complex_mps_orthogonalize_and_svd_route(dev):
    """A complex random MPS orthogonalised by complex QR keeps its norm and
    inner products; complex DMRG through the SVD route (cutoff below 1e-12)
    reaches the same energy as the density-matrix route."""
    from itg import mpo
    from itg.itensor import Index
    from itg.mps import MPO, OpSum, Sweeps, dmrg, random_mps
    sites = [Index(2, "Site,n=%d" % j) for j in range(6)]
    psi = mpo.random_mps(sites, linkdims=4, seed=3, dtype="complex128")
    assert psi[0].is_complex()
    assert abs(mpo.inner(psi, psi) - 1.0) < 1e-12
    phi = psi.copy()
    mpo.orthogonalize(phi, 3)
    assert abs(mpo.inner(psi, phi) - 1.0) < 1e-12
    N, D = 8, 0.4
    os_ = OpSum()
    for j in range(N - 1):
        os_.add(1.0, "Sz", j, "Sz", j + 1)
        os_.add(0.5, "S+", j, "S-", j + 1)
        os_.add(0.5, "S-", j, "S+", j + 1)
        os_.add(0.5j * D, "S+", j, "S-", j + 1)
        os_.add(-0.5j * D, "S-", j, "S+", j + 1)
    H = MPO(os_.to_mpo_arrays(N))
    E_eig = dmrg(H, random_mps(N, 2, 4, seed=5), Sweeps(6, maxdim=[8, 16, 32], cutoff=[1e-12]))
    E_svd = dmrg(H, random_mps(N, 2, 4, seed=5), Sweeps(6, maxdim=[8, 16, 32], cutoff=[1e-14]))
    assert abs(E_eig - E_svd) < 1e-9
