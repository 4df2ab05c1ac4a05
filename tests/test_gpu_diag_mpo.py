"""GPU tests of Diag and Combiner storage, MPS/MPO algebra and gate
application through the C ABI, in the reference's test shapes
(test/test_cudiag.jl, test/test_cumpo.jl, examples/gate_evolution.jl).
Expected values come from numpy on the host: dense diagonal matrices for
Diag, and, for the MPS/MPO algebra, the full state vectors and operator
matrices contracted from the same site tensors."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def it():
    import itg
    from itg import itensor
    itg._lib.ctx()
    return itensor


def rnd(rng, dtype, *shape):
    x = rng.standard_normal(shape)
    if dtype == "complex128":
        x = x + 1j * rng.standard_normal(shape)
    return x


# ------------------------------------------------------------------- Diag
@pytest.mark.parametrize("T1,T2", [("float64", "float64"), ("complex128", "complex128"),
                                   ("float64", "complex128"), ("complex128", "float64")])
def test_diag_contractions(it, T1, T2):
    """test_cudiag.jl: Matrix*Diag -> Matrix, Diag*Diag -> Diag,
    UniformDiag*Diag -> Diag (both orders), Matrix*UniformDiag -> Matrix
    (both orders, the second following the index labels where the
    reference's A.*B is @test_broken), for Float64/ComplexF64 and mixed."""
    rng = np.random.default_rng(5)
    i, j = it.Index(2, "i"), it.Index(3, "j")
    ip, ipp = i.prime(), i.prime(2)
    A = rnd(rng, T1, 2, 3)
    Aij = it.itensor(A, i, j, dtype=T1).roc()
    Dv, Ev = rnd(rng, T2, 2), rnd(rng, T2, 2)
    D = it.diag_itensor(Dv, i, ip)
    E = it.diag_itensor(Ev, i, ipp)
    C = Aij * D
    np.testing.assert_allclose(C.array(j, ip), A.T @ np.diag(Dv), rtol=1e-14, atol=1e-15)
    C = E * D
    assert isinstance(C, it.DiagITensor)
    np.testing.assert_allclose(C.array(ipp, ip), np.diag(Ev) @ np.diag(Dv), rtol=1e-14)
    scal = it.diag_itensor(2.0, i, ipp)
    for C in (scal * D, D * scal):
        np.testing.assert_allclose(C.array(ipp, ip), 2.0 * np.diag(Dv), rtol=1e-15)
    scal = it.diag_itensor(2.0, i, ip)
    C = scal * Aij
    np.testing.assert_allclose(C.array(ip, j), 2.0 * A, rtol=1e-15)
    C = Aij * scal
    np.testing.assert_allclose(C.array(j, ip), 2.0 * A.T, rtol=1e-15)


def test_diag_dense_trace_and_rank3(it):
    """Every Diag label contracted (a weighted trace) and a rank-3 Diag
    (a delta tensor with weights) against einsum on the dense diagonal."""
    rng = np.random.default_rng(6)
    i, k = it.Index(4, "i"), it.Index(5, "k")
    ip = i.prime()
    dv = rng.standard_normal(4)
    D = it.diag_itensor(dv, i, ip)
    A = rng.standard_normal((4, 4, 5))
    At = it.itensor(A, i, ip, k, dtype="float64").roc()
    C = D * At
    np.testing.assert_allclose(C.array(k), np.einsum("i,iik->k", dv, A), rtol=1e-13)
    a, b, c = it.Index(3, "a"), it.Index(3, "b"), it.Index(3, "c")
    wv = rng.standard_normal(3)
    W = it.diag_itensor(wv, a, b, c)
    dense = np.zeros((3, 3, 3))
    for q in range(3):
        dense[q, q, q] = wv[q]
    np.testing.assert_allclose(W.dense().array(a, b, c), dense, rtol=1e-15)
    X = rng.standard_normal((3, 6))
    m = it.Index(6, "m")
    Xt = it.itensor(X, b, m, dtype="float64").roc()
    np.testing.assert_allclose((W * Xt).array(a, c, m), np.einsum("abc,bm->acm", dense, X),
                               rtol=1e-14, atol=1e-15)


def test_combiner_roundtrip(it):
    """combiner(i, k): T(i, j, k) * C has the combined index (i fastest) and
    j; contracting C again restores T exactly."""
    rng = np.random.default_rng(7)
    i, j, k = it.Index(2, "i"), it.Index(3, "j"), it.Index(4, "k")
    A = rng.standard_normal((2, 3, 4))
    T = it.itensor(A, i, j, k, dtype="float64").roc()
    Cm = it.combiner(i, k)
    TC = T * Cm
    c = Cm.combinedind()
    assert c.dim == 8 and set(TC.inds) == {c, j}
    np.testing.assert_array_equal(TC.array(c, j),
                                  np.reshape(np.transpose(A, (0, 2, 1)), (8, 3), order="F"))
    back = TC * Cm
    np.testing.assert_array_equal(back.array(i, j, k), A)


# --------------------------------------------------------------- MPS / MPO
def state_vector(psi):
    """Full state (site 0 most significant) contracted on the host."""
    N = len(psi)
    v = None
    for j in range(N):
        s = psi.siteinds(j)[0]
        ls = [psi.linkind(j - 1), psi.linkind(j)]
        order = [x for x in (ls[0], s, ls[1]) if x is not None]
        a = psi[j].array(*order)
        if ls[0] is None:
            a = a[None]
        if ls[1] is None:
            a = a[..., None]
        v = a[0] if v is None else np.einsum("pl,lsr->psr", v, a).reshape(-1, a.shape[2])
    return v[:, 0]


def mpo_matrix(K):
    N = len(K)
    M = None
    for j in range(N):
        s = [x for x in K.siteinds(j) if x.plev == 0][0]
        ls = [K.linkind(j - 1), K.linkind(j)]
        order = [x for x in (ls[0], s.prime(), s, ls[1]) if x is not None]
        a = K[j].array(*order)
        if ls[0] is None:
            a = a[None]
        if ls[1] is None:
            a = a[..., None]
        if M is None:
            M = a[0]
        else:
            M = np.einsum("OIl,loir->OoIir", M, a)
            M = M.reshape(M.shape[0] * M.shape[1], M.shape[2] * M.shape[3], M.shape[4])
    return M[:, :, 0]


def test_mpo_basics_and_inner(it):
    """test_cumpo.jl "CuMPO Basics": length and layout of a random MPO,
    inner(phi, K, psi) against the explicit site-by-site contraction and the
    dense vectors, invariance under orthogonalize! of phi and K."""
    from itg import mpo
    N = 6
    sites = [it.Index(2, "Site,n=%d" % n) for n in range(N)]
    K = mpo.random_mpo(sites, seed=1)
    assert len(K) == N and K.maxlinkdim() == 1
    assert len(repr(K).split("\n")) == N + 2
    phi = mpo.random_mps(sites, seed=2)
    psi = mpo.random_mps(sites, seed=3)
    want = np.vdot(state_vector(phi), mpo_matrix(K) @ state_vector(psi))
    got = mpo.inner(phi, K, psi)
    assert abs(got - want) < 1e-12 * max(1.0, abs(want))
    pd = phi.dag().prime()
    E = pd[0] * K[0] * psi[0]
    for j in range(1, N):
        E = E * (pd[j] * K[j] * psi[j])
    assert abs(E.scalar() - got) < 1e-12 * max(1.0, abs(got))
    orig = mpo.inner(phi, K, phi)
    mpo.orthogonalize(phi, N // 2)
    mpo.orthogonalize(K, N // 2)
    assert abs(mpo.inner(phi, K, phi) - orig) < 1e-12 * max(1.0, abs(orig))
    bad = mpo.random_mps([it.Index(2, "Site") for _ in range(N + 1)], seed=4)
    with pytest.raises(ValueError):
        mpo.inner(phi, K, bad)


@pytest.mark.parametrize("link_dim", [2, 3, 5])
def test_mpo_contract_and_inner_bigger_links(it, link_dim):
    """test_cumpo.jl "inner <y|A|x>" and "contract" with random MPO/MPS of
    link dimension 2..5, orthogonalised with maxdim = link_dim: inner equals
    the explicit contraction; contract(K, psi; maxdim = 10 link_dim,
    cutoff = 0) reproduces K|psi>."""
    from itg import mpo
    N = 6
    rng = np.random.default_rng(link_dim)
    sites = [it.Index(2, "Site,n=%d" % n) for n in range(N)]
    kl = [it.Index(link_dim, "r%d,Link" % n) for n in range(N - 1)]
    pl = [it.Index(link_dim, "r%d,Link" % n) for n in range(N - 1)]
    ql = [it.Index(link_dim, "r%d,Link" % n) for n in range(N - 1)]

    def rand_t(*inds):
        return it.itensor(rng.standard_normal([i.dim for i in inds]), *inds,
                          dtype="float64").roc()

    Kt, Pt, Qt = [], [], []
    for j in range(N):
        lk = ([kl[j - 1]] if j else []) + ([kl[j]] if j < N - 1 else [])
        lp = ([pl[j - 1]] if j else []) + ([pl[j]] if j < N - 1 else [])
        lq = ([ql[j - 1]] if j else []) + ([ql[j]] if j < N - 1 else [])
        Kt.append(rand_t(*(lk + [sites[j], sites[j].prime()])))
        Pt.append(rand_t(*(lp + [sites[j]])))
        Qt.append(rand_t(*(lq + [sites[j]])))
    K, psi, phi = mpo.MPO(Kt), mpo.MPS(Pt), mpo.MPS(Qt)
    mpo.orthogonalize(psi, 0, maxdim=link_dim)
    mpo.orthogonalize(K, 0, maxdim=link_dim)
    mpo.orthogonalize(phi, 0, maxdim=link_dim, normalize=True)
    pd = phi.dag().prime()
    E = pd[0] * K[0] * psi[0]
    for j in range(1, N):
        E = E * pd[j] * K[j] * psi[j]
    want = E.scalar()
    assert abs(mpo.inner(phi, K, psi) - want) < 1e-11 * max(1.0, abs(want))
    out = mpo.contract_mpo(K, psi, maxdim=10 * link_dim, cutoff=0.0)
    assert abs(mpo.inner(phi, out) - want) < 1e-11 * max(1.0, abs(want))
    np.testing.assert_allclose(state_vector(out), mpo_matrix(K) @ state_vector(psi),
                               rtol=1e-10, atol=1e-12 * np.linalg.norm(state_vector(psi)))


def test_mpo_contract_add_and_mpo_mpo(it):
    """test_cumpo.jl "contract" (link dimension 1, maxdim 1, bad method and
    bad sites rejected), "add" (<psi|K psi + L psi> = <psi|K + L|psi>) and
    "contract(::CuMPO, ::CuMPO)" (prime(K) L applied to psi, and K, L on
    differently labelled sites)."""
    from itg import mpo
    from itg._lib import ArgumentError, DimensionMismatch
    N = 6
    sites = [it.Index(2, "Site,n=%d" % n) for n in range(N)]
    phi, psi = mpo.random_mps(sites, seed=5), mpo.random_mps(sites, seed=6)
    K, L = mpo.random_mpo(sites, seed=7), mpo.random_mpo(sites, seed=8)
    out = mpo.contract_mpo(K, psi, maxdim=1)
    assert abs(mpo.inner(phi, out) - mpo.inner(phi, K, psi)) < 1e-12
    with pytest.raises(ArgumentError):
        mpo.contract_mpo(K, psi, method="fakemethod")
    with pytest.raises(DimensionMismatch):
        mpo.contract_mpo(K, mpo.random_mps([it.Index(2, "Site") for _ in range(N + 1)]))
    M = mpo.add(K, L)
    assert len(M) == N and M.maxlinkdim() == 2
    kp, lp = mpo.contract_mpo(K, psi, maxdim=1), mpo.contract_mpo(L, psi, maxdim=1)
    s = mpo.add(kp, lp)
    assert abs(mpo.inner(psi, s) - mpo.inner(psi, M, psi)) < 1e-12
    np.testing.assert_allclose(mpo_matrix(M), mpo_matrix(K) + mpo_matrix(L), rtol=1e-13,
                               atol=1e-14)
    KL = mpo.contract_mpo(K.prime(), L, maxdim=1)
    via = mpo.contract_mpo(K.prime(), mpo.contract_mpo(L, psi, maxdim=1), maxdim=1)
    assert abs(mpo.inner(psi, KL, psi) - mpo.inner(psi, via)) < 1e-12
    # K and L on differently labelled sites
    sk = [it.Index(2, "Site,aaa") for _ in range(N)]
    sl = [it.Index(2, "Site,bbb") for _ in range(N)]
    K2, L2 = mpo.random_mpo(sites, seed=9), mpo.random_mpo(sites, seed=10)
    for j in range(N):
        K2[j] = it.ITensor([sk[j] if x == sites[j].prime() else x for x in K2[j].inds], K2[j].data)
        L2[j] = it.ITensor([sl[j] if x == sites[j].prime() else x for x in L2[j].inds], L2[j].data)
    KL = mpo.contract_mpo(K2, L2, maxdim=1)
    pk, pl_ = mpo.random_mps(sk, seed=11), mpo.random_mps(sl, seed=12)
    via = mpo.contract_mpo(K2, mpo.contract_mpo(L2, pl_, maxdim=1), maxdim=1)
    assert abs(mpo.inner(pk, KL, pl_) - mpo.inner(pk, via)) < 1e-12


def test_gate_evolution(it):
    """examples/gate_evolution.jl: X on every other qubit of |0...0> (N=10)
    gives |1010101010>; and H, CNOT, X, SWAP on N = 6 qubits against the
    dense state vector."""
    from itg import mpo
    N = 10
    s = [it.Index(2, "Qubit,n=%d" % n) for n in range(N)]
    psi0 = mpo.product_mps(s, [0] * N)
    gates = mpo.ops_list(s, [("X", n) for n in range(1, N + 1, 2)])
    psi = mpo.apply(gates, psi0)
    want = np.zeros(2 ** N)
    want[int("10" * (N // 2), 2)] = 1.0
    np.testing.assert_allclose(state_vector(psi), want, atol=1e-15)
    N = 6
    s = [it.Index(2, "Qubit,n=%d" % n) for n in range(N)]
    psi = mpo.product_mps(s, [0] * N)
    seq = [("H", 1), ("CNOT", 1, 2), ("X", 4), ("H", 3), ("SWAP", 3, 4), ("CNOT", 5, 6),
           ("H", 5), ("CNOT", 2, 3)]
    psi = mpo.apply(mpo.ops_list(s, seq), psi)
    v = np.zeros(2 ** N)
    v[0] = 1.0
    I2 = np.eye(2)
    for name, *ns in seq:
        m = mpo.QUBIT_OPS[name] if len(ns) == 1 else mpo.TWO_QUBIT_OPS[name]
        ops_ = [I2] * (ns[0] - 1) + [m] + [I2] * (N - ns[0] - len(ns) + 1)
        full = ops_[0]
        for o in ops_[1:]:
            full = np.kron(full, o)
        v = full @ v
    np.testing.assert_allclose(state_vector(psi), v, atol=1e-14)
