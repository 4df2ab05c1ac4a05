"""GPU parity of the ITensor-level mirror, random MPS, Davidson and two-site
DMRG against the CPU oracle, exact diagonalisation and the reference's own
physics tests (test/dmrg.jl)."""
import math

import numpy as np
import pytest

from oracle import dmrg as odmrg
from oracle import models
from oracle import tensor_ops as ops

pytestmark = pytest.mark.gpu

E0_N20 = -8.682473334398939


@pytest.fixture(scope="module")
def itg():
    import itg as m
    from itg import itensor, mps
    m._lib.ctx()
    return m


# ------------------------------------------------------------ ITensor layer
def test_itensor_roundtrip_and_permute(itg):
    """test/test_cuitensor.jl:17-43 and test_cudense.jl:43-48."""
    from itg.itensor import Index, itensor
    i, j, k = Index(2, "i"), Index(3, "j"), Index(4, "k")
    rng = np.random.default_rng(0)
    a = rng.standard_normal((2, 4, 3))
    T = itensor(a, i, k, j)
    D = T.roc()
    assert np.array_equal(D.cpu().data, T.data)
    P = D.permute(k, j, i)
    assert P.inds == (k, j, i)
    assert np.array_equal(P.array(), np.transpose(a, (1, 2, 0)))
    assert np.array_equal(P.array(i, k, j), a)


def test_itensor_contract_add_norm(itg):
    from itg.itensor import Index, itensor, random_roc_itensor
    i, j, k, l = Index(2, "i"), Index(3, "j"), Index(4, "k"), Index(5, "l")
    A = random_roc_itensor(i, j, k, seed=1)
    B = random_roc_itensor(k, l, j, seed=2)
    C = A * B
    assert set(C.inds) == {i, l} and C.inds == (i, l)
    want = np.einsum("ijk,klj->il", A.array(), B.array())
    np.testing.assert_allclose(C.array(), want, rtol=1e-13)
    # add with permutation (test_cuitensor.jl:89-99)
    A2 = random_roc_itensor(k, i, j, seed=3)
    S = A + A2
    np.testing.assert_allclose(S.array(i, j, k), A.array(i, j, k) + A2.array(i, j, k), rtol=0,
                               atol=1e-15)
    D = A - A2
    np.testing.assert_allclose(D.array(i, j, k), A.array(i, j, k) - A2.array(i, j, k), atol=1e-15)
    # norm(A) = sqrt(scalar(A*A)) (test_cuitensor.jl:74-78)
    assert math.isclose(A.norm(), math.sqrt((A * A).scalar()), rel_tol=1e-13)


def test_itensor_factorizations(itg):
    """test/test_cuitensor.jl:101-130 at the ITensor level."""
    from itg.itensor import Index, qr, random_roc_itensor, svd, contract
    i, j, k, l = Index(2, "i"), Index(3, "j"), Index(4, "k"), Index(5, "l")
    A = random_roc_itensor(i, j, k, l, seed=4)
    U, S, V, spec = svd(A, (j, l))
    R = contract(contract(U, S), V)
    np.testing.assert_allclose(R.array(i, j, k, l), A.array(i, j, k, l), atol=1e-13)
    u = [x for x in U.inds if x not in A.inds][0]
    UU = contract(U, U.prime(1, [u]))
    np.testing.assert_allclose(UU.array(), np.eye(u.dim), atol=1e-14)
    Q, Rr = qr(A, (i, l))
    np.testing.assert_allclose(contract(Q, Rr).array(i, j, k, l), A.array(i, j, k, l), atol=1e-13)


def test_contract_errors(itg):
    from itg import DimensionMismatch
    from itg.itensor import Index, random_roc_itensor
    i, j = Index(2, "i"), Index(3, "j")
    A = random_roc_itensor(i, j, seed=1)
    B = random_roc_itensor(j, seed=2)
    with pytest.raises(DimensionMismatch):
        A + B


# ----------------------------------------------------------------- davidson
def test_davidson_10x10(itg):
    """test/test_cuiterativesolvers.jl:13-27: M(v) ~ lambda v."""
    from itg import ops as dev
    from itg.mps import davidson
    rng = np.random.default_rng(2)
    A = rng.standard_normal((10, 10))
    M = A @ A.T
    dM = dev.to_device(ops.to_flat(M))

    def mv(x):
        y = dev.empty(10)
        dev.gemm("N", "N", 10, 1, 10, 1.0, dM, 10, x, 10, 0.0, y, 10)
        return y

    lam, v = davidson(mv, dev.to_device(rng.standard_normal(10)), maxiter=10)
    vh = dev.to_host(v)
    np.testing.assert_allclose(M @ vh, lam * vh, atol=1e-9)
    assert math.isclose(lam, np.linalg.eigvalsh(M)[0], rel_tol=1e-9)


# -------------------------------------------------------------- random MPS
def test_random_mps_matches_oracle(itg):
    from itg.mps import random_mps
    psi = random_mps(12, 2, 10, seed=0)
    ref = odmrg.random_mps(12, 2, 10, seed=0)
    for a, b in zip(psi.to_host(), ref):
        assert a.shape == b.shape
        np.testing.assert_allclose(a, b, atol=1e-12)


# --------------------------------------------------------------------- dmrg
def heis_mpo(N):
    from itg.mps import MPO, heisenberg_opsum
    return MPO.from_opsum(heisenberg_opsum([(j, j + 1) for j in range(N - 1)]), N)


def test_dmrg_config1_n20_fp64(itg):
    """S=1/2 Heisenberg N=20 (config 1): random MPS linkdims 10 seed 0,
    5 sweeps maxdim [10,20,100,100,200], cutoff 1e-10, 2 Krylov vectors.
    Compared with the oracle's identical run and with ED."""
    from itg.mps import Sweeps, dmrg, expect_energy, random_mps
    N = 20
    H = heis_mpo(N)
    psi = random_mps(N, 2, 10, seed=0)
    rec = []
    sw = Sweeps(5, maxdim=[10, 20, 100, 100, 200], cutoff=[1e-10])
    E = dmrg(H, psi, sw, krylovdim=2, record=rec)
    print(rec)
    Eo = odmrg.dmrg(models.heisenberg_mpo(N), odmrg.random_mps(N, 2, 10, seed=0),
                    maxdims=[10, 20, 100, 100, 200], cutoffs=[1e-10] * 5, krylovdim=2)
    assert E0_N20 - 1e-10 <= E < E0_N20 + 1e-7
    assert abs(E - Eo) < 1e-8, (E, Eo)
    assert abs(expect_energy(H, psi) - E) < 1e-9


def test_dmrg_tfim_reference(itg):
    """test/dmrg.jl:58-81: critical TFIM N=32, relative error < 1e-2."""
    from itg.mps import MPO, Sweeps, dmrg, random_mps, tfim_opsum
    N = 32
    H = MPO.from_opsum(tfim_opsum(N), N)
    psi = random_mps(N, 2, 10, seed=432)
    E = dmrg(H, psi, Sweeps(5, maxdim=[10, 20], cutoff=[1e-12], noise=[1e-10]))
    Ex = models.tfim_exact_energy(N)
    print(E, Ex)
    assert abs((E - Ex) / Ex) < 1e-2
    assert E >= Ex - 1e-9


@pytest.mark.parametrize("ortho", ["left", "right"])
@pytest.mark.parametrize("mode", ["native", "bf16x3"])
def test_noiseterm_matches_oracle(itg, ortho, mode):
    """The density-matrix noise term on the device against the oracle's
    (itself pinned by the operator-sandwich form), at a mid-chain bond with
    chi 24 x 32: FP64 to 1e-12, split-BF16 to 5e-5 (it is scaled by the
    noise before it enters the density matrix)."""
    from itg import ops as dev
    from itg.mps import ProjMPO, noiseterm, random_mps
    N, b = 12, 5
    H = heis_mpo(N)
    psi = random_mps(N, 2, 32, seed=9)
    PH = ProjMPO(H, mode)
    PH.build_right(psi, upto=b + 2)
    from itg.mps import update_left
    for j in range(b):
        L, Ld = PH.L[j]
        PH.L[j + 1] = update_left(L, Ld, psi.data[j], psi.dims[j], H.data[j], H.dims[j])
    phi, pd = dev.contract(psi.data[b], psi.dims[b], [1, 2, 3], psi.data[b + 1],
                           psi.dims[b + 1], [3, 4, 5], [1, 2, 4, 5])
    drho = dev.to_host(noiseterm(PH, b, phi, pd, ortho, mode))
    # the MPO the environments were built with (its channel order differs
    # from the oracle's own Heisenberg MPO, so that one cannot be mixed in)
    Hh = {j: ops.from_flat(dev.to_host(H.data[j]), H.dims[j]) for j in (b, b + 1)}
    Lh = ops.from_flat(dev.to_host(PH.L[b][0]), PH.L[b][1])
    Rh = ops.from_flat(dev.to_host(PH.R[b + 2][0]), PH.R[b + 2][1])
    ph = ops.from_flat(dev.to_host(phi), pd)
    ref = odmrg.noiseterm(Lh, Hh[b], Hh[b + 1], Rh, ph, ortho)
    got = drho.reshape(ref.shape, order="F")
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(ortho, mode, pd, err)
    assert err < (1e-12 if mode == "native" else 5e-5)
    np.testing.assert_array_equal(got, got.T)


def test_dmrg_with_noise_matches_oracle(itg):
    """Heisenberg N=20 with a decaying noise schedule (noise! of
    test/dmrg.jl:24): the device run follows the oracle's identical run
    (energies per sweep within 1e-8) and ends at the ED energy."""
    from itg.mps import Sweeps, dmrg, random_mps
    N = 20
    maxd, noises = [10, 20, 100, 100, 200, 200], [1e-5, 1e-7, 1e-9, 0.0, 0.0, 0.0]
    rec = []
    E = dmrg(heis_mpo(N), random_mps(N, 2, 10, seed=0),
             Sweeps(6, maxdim=maxd, cutoff=[1e-10], noise=noises), krylovdim=2, record=rec)
    reco = []
    # the oracle runs on the same MPO arrays: the noise term sums over the
    # open MPO link, so it depends on the MPO's channel basis
    from itg.mps import heisenberg_opsum
    Ws = heisenberg_opsum([(j, j + 1) for j in range(N - 1)]).to_mpo_arrays(N, 0.5)
    Eo = odmrg.dmrg(Ws, odmrg.random_mps(N, 2, 10, seed=0),
                    maxdims=maxd, cutoffs=[1e-10] * 6, krylovdim=2, noises=noises, record=reco)
    for r, ro in zip(rec, reco):
        print(r["energy_davidson"], ro["energy"], r["maxlinkdim"], ro["maxlinkdim"])
        assert abs(r["energy_davidson"] - ro["energy"]) < 1e-8
    assert abs(E - Eo) < 1e-8
    assert E0_N20 - 1e-10 <= E < E0_N20 + 1e-7


def test_dmrg_spin_one(itg):
    """test/dmrg.jl:5-29: S=1 Heisenberg N=10, maxdim 10,20,40: E < -12."""
    from itg.mps import MPO, Sweeps, dmrg, heisenberg_opsum, random_mps
    N = 10
    H = MPO.from_opsum(heisenberg_opsum([(j, j + 1) for j in range(N - 1)]), N, S=1.0)
    psi = random_mps(N, 3, 10, seed=1)
    E = dmrg(H, psi, Sweeps(3, maxdim=[10, 20, 40], mindim=[1, 10], cutoff=[1e-11]))
    print(E)
    assert E < -12.0


@pytest.mark.parametrize("mode", ["fp32", "bf16x3"])
def test_dmrg_mixed_precision_n20(itg, mode):
    """Contractions in FP32 / split-BF16 with FP64 Davidson, density matrix
    and energy: |dE|/N against the FP64 run."""
    from itg.mps import Sweeps, dmrg, random_mps
    N = 20
    sw = Sweeps(6, maxdim=[10, 20, 100, 100, 200], cutoff=[1e-12])
    E64 = dmrg(heis_mpo(N), random_mps(N, 2, 10, seed=0), sw)
    Em = dmrg(heis_mpo(N), random_mps(N, 2, 10, seed=0), sw, mode=mode)
    print(mode, E64, Em, abs(Em - E64) / N)
    # the precision target, |dE|/N <= 1e-9 (measured ~6e-14 fp32, ~2e-11 bf16x3)
    tol = 1e-9
    assert abs(Em - E64) / N < tol


def test_dmrg_diag_preconditioner(itg):
    from itg.mps import Sweeps, dmrg, random_mps
    N = 16
    E = dmrg(heis_mpo(N), random_mps(N, 2, 10, seed=3),
             Sweeps(5, maxdim=[10, 20, 64], cutoff=[1e-12]), precond="diag")
    Eed = models.ground_energy(models.bond_hamiltonian(N, [(j, j + 1) for j in range(N - 1)]))
    assert Eed - 1e-10 <= E < Eed + 1e-6


# ------------------------------------------------------ device Davidson
def _bond_setup(N=12, chi=16, seed=5, mode="native"):
    from itg import ops
    from itg.mps import MPO, ProjMPO, heisenberg_opsum, random_mps
    H = MPO.from_opsum(heisenberg_opsum([(j, j + 1) for j in range(N - 1)]), N)
    psi = random_mps(N, 2, chi, seed)
    PH = ProjMPO(H, mode)
    PH.build_right(psi)
    b = 0
    A, Ad, B, Bd = psi.data[b], psi.dims[b], psi.data[b + 1], psi.dims[b + 1]
    phi, pdims = ops.contract(A, Ad, [1, 2, 3], B, Bd, [3, 4, 5], [1, 2, 4, 5])
    return PH, b, phi, pdims


@pytest.mark.parametrize("mode", ["native", "bf16x3"])
@pytest.mark.parametrize("precond", [None, "diag"])
def test_davidson2_matches_host_davidson(itg, mode, precond):
    """itg_davidson2 against the host-orchestrated Davidson on the same
    prepared operator and start vector: same Ritz value and vector."""
    from itg import ops
    from itg.mps import davidson, diag_preconditioner
    PH, b, phi, pdims = _bond_setup(mode=mode)
    PH.prepare(b, pdims)
    dg = PH.diag(b, pdims) if precond == "diag" else None
    pre = diag_preconditioner(dg) if dg is not None else None
    for maxiter in (1, 3, 6):
        lam_h, v_h = davidson(lambda x: PH.product(b, x, pdims), ops.copy(phi), maxiter=maxiter,
                              precond=pre)
        v_n = ops.copy(phi)
        lam_n, v_n = PH.eigsolve(b, v_n, pdims, maxiter=maxiter, diag=dg)
        assert abs(lam_n - lam_h) < 1e-10 * abs(lam_h), (maxiter, lam_n, lam_h)
        ov = abs(ops.dot(v_n, v_h))
        assert abs(ov - 1.0) < 1e-9, (maxiter, ov)


@pytest.mark.parametrize("mode", ["native", "bf16x3"])
def test_davidson2_mlp_matches_host_guarded_davidson(itg, mode):
    """itg_davidson2_mlp (the guarded learned step inside the device loop)
    against the host Davidson with the same network correction and the
    diagonal fallback, for a network whose layers run as FP32 products
    (hidden 32): same Ritz value, vector, matvec and revert counts."""
    from itg import ops
    from itg.mps import davidson, diag_preconditioner
    from itg.precond import MLPPreconditioner
    PH, b, phi, pdims = _bond_setup(mode=mode)
    PH.prepare(b, pdims)
    dg = PH.diag(b, pdims)
    net = MLPPreconditioner(2, hidden=32, seed=3)
    net.set_params(np.random.default_rng(3).standard_normal(net.nparam).astype(np.float32) * 0.3)
    pre = lambda q, lam: net.correction(q, lam, dg, pdims)
    for maxiter in (1, 3, 6):
        st_h, st_n = {}, {}
        lam_h, v_h = davidson(lambda x: PH.product(b, x, pdims), ops.copy(phi), maxiter=maxiter,
                              precond=pre, fallback=diag_preconditioner(dg), stats=st_h)
        v_n = ops.copy(phi)
        lam_n, v_n = PH.eigsolve(b, v_n, pdims, maxiter=maxiter, diag=dg, stats=st_n, net=net)
        assert abs(lam_n - lam_h) < 1e-10 * abs(lam_h), (maxiter, lam_n, lam_h)
        assert abs(abs(ops.dot(v_n, v_h)) - 1.0) < 1e-9
        assert st_n["matvecs"] == st_h["matvecs"] and st_n["reverted"] == st_h["reverted"], \
            (maxiter, st_n, st_h)


def test_davidson2_mlp_zero_network_is_the_diagonal_run(itg):
    """A zero network without skip gives a zero correction: every learned
    step is taken back and the device run is the diagonal-preconditioned
    itg_davidson2 run bit for bit."""
    from itg import ops
    from itg.precond import MLPPreconditioner
    PH, b, phi, pdims = _bond_setup()
    PH.prepare(b, pdims)
    dg = PH.diag(b, pdims)
    net = MLPPreconditioner(2, hidden=256, skip="none")
    net.set_params(np.zeros(net.nparam, dtype=np.float32))
    st = {}
    v_n = ops.copy(phi)
    lam_n, v_n = PH.eigsolve(b, v_n, pdims, maxiter=3, diag=dg, stats=st, net=net)
    v_d = ops.copy(phi)
    lam_d, v_d = PH.eigsolve(b, v_d, pdims, maxiter=3, diag=dg)
    assert st["reverted"] == 3
    assert lam_n == lam_d
    assert np.array_equal(ops.to_host(v_n), ops.to_host(v_d))


def test_dmrg_learned_device_path_matches_host(itg):
    """dmrg with the trained network's factory on the device Davidson
    (itg_davidson2_mlp) and on the host Davidson: both converge to the same
    energy (N = 12 Heisenberg, chi 32, 4 sweeps), and the device run counts
    its reverted steps."""
    from itg.mps import MPO, Sweeps, dmrg, heisenberg_opsum, random_mps
    from itg.precond import MLPPreconditioner, train_sweep
    N = 12
    H = MPO.from_opsum(heisenberg_opsum([(j, j + 1) for j in range(N - 1)]), N)
    net = MLPPreconditioner(2, hidden=256, seed=1)
    train_sweep(net, H, random_mps(N, 2, 16, seed=4), 16, steps_per_bond=3, sweeps=1, lr=3e-4)
    sw = Sweeps(4, maxdim=[32], cutoff=[1e-12])
    rec_d, rec_h = [], []
    Ed = dmrg(H, random_mps(N, 2, 16, seed=2), sw, precond=net.factory(), record=rec_d)
    Eh = dmrg(H, random_mps(N, 2, 16, seed=2), sw, precond=net.factory(), native=False,
              record=rec_h)
    assert abs(Ed - Eh) < 1e-9 * abs(Eh)
    E0 = models.ground_energy(models.bond_hamiltonian(N, [(j, j + 1) for j in range(N - 1)]))
    assert abs(Ed - E0) < 1e-8
    assert all("reverted" in r for r in rec_d)


def test_davidson2_graph_replays_match_eager(itg):
    """Split-BF16 Heff applies are captured into HIP graphs on their second
    use: five calls with eager library work and synchronisations between
    them give bit-identical results to the first (eager) call."""
    import torch
    from itg import _lib, ops
    PH, b, phi, pdims = _bond_setup(N=14, chi=32, seed=9, mode="bf16x3")
    PH.prepare(b, pdims)
    scratch = ops.randn(1 << 16, 3)
    ref = None
    for call in range(5):
        v = ops.copy(phi)
        lam, v = PH.eigsolve(b, v, pdims, maxiter=4)
        host = ops.to_host(v)
        if ref is None:
            ref = (lam, host)
        else:
            assert lam == ref[0]
            assert np.array_equal(host, ref[1])
        ops.scal(scratch, 1.0001)          # eager work between replays
        _lib.ctx().sync()
        torch.cuda.synchronize()


def test_truncating_sweeps_memory_bounded(itg):
    """Truncating split-BF16 sweeps (N=40 from a random MPS at the full link
    dimensions min(2^j, 2^(N-j), 256), maxdim 256, cutoff 1e-9: the first
    sweep runs at the full shapes, the next ones truncate, so nearly every
    bond gets its own (chiL, chiR)): the ProjMPO keeps one grow-only set of
    work buffers, and the peak of device memory stays under a bound computed
    from the largest shape (environments, MPS, one Heff work + Krylov basis,
    eigensolver buffers)."""
    import torch
    from itg import _lib, ops
    from itg.mps import Sweeps, dmrg, random_mps
    N, chi, w, d = 40, 256, 5, 2
    torch.cuda.synchronize()
    base = torch.cuda.memory_allocated()
    torch.cuda.reset_peak_memory_stats()
    psi = random_mps(N, 2, chi, seed=4)
    rec = []
    dmrg(heis_mpo(N), psi, Sweeps(4, maxdim=[chi], cutoff=[1e-9]), mode="bf16x3", record=rec)
    torch.cuda.synchronize()
    PH = __import__("itg.mps", fromlist=["dmrg"]).dmrg.last_projmpo
    shapes = {dm for dm in psi.dims}
    peak = torch.cuda.max_memory_allocated() - base
    env = 2 * (N + 1) * chi * w * chi * 8
    mps = 2 * N * chi * d * chi * 8
    heff = ops.heff2_work_bytes(chi, w, w, w, d, chi) + \
        4 * _lib.lib().itg_heff2_prep_elems(chi, w, w, w, d, chi)
    kw = 8 * _lib.lib().itg_davidson2_work_elems(chi, d, chi, 2)
    eig = 6 * (d * chi) ** 2 * 8
    bound = 1.5 * (env + mps + heff + kw + eig) + (64 << 20)
    print("distinct site shapes %d, peak %.1f MB, bound %.1f MB, buffers %s"
          % (len(shapes), peak / 1e6, bound / 1e6, sorted(PH._bufs)))
    print("maxlinkdim per sweep", [r["maxlinkdim"] for r in rec])
    assert rec[0]["maxlinkdim"] == chi
    assert len(shapes) > 20
    assert len(PH._bufs) <= 3
    assert peak < bound
