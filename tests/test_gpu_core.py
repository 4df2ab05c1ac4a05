"""GPU parity: RNG, permutation, permuted add, BLAS-1, truncation and
contraction through the C ABI against the CPU oracle on seeded inputs."""
import itertools
import json
import math
import os

import numpy as np
import pytest

from oracle import philox
from oracle import tensor_ops as ops
from oracle.truncate import ndtensors_truncate, refgpu_truncate

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def itg():
    import itg as m
    from itg import ops as dev
    m._lib.ctx()
    return dev


def dev_tensor(dev, T, dtype="float64"):
    return dev.to_device(ops.to_flat(T), dtype)


def host_tensor(dev, t, dims):
    return ops.from_flat(dev.to_host(t), dims)


def test_device_is_gfx950(itg):
    from itg import _lib
    info = _lib.ctx().info()
    print(info)
    assert info["gfx"] == "0x950"


# ---------------------------------------------------------------------- rng
def test_randn_matches_host_stream(itg):
    for seed, off, n in [(0, 0, 1000), (7, 1001, 4097), (2 ** 40 + 3, 5, 333)]:
        x = itg.to_host(itg.randn(n, seed, off))
        y = philox.randn(n, seed, off)
        np.testing.assert_allclose(x, y, rtol=1e-13, atol=1e-13)


# ------------------------------------------------------------------ permute
def test_permute_closed_form_transpose(itg):
    """test/test_cudense.jl:34-42, bit-exact."""
    A = np.array([[float((i + 1) * (j + 1)) for j in range(3)] for i in range(2)])
    B, dB = itg.permute(dev_tensor(itg, A), (2, 3), (1, 0))
    assert dB == (3, 2)
    assert np.array_equal(itg.to_host(B), ops.to_flat(A.T))


def test_permute_14_index_shuffles_bit_exact(itg):
    """test/test_cuitensor.jl:44-57: 14 indices of dim 2, 20 shuffles."""
    rng = np.random.default_rng(11)
    dims = (2,) * 14
    A = rng.standard_normal(dims)
    dA = dev_tensor(itg, A)
    for _ in range(20):
        perm = rng.permutation(14)
        B, dB = itg.permute(dA, dims, perm)
        assert np.array_equal(host_tensor(itg, B, dB), np.transpose(A, perm))


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_permute_random_shapes_bit_exact(itg, dtype):
    rng = np.random.default_rng(12)
    shapes = [(1,), (7, 1, 3), (64, 48), (33, 17, 65), (2, 3, 4, 5), (40, 2, 2, 37),
              (100, 5, 100), (3, 128, 2, 64), (1, 1, 1)]
    for dims in shapes:
        A = rng.standard_normal(dims).astype(dtype)
        dA = dev_tensor(itg, A, dtype)
        for perm in list(itertools.permutations(range(len(dims))))[:24]:
            B, dB = itg.permute(dA, dims, perm)
            assert np.array_equal(host_tensor(itg, B, dB), np.transpose(A, perm)), (dims, perm)


def test_permute_alpha_and_add(itg):
    rng = np.random.default_rng(13)
    A = rng.standard_normal((30, 20, 10))
    B = rng.standard_normal((10, 30, 20))
    dA, dB = dev_tensor(itg, A), dev_tensor(itg, B)
    P, _ = itg.permute(dA, A.shape, (2, 0, 1), alpha=-2.5)
    np.testing.assert_allclose(host_tensor(itg, P, B.shape), -2.5 * np.transpose(A, (2, 0, 1)),
                               rtol=0, atol=1e-15)
    itg.add_permuted(dA, A.shape, (2, 0, 1), 1.0, dB, -1.0)
    ref = ops.add_permuted(A, (2, 0, 1), 1.0, B, -1.0)
    np.testing.assert_allclose(host_tensor(itg, dB, B.shape), ref, rtol=0, atol=1e-15)


def test_add_closed_form_ones_twos(itg):
    """test/test_cudense.jl:16-33: ones(i,j) + twos(j,i) permuted -> threes."""
    A = np.ones((2, 3))
    dA = dev_tensor(itg, A)
    dB = itg.fill(itg.empty(6), 2.0)
    itg.add_permuted(dA, (2, 3), (1, 0), 1.0, dB, 1.0)
    assert np.array_equal(itg.to_host(dB), 3.0 * np.ones(6))


# ------------------------------------------------------------------- BLAS-1
def test_blas1(itg):
    rng = np.random.default_rng(14)
    for n in [1, 255, 256, 257, 100003]:
        x, y = rng.standard_normal(n), rng.standard_normal(n)
        dx, dy = itg.to_device(x), itg.to_device(y)
        assert math.isclose(itg.dot(dx, dy), float(x @ y), rel_tol=1e-12, abs_tol=1e-12)
        assert math.isclose(itg.nrm2(dx), float(np.linalg.norm(x)), rel_tol=1e-13)
        itg.axpby(0.5, dx, -2.0, dy)
        np.testing.assert_allclose(itg.to_host(dy), 0.5 * x - 2.0 * y, rtol=1e-15, atol=1e-15)
        itg.scal(dx, 3.0)
        np.testing.assert_allclose(itg.to_host(dx), 3.0 * x, rtol=1e-15)


# ----------------------------------------------------------------- truncate
def test_truncate_golden_vectors(itg):
    cases = json.load(open(os.path.join(GOLD, "truncate_ref_vectors.json")))["cases"]
    for c in cases:
        for fn in (itg.truncate, lambda P, **kw: itg.truncate_device(itg.to_device(P), **kw)):
            n, terr, docut = fn(c["P"], semantics="refgpu", **c["kwargs"])
            assert n == len(c["kept"])
            assert math.isclose(terr, c["truncerr"], rel_tol=1e-12, abs_tol=1e-15)
            assert math.isclose(docut, c["docut"], rel_tol=1e-12, abs_tol=1e-15)


def test_truncate_random_vs_oracle(itg):
    rng = np.random.default_rng(15)
    for trial in range(60):
        m = int(rng.integers(1, 300))
        P = np.sort(rng.random(m) ** 8)[::-1].copy()
        if trial % 7 == 0:
            P[-3:] = -1e-18
        kw = dict(maxdim=int(rng.integers(1, m + 1)), cutoff=10.0 ** rng.uniform(-12, -1),
                  absolute_cutoff=bool(trial % 5 == 0))
        for sem, orc in (("ndtensors", ndtensors_truncate), ("refgpu", refgpu_truncate)):
            want = orc(P, **kw)
            for got in (itg.truncate(P, semantics=sem, **kw),
                        itg.truncate_device(itg.to_device(P), semantics=sem, **kw)):
                assert got[0] == want[0], (sem, kw)
                assert math.isclose(got[1], want[1], rel_tol=1e-12, abs_tol=1e-300)
                assert math.isclose(got[2], want[2], rel_tol=1e-12, abs_tol=1e-300)


# -------------------------------------------------------------- contraction
TOL = {"native": 1e-12, "fp32": 2e-5, "bf16x3": 5e-5}


def rel_err(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def check_contract(itg, A, la, B, lb, mode="native", lc=None, alpha=1.0, beta=0.0, C0=None):
    lc = lc if lc is not None else ops.itensor_output_labels(la, lb)
    labs = {x: i + 1 for i, x in enumerate(dict.fromkeys(list(la) + list(lb)))}
    out = None
    if C0 is not None:
        out = dev_tensor(itg, C0)
    dC, dims = itg.contract(dev_tensor(itg, A), A.shape, [labs[x] for x in la],
                            dev_tensor(itg, B), B.shape, [labs[x] for x in lb],
                            [labs[x] for x in lc], mode=mode, alpha=alpha, beta=beta, out=out)
    got = host_tensor(itg, dC, dims)
    want = ops.contract(A, la, B, lb, lc, alpha=alpha, beta=beta, C=C0)
    err = rel_err(got, np.asarray(want))
    assert err <= TOL[mode], (la, lb, lc, mode, err)
    return err


def test_contract_rank_pairings(itg):
    """test/test_cucontract.jl:33-159: scalar/vector/matrix/3-/4-tensor pairs."""
    rng = np.random.default_rng(16)
    dims = {"i": 2, "j": 3, "k": 4, "l": 5, "a": 6}
    pairs = [("", ""), ("", "i"), ("i", ""), ("j", ""), ("i", "i"), ("i", "j"),
             ("ijk", "kl"), ("ij", "ij"), ("ij", "ji"), ("ijk", "jkl"), ("ijk", "kla"),
             ("ijkl", "kla"), ("ijkl", "jkl"), ("ik", "kj"), ("ij", "jk"), ("jk", "ij")]
    for la, lb in pairs:
        A = rng.standard_normal([dims[x] for x in la])
        B = rng.standard_normal([dims[x] for x in lb])
        check_contract(itg, A, list(la), B, list(lb))


def test_contract_all_permutations_3x3_and_4x3(itg):
    """test/test_cucontract.jl:170-196: every operand index order."""
    rng = np.random.default_rng(17)
    dims = {"i": 2, "j": 3, "k": 4, "l": 5, "a": 6}
    for base_a, base_b in [("ijk", "jkl"), ("ijkl", "kla"), ("ijkl", "jkl")]:
        A0 = rng.standard_normal([dims[x] for x in base_a])
        B0 = rng.standard_normal([dims[x] for x in base_b])
        for pa in itertools.permutations(range(len(base_a))):
            for pb in itertools.permutations(range(len(base_b))):
                la = [base_a[p] for p in pa]
                lb = [base_b[p] for p in pb]
                # B*A order as in the reference (C = Ajkl*Aijk)
                check_contract(itg, np.transpose(B0, pb), lb, np.transpose(A0, pa), la)


def test_contract_supersized_14_index(itg):
    """test/test_cucontract.jl:197-221: 14-index tensors of dim 2."""
    rng = np.random.default_rng(18)
    a_only = ["a%d" % i for i in range(7)]
    b_only = ["b%d" % i for i in range(7)]
    shared = ["s%d" % i for i in range(7)]
    A = rng.standard_normal((2,) * 14)
    B = rng.standard_normal((2,) * 14)
    check_contract(itg, A, a_only + shared, B, b_only + shared)
    for _ in range(3):
        pa, pb = rng.permutation(14), rng.permutation(14)
        la = [(a_only + shared)[p] for p in pa]
        lb = [(b_only + shared)[p] for p in pb]
        check_contract(itg, np.transpose(A, pa), la, np.transpose(B, pb), lb)


@pytest.mark.parametrize("mode", ["native", "fp32", "bf16x3"])
def test_contract_modes_medium(itg, mode):
    rng = np.random.default_rng(19)
    errs = []
    for (la, lb, shape) in [("pwa", "asub", dict(p=300, w=5, a=257, s=2, u=2, b=190)),
                            ("ptzvb", "qvb", dict(p=130, t=2, z=2, v=5, b=129, q=140)),
                            ("ab", "bc", dict(a=513, b=700, c=255)),
                            ("ba", "cb", dict(a=200, b=1030, c=64))]:
        A = rng.standard_normal([shape[x] for x in la])
        B = rng.standard_normal([shape[x] for x in lb])
        errs.append(check_contract(itg, A, list(la), B, list(lb), mode=mode))
    print(mode, errs)


def test_contract_alpha_beta_and_split_k(itg):
    rng = np.random.default_rng(20)
    A = rng.standard_normal((40, 6000))
    B = rng.standard_normal((6000, 30))
    C0 = rng.standard_normal((40, 30))
    check_contract(itg, A, ["i", "k"], B, ["k", "j"], alpha=0.7, beta=-1.3, C0=C0)
    # dot-like: scalar output with a large contracted extent
    x = rng.standard_normal((50, 40, 30))
    y = rng.standard_normal((30, 50, 40))
    check_contract(itg, x, list("abc"), y, list("cab"))


def test_contract_small_operand(itg):
    """Contractions with one small operand (the MPO tensor of an environment
    update): the big free modes do not merge, so they run as a stream over
    the big operand. Both operand roles, odd unequal bond dimensions, a
    non-mergeable summed group, and beta != 0."""
    rng = np.random.default_rng(22)
    w, d = 5, 2
    for chip, chi in ((97, 131), (131, 97), (64, 70)):
        T1 = rng.standard_normal((chip, w, d, chi))          # (a', w, s, b)
        W = rng.standard_normal((w, w, d, d))                # (w, w', s', s)
        # the left update's order (a', s', w', b) and the right one's (s', a', w', b)
        check_contract(itg, T1, list("pwsb"), W, list("wvts"), lc=list("ptvb"))
        check_contract(itg, T1, list("pwsb"), W, list("wvts"), lc=list("tpvb"))
        C0 = rng.standard_normal((chip, d, w, chi))
        check_contract(itg, T1, list("pwsb"), W, list("wvts"), lc=list("ptvb"), alpha=0.5,
                       beta=-2.0, C0=C0)
    for mode in ["fp32", "bf16x3"]:
        T1 = rng.standard_normal((97, w, d, 131))
        check_contract(itg, T1, list("pwsb"), rng.standard_normal((w, w, d, d)), list("wvts"),
                       lc=list("ptvb"), mode=mode)


def test_contract_config2_lpsi(itg):
    """Effective-Hamiltonian step 1 at chi=512: L(chi,w,chi) . psi(chi,d,d,chi)."""
    rng = np.random.default_rng(21)
    chi, w, d = 512, 5, 2
    L = rng.standard_normal((chi, w, chi))
    psi = rng.standard_normal((chi, d, d, chi))
    for mode in ["native", "fp32", "bf16x3"]:
        print(mode, check_contract(itg, L, list("pwa"), psi, list("asub"), mode=mode))


def test_contract_errors(itg):
    from itg import DimensionMismatch
    a = itg.to_device(np.ones(6))
    with pytest.raises(DimensionMismatch):
        itg.contract(a, (2, 3), [1, 2], a, (3, 2), [1, 3], [2, 3])   # dims of label 1 differ


def test_gemm_all_transposes(itg):
    rng = np.random.default_rng(22)
    m, n, k = 150, 70, 300
    for ta in "NT":
        for tb in "NT":
            A = rng.standard_normal((m, k) if ta == "N" else (k, m))
            B = rng.standard_normal((k, n) if tb == "N" else (n, k))
            C0 = np.zeros((m, n))
            dC = itg.to_device(ops.to_flat(C0))
            itg.gemm(ta, tb, m, n, k, 1.0, itg.to_device(ops.to_flat(A)), A.shape[0],
                     itg.to_device(ops.to_flat(B)), B.shape[0], 0.0, dC, m)
            ref = (A if ta == "N" else A.T) @ (B if tb == "N" else B.T)
            assert rel_err(ops.from_flat(itg.to_host(dC), (m, n)), ref) < 1e-13


@pytest.mark.parametrize("m,k,ta,tb", [(256, 300, "N", "T"), (300, 257, "T", "N"),
                                       (1000, 64, "N", "T")])
def test_gemm_symmetric_product(itg, m, k, ta, tb):
    """C = A A^T (or A^T A) with one buffer for both operands: the lower
    triangle is multiplied and mirrored, so C is exactly symmetric and
    matches the product to FP64 rounding; leading dimension larger than m."""
    rng = np.random.default_rng(m + k)
    lda = (m if ta == "N" else k) + 3
    A = rng.standard_normal((lda, k if ta == "N" else m))
    dA = itg.to_device(ops.to_flat(A))
    ldc = m + 5
    dC = itg.to_device(ops.to_flat(np.full((ldc, m), 7.0)))
    itg.gemm(ta, tb, m, m, k, 0.5, dA, lda, dA, lda, 0.0, dC, ldc)
    C = ops.from_flat(itg.to_host(dC), (ldc, m))
    Av = A[:m, :] if ta == "N" else A[:k, :]
    ref = 0.5 * (Av @ Av.T if ta == "N" else Av.T @ Av)
    assert rel_err(C[:m], ref) < 1e-13
    assert np.array_equal(C[:m], C[:m].T)
    assert np.all(C[m:] == 7.0)
