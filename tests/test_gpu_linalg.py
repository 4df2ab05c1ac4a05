"""GPU parity: effective-Hamiltonian apply/diagonal, truncated eigen, SVD and
QR through the C ABI against the CPU oracle (numpy LAPACK) on seeded inputs."""
import math

import numpy as np
import pytest

from oracle import dmrg as odmrg
from oracle import linalg as olin
from oracle import tensor_ops as ops
from oracle.truncate import ndtensors_truncate

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import itg
    from itg import ops as d
    itg._lib.ctx()
    return d


def up(dev, a, dtype="float64"):
    return dev.to_device(ops.to_flat(a), dtype)


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def sym_env(rng, chi, w):
    L = rng.standard_normal((chi, w, chi))
    return L + np.transpose(L, (2, 1, 0))


# --------------------------------------------------------------------- Heff
@pytest.mark.parametrize("chiL,wl,wm,wr,d,chiR", [(16, 1, 5, 5, 2, 24), (40, 5, 5, 5, 2, 40),
                                                  (33, 5, 5, 1, 2, 17), (64, 14, 14, 14, 2, 48),
                                                  (20, 5, 5, 5, 3, 20), (257, 5, 5, 5, 2, 130)])
def test_heff2_apply_matches_oracle(dev, chiL, wl, wm, wr, d, chiR):
    rng = np.random.default_rng(chiL + 7 * wl)
    L = sym_env(rng, chiL, wl)
    R = sym_env(rng, chiR, wr)
    W1 = rng.standard_normal((wl, wm, d, d))
    W2 = rng.standard_normal((wm, wr, d, d))
    psi = rng.standard_normal((chiL, d, d, chiR))
    want = odmrg.heff_apply(L, W1, W2, R, psi)
    for mode, tol in [("native", 1e-13), ("fp32", 2e-5), ("bf16x3", 5e-5)]:
        y = dev.heff2_apply(up(dev, L), up(dev, W1), up(dev, W2), up(dev, R), up(dev, psi),
                            chiL, wl, wm, wr, d, chiR, mode=mode)
        err = rel(ops.from_flat(dev.to_host(y), want.shape), want)
        print(mode, (chiL, wl, wm, wr, d, chiR), err)
        assert err < tol
    dg = dev.heff2_diag(up(dev, L), up(dev, W1), up(dev, W2), up(dev, R), chiL, wl, wm, wr, d, chiR)
    np.testing.assert_allclose(ops.from_flat(dev.to_host(dg), want.shape),
                               odmrg.heff_diag(L, W1, W2, R), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("mode", ["native", "fp32", "bf16x3"])
def test_heff2_prepared_equals_one_shot(dev, dtype, mode):
    """prepare once + apply per vector runs the same kernels as the one-shot
    apply: bit-identical output, and within the mode's tolerance of the
    oracle for FP64 and FP32 storage."""
    rng = np.random.default_rng(5)
    chiL, wl, wm, wr, d, chiR = 72, 5, 5, 5, 2, 136
    L = sym_env(rng, chiL, wl)
    R = sym_env(rng, chiR, wr)
    W1 = rng.standard_normal((wl, wm, d, d))
    W2 = rng.standard_normal((wm, wr, d, d))
    dims = (chiL, wl, wm, wr, d, chiR)
    args = [up(dev, a, dtype) for a in (L, W1, W2, R)]
    prep = dev.heff2_prepare(*args, *dims, mode=mode)
    tol = {"native": 1e-13 if dtype == "float64" else 2e-6, "fp32": 2e-5, "bf16x3": 5e-5}[mode]
    for seed in (1, 2):
        psi = np.random.default_rng(seed).standard_normal((chiL, d, d, chiR))
        x = up(dev, psi, dtype)
        y1 = dev.to_host(dev.heff2_apply(*args, x, *dims, mode=mode))
        y2 = dev.to_host(dev.heff2_apply_prepared(args[0], args[3], prep, x, *dims, mode=mode))
        assert np.array_equal(y1, y2)
        want = odmrg.heff_apply(L, W1, W2, R, psi)
        assert rel(ops.from_flat(y2, want.shape), want) < tol


def test_heff2_config2_chi1024(dev):
    """Effective-Hamiltonian microbench shapes: chi=1024, w=5, d=2."""
    rng = np.random.default_rng(2)
    chi, w, d = 1024, 5, 2
    L = rng.standard_normal((chi, w, chi))
    R = rng.standard_normal((chi, w, chi))
    W1 = rng.standard_normal((w, w, d, d))
    W2 = rng.standard_normal((w, w, d, d))
    psi = rng.standard_normal((chi, d, d, chi))
    want = odmrg.heff_apply(L, W1, W2, R, psi)
    for mode, tol in [("native", 1e-13), ("fp32", 2e-5), ("bf16x3", 5e-5)]:
        y = dev.heff2_apply(up(dev, L), up(dev, W1), up(dev, W2), up(dev, R), up(dev, psi),
                            chi, w, w, w, d, chi, mode=mode)
        err = rel(ops.from_flat(dev.to_host(y), want.shape), want)
        print("chi=1024", mode, err)
        assert err < tol


@pytest.mark.parametrize("chiL,w,d,chiR,dtype", [(800, 5, 2, 1000, "float64"),
                                                 (800, 5, 2, 1000, "float32"),
                                                 (500, 2, 4, 1900, "float64")])
def test_heff2_bf16x3_large_ragged(dev, chiL, w, d, chiR, dtype):
    """Split-BF16 products large enough for the 256 x 256 tile, ragged in both
    directions: C1 at 4000 x 4000 (K = 800), and in the last case C1 at
    1000 x 30400 (K = 500) and C4 at 8000 x 1900 (K = 3800), against the
    oracle."""
    rng = np.random.default_rng(chiL + chiR)
    L = sym_env(rng, chiL, w)
    R = sym_env(rng, chiR, w)
    W1 = rng.standard_normal((w, w, d, d))
    W2 = rng.standard_normal((w, w, d, d))
    psi = rng.standard_normal((chiL, d, d, chiR))
    want = odmrg.heff_apply_blas(L, W1, W2, R, psi)   # equal to heff_apply (test_oracle)
    y = dev.heff2_apply(*(up(dev, a, dtype) for a in (L, W1, W2, R, psi)),
                        chiL, w, w, w, d, chiR, mode="bf16x3")
    err = rel(ops.from_flat(dev.to_host(y), want.shape), want)
    print((chiL, w, d, chiR, dtype), err)
    assert err < 5e-5


def test_heff2_bf16x3_headline_shape_against_fp64(dev):
    """The split-BF16 apply at the bulk-bond shape of the sweep benchmark
    (chi=2048, w=5, d=2: C1 is 10240 x 8192 x 2048 and C4 8192 x 2048 x
    10240, both on 256 x 256 tiles) against the FP64 apply of the same
    device operands, which matches the oracle at chi=1024 above; the
    tolerance is the oracle tests'."""
    chi, w, d = 2048, 5, 2
    L = dev.randn(chi * w * chi, 31)
    R = dev.randn(chi * w * chi, 32)
    W1 = dev.randn(w * w * d * d, 33)
    W2 = dev.randn(w * w * d * d, 34)
    x = dev.randn(chi * d * d * chi, 35)
    y64 = dev.heff2_apply(L, W1, W2, R, x, chi, w, w, w, d, chi, mode="native")
    y3 = dev.heff2_apply(L, W1, W2, R, x, chi, w, w, w, d, chi, mode="bf16x3")
    ref = dev.nrm2(y64)
    dev.axpby(-1.0, y64, 1.0, y3)
    err = dev.nrm2(y3) / ref
    print("chi=2048 bf16x3 against fp64", err)
    assert 0.0 < err < 5e-5


# -------------------------------------------------------------------- eigen
def density_matrix(rng, n, decades=14):
    lam = 10.0 ** (-np.linspace(0, decades, n)) * rng.uniform(0.5, 1.0, n)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (Q * lam) @ Q.T


@pytest.mark.parametrize("n", [1, 2, 3, 31, 32, 33, 100, 257, 600])
def test_eigh_full_spectrum(dev, n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    A = A + A.T
    w, U, k, te, dc = dev.eigh_trunc(up(dev, A), n)
    assert k == n
    wh = dev.to_host(w)
    Uh = ops.from_flat(dev.to_host(U), (n, n))
    wr = np.linalg.eigvalsh(A)[::-1]
    scale = np.abs(wr).max()
    assert np.abs(wh - wr).max() <= 1e-13 * scale * max(1, n / 64)
    assert np.abs(Uh.T @ Uh - np.eye(n)).max() < 1e-12
    assert np.abs(A @ Uh - Uh * wh).max() < 1e-12 * scale * max(1, n / 64)


@pytest.mark.parametrize("n,maxdim,cutoff", [(200, 100, 1e-10), (512, 512, 1e-12), (1024, 300, 1e-8)])
def test_eigh_truncated_density_matrix(dev, n, maxdim, cutoff):
    rng = np.random.default_rng(n + 1)
    rho = density_matrix(rng, n)
    w, U, k, te, dc = dev.eigh_trunc(up(dev, rho), n, maxdim=maxdim, cutoff=cutoff)
    wr, Ur, ter, dcr = olin.eigh_trunc(rho, maxdim=maxdim, cutoff=cutoff)
    assert k == Ur.shape[1]
    # discarded weights are eigenvalues known to eps*|rho| each (absolute)
    assert abs(te - ter) <= (n - k) * 4e-16 * wr[0] / np.trace(rho) + 1e-15 * ter
    wh = dev.to_host(w)[:k]
    assert np.abs(wh - wr).max() <= 1e-14 * wr[0] * 10
    Uh = ops.from_flat(dev.to_host(U)[: n * k], (n, k))
    # eigenvectors up to sign, for well-separated eigenvalues
    gaps = np.abs(np.diff(wr))
    ok = np.ones(k, bool)
    ok[:-1] &= gaps > 1e-6 * wr[0]
    ok[1:] &= gaps > 1e-6 * wr[0]
    dots = np.abs(np.sum(Uh * Ur, axis=0))
    assert np.all(np.abs(dots[ok] - 1) < 1e-8)
    assert np.abs(Uh.T @ Uh - np.eye(k)).max() < 1e-12


@pytest.mark.parametrize("n,cols", [(257, 257), (600, 300), (1024, 1024), (4096, 2048)])
def test_eigh_two_stage_reduction(dev, monkeypatch, n, cols):
    """The two-stage reduction (dense -> band of width 32 -> tridiagonal by
    bulge chasing, then the two back-transformations) against LAPACK on a
    density matrix M M^T: full rank, and with a (n - cols)-fold cluster at
    zero; every eigenvalue to 1e-13 of the largest, the kept vectors'
    eigen-residual and orthonormality."""
    monkeypatch.setenv("ITG_EIGH_TWOSTAGE", "1")
    rng = np.random.default_rng(n + cols)
    M = rng.standard_normal((n, cols))
    M /= np.linalg.norm(M)
    rho = M @ M.T
    maxdim = n // 2
    w, U, k, te, dc = dev.eigh_trunc(up(dev, rho), n, maxdim=maxdim, cutoff=1e-12)
    wr, Ur, ter, dcr = olin.eigh_trunc(rho, maxdim=maxdim, cutoff=1e-12)
    assert k == Ur.shape[1]
    scale = wr[0]
    wh = dev.to_host(w)[:k]
    assert np.abs(wh - wr).max() <= 1e-13 * scale
    assert abs(te - ter) <= n * 1e-16 * scale + 1e-12 * ter
    Uh = ops.from_flat(dev.to_host(U)[: n * k], (n, k))
    assert np.abs(Uh.T @ Uh - np.eye(k)).max() < 1e-12
    assert np.abs(rho @ Uh - Uh * wh).max() < 1e-13 * scale


def test_eigh_degenerate_multiplets(dev):
    """Exactly degenerate eigenvalues (SU(2) multiplets in Heisenberg density
    matrices): the kept invariant subspace must be exact."""
    rng = np.random.default_rng(9)
    lam = np.repeat(10.0 ** -np.arange(0, 12, 1.0), [1, 3, 5, 3, 1, 3, 5, 7, 3, 1, 3, 5])
    n = lam.size
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    rho = (Q * lam) @ Q.T
    w, U, k, te, dc = dev.eigh_trunc(up(dev, rho), n, maxdim=n, cutoff=1e-12)
    Uh = ops.from_flat(dev.to_host(U)[: n * k], (n, k))
    P = Uh @ Uh.T
    Pr = Q[:, :k] @ Q[:, :k].T
    assert np.abs(P - Pr).max() < 1e-9
    assert np.abs(Uh.T @ Uh - np.eye(k)).max() < 1e-12


def test_eigh_degenerate_large(dev):
    """Twenty 30-fold degenerate eigenvalues at n=600: the merges deflate by
    Givens rotations (the sequential deflation pass), the spectrum and the
    kept invariant subspaces must come out exact."""
    rng = np.random.default_rng(17)
    vals = 10.0 ** -np.linspace(0, 10, 20)
    lam = np.repeat(vals, 30)
    n = lam.size
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    rho = (Q * lam) @ Q.T
    w, U, k, te, dc = dev.eigh_trunc(up(dev, rho), n, maxdim=300, cutoff=0.0)
    assert k == 300
    wh = dev.to_host(w)[:k]
    assert np.abs(wh - lam[:k]).max() <= 1e-14 * 10
    Uh = ops.from_flat(dev.to_host(U)[: n * k], (n, k))
    assert np.abs(Uh.T @ Uh - np.eye(k)).max() < 1e-12
    P = Uh @ Uh.T
    Pr = Q[:, :k] @ Q[:, :k].T
    assert np.abs(P - Pr).max() < 1e-9


def test_eigh_fp32_storage(dev):
    rng = np.random.default_rng(10)
    n = 150
    A = rng.standard_normal((n, n))
    A = A + A.T
    w, U, k, te, dc = dev.eigh_trunc(up(dev, A, "float32"), n)
    wr = np.linalg.eigvalsh(A.astype(np.float32).astype(np.float64))[::-1]
    assert np.abs(dev.to_host(w) - wr).max() < 1e-12 * np.abs(wr).max()


# ---------------------------------------------------------------------- svd
@pytest.mark.parametrize("m,n", [(15, 8), (8, 15), (40, 40), (100, 37), (1, 5), (5, 1)])
def test_svd_invariants(dev, m, n):
    """test/test_cuitensor.jl:105-112: A = U S V^T, U^T U = I, V^T V = I
    (rtol 1e-14), singular values against LAPACK."""
    rng = np.random.default_rng(m * 100 + n)
    A = rng.standard_normal((m, n))
    U, S, V, k, te, dc = dev.svd_trunc(up(dev, A), m, n)
    r = min(m, n)
    assert k == r
    Uh = ops.from_flat(dev.to_host(U)[: m * k], (m, k))
    Vh = ops.from_flat(dev.to_host(V)[: n * k], (n, k))
    Sh = dev.to_host(S)[:k]
    np.testing.assert_allclose(Uh @ np.diag(Sh) @ Vh.T, A, atol=1e-13 * np.abs(A).max() * 10)
    assert np.abs(Uh.T @ Uh - np.eye(k)).max() < 1e-14 * 10
    assert np.abs(Vh.T @ Vh - np.eye(k)).max() < 1e-14 * 10
    np.testing.assert_allclose(Sh, np.linalg.svd(A, compute_uv=False), rtol=1e-13)


@pytest.mark.parametrize("m,n,r", [(300, 700, 50), (900, 400, 120)])
def test_svd_rank_deficient_converges(dev, m, n, r):
    """Columns that rotate down to the rounding level of ||A||_F must not keep
    the one-sided Jacobi from converging (two-site tensors of DMRG are often
    rank deficient); the rank-r part is reproduced and truncation at 1e-14
    keeps exactly r values."""
    rng = np.random.default_rng(m + n + r)
    U0, _ = np.linalg.qr(rng.standard_normal((m, r)))
    V0, _ = np.linalg.qr(rng.standard_normal((n, r)))
    s = np.sort(rng.uniform(0.5, 2.0, r))[::-1]
    A = (U0 * s) @ V0.T
    U, S, V, k, te, dc = dev.svd_trunc(up(dev, A), m, n, cutoff=1e-14)
    assert k == r
    Uh = ops.from_flat(dev.to_host(U)[: m * k], (m, k))
    Vh = ops.from_flat(dev.to_host(V)[: n * k], (n, k))
    Sh = dev.to_host(S)[:k]
    np.testing.assert_allclose(Sh, s, rtol=1e-12)
    np.testing.assert_allclose(Uh @ np.diag(Sh) @ Vh.T, A, atol=1e-12)
    assert np.abs(Uh.T @ Uh - np.eye(k)).max() < 1e-12
    assert np.abs(Vh.T @ Vh - np.eye(k)).max() < 1e-12


def test_svd_truncation(dev):
    rng = np.random.default_rng(3)
    m, n = 60, 50
    U0, _ = np.linalg.qr(rng.standard_normal((m, n)))
    V0, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = 10.0 ** -np.linspace(0, 10, n)
    A = (U0 * s) @ V0.T
    U, S, V, k, te, dc = dev.svd_trunc(up(dev, A), m, n, maxdim=40, cutoff=1e-12)
    kr, ter, dcr = ndtensors_truncate(s ** 2, maxdim=40, cutoff=1e-12)
    assert k == kr
    assert math.isclose(te, ter, rel_tol=1e-6, abs_tol=1e-24)
    np.testing.assert_allclose(dev.to_host(S)[:k], s[:k], rtol=1e-10)


# ----------------------------------------------------------------------- qr
@pytest.mark.parametrize("m,n", [(20, 12), (12, 20), (64, 64), (300, 130), (1, 1)])
def test_qr(dev, m, n):
    """test/test_cuitensor.jl:125-130: A = QR, Q^T Q = I (atol 1e-14)."""
    rng = np.random.default_rng(m + n)
    A = rng.standard_normal((m, n))
    Q, R = dev.qr(up(dev, A), m, n)
    r = min(m, n)
    Qh = ops.from_flat(dev.to_host(Q), (m, r))
    Rh = ops.from_flat(dev.to_host(R), (r, n))
    np.testing.assert_allclose(Qh @ Rh, A, atol=1e-13 * max(1, np.abs(A).max()))
    assert np.abs(Qh.T @ Qh - np.eye(r)).max() < 1e-14 * 10
    assert np.allclose(np.tril(Rh, -1), 0)
    Qr, Rr = np.linalg.qr(A)
    np.testing.assert_allclose(np.abs(np.diag(Rh)), np.abs(np.diag(Rr)), rtol=1e-12)
