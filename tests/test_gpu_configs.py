"""GPU parity at the workloads' own sizes: the bulk-bond density-matrix
eigensolver (n = 4096), the effective-Hamiltonian microbench's truncated SVD
of the 2048 x 2048 result, the 4-leg cylinder (MPO bond dimension 14) at
small chi against exact diagonalisation and the oracle's identical run, and
the w = 14 effective Hamiltonian at the chi = 3000 shape on a sub-block."""
import time

import numpy as np
import pytest

from oracle import dmrg as odmrg
from oracle import linalg as olin
from oracle import models
from oracle import tensor_ops as ops
from oracle.truncate import ndtensors_truncate

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import itg
    from itg import ops as d
    itg._lib.ctx()
    return d


def up(dev, a):
    return dev.to_device(ops.to_flat(a))


# --------------------------------------------------- bulk-bond eigensolver
@pytest.mark.parametrize("cols", [4096, 2048])
def test_eigh_bulk_bond_n4096_vs_lapack(dev, cols):
    """The density matrix of a bulk bond at chi = 2048 (n = d chi = 4096),
    maxdim 2048, cutoff 1e-12, as every bulk bond of the N=100 sweep: rho =
    M M^T / |M|^2 with M n x cols Gaussian. cols = 4096 is full rank;
    cols = 2048 leaves a 2048-fold cluster at zero (the sweep's density
    matrices have such clusters, which deflate in the divide and conquer).
    Kept eigenvalues, k and the truncation error against LAPACK (syevd);
    orthonormality and the eigen-residual of the kept vectors."""
    n, maxdim = 4096, 2048
    rng = np.random.default_rng(cols)
    M = rng.standard_normal((n, cols))
    M /= np.linalg.norm(M)
    rho = M @ M.T
    t0 = time.perf_counter()
    w, U, k, te, dc = dev.eigh_trunc(up(dev, rho), n, maxdim=maxdim, cutoff=1e-12)
    t_gpu = time.perf_counter() - t0
    wr, Ur, ter, dcr = olin.eigh_trunc(rho, maxdim=maxdim, cutoff=1e-12)
    print("eigh n=%d cols=%d: k=%d (LAPACK %d), truncerr %.6e (%.6e), %.1f ms"
          % (n, cols, k, Ur.shape[1], te, ter, t_gpu * 1e3))
    assert k == Ur.shape[1]
    scale = wr[0]
    wh = dev.to_host(w)[:k]
    assert np.abs(wh - wr).max() <= 1e-13 * scale
    assert abs(te - ter) <= n * 1e-16 * scale + 1e-12 * ter
    Uh = ops.from_flat(dev.to_host(U)[: n * k], (n, k))
    assert np.abs(Uh.T @ Uh - np.eye(k)).max() < 1e-12
    assert np.abs(rho @ Uh - Uh * wh).max() < 1e-13 * scale


# ------------------------------------------- microbench: the truncated SVD
def test_config2_chain_and_truncated_svd(dev):
    """Effective-Hamiltonian microbench: L(1024,5,1024), W1, W2 (5,5,2,2),
    psi(1024,2,2,1024), R(1024,5,1024) drawn in that order from one seeded
    stream; the fixed-order product in FP64 and FP32, then the truncated SVD
    of the (chi d) x (d chi) = 2048 x 2048 result at cutoff 1e-10, maxdim
    1024 against LAPACK: k and truncation error as truncate! decides them
    on LAPACK's values, kept singular values to 1e-12 relative, orthonormal
    factors, and the truncated reconstruction error equal to the discarded
    weight."""
    chi, w, d = 1024, 5, 2
    rng = np.random.default_rng(2)
    L = rng.standard_normal((chi, w, chi))
    W1 = rng.standard_normal((w, w, d, d))
    W2 = rng.standard_normal((w, w, d, d))
    psi = rng.standard_normal((chi, d, d, chi))
    R = rng.standard_normal((chi, w, chi))
    want = odmrg.heff_apply(L, W1, W2, R, psi)
    m = chi * d
    A = np.reshape(want, (m, m), order="F")
    ys = {}
    for mode, tol in (("native", 1e-13), ("fp32", 2e-5)):
        y = dev.heff2_apply(up(dev, L), up(dev, W1), up(dev, W2), up(dev, R), up(dev, psi),
                            chi, w, w, w, d, chi, mode=mode)
        err = np.linalg.norm(dev.to_host(y) - ops.to_flat(want)) / np.linalg.norm(want)
        print("config2 Heff %s rel err %.2e" % (mode, err))
        assert err < tol
        ys[mode] = y
    t0 = time.perf_counter()
    U, S, V, k, te, dc = dev.svd_trunc(ys["native"], m, m, maxdim=1024, cutoff=1e-10)
    t_gpu = time.perf_counter() - t0
    s = np.linalg.svd(A, compute_uv=False)
    kr, ter, _ = ndtensors_truncate(s ** 2, maxdim=1024, cutoff=1e-10)
    print("config2 SVD 2048x2048: k=%d (LAPACK %d), truncerr %.6e (%.6e), %.2f s"
          % (k, kr, te, ter, t_gpu))
    assert k == kr == 1024
    assert abs(te - ter) <= 1e-10 * ter
    Sh = dev.to_host(S)[:k]
    np.testing.assert_allclose(Sh, s[:k], rtol=1e-12)
    Uh = ops.from_flat(dev.to_host(U)[: m * k], (m, k))
    Vh = ops.from_flat(dev.to_host(V)[: m * k], (m, k))
    assert np.abs(Uh.T @ Uh - np.eye(k)).max() < 1e-12
    assert np.abs(Vh.T @ Vh - np.eye(k)).max() < 1e-12
    resid = np.linalg.norm(A - (Uh * Sh) @ Vh.T) ** 2
    assert abs(resid - np.sum(s[k:] ** 2)) <= 1e-10 * np.sum(s ** 2)


# ------------------------------------------------ the 4-leg cylinder (w=14)
def cylinder(Lx, Ly):
    from itg.mps import MPO, cylinder_bonds, heisenberg_opsum
    bonds = cylinder_bonds(Lx, Ly)
    os_ = heisenberg_opsum(bonds)
    N = Lx * Ly
    return bonds, N, MPO.from_opsum(os_, N), os_.to_mpo_arrays(N)


def test_cylinder_4x4_dmrg_vs_ed_and_oracle(dev):
    """J1 Heisenberg on a 4 x 4 cylinder (periodic in y, snake order, MPO
    bond dimension 14): FP64 DMRG from a random MPS (seed 2), maxdim ramp to
    256 (the largest the 16 sites need), cutoff 1e-10, ends within 1e-8 of
    the exact ground energy and of the oracle's identical run; split-BF16
    and FP32 contractions end within 1e-9 per site of FP64."""
    from itg.mps import Sweeps, dmrg, random_mps
    bonds, N, H, Ws = cylinder(4, 4)
    assert max(W.shape[1] for W in Ws) == 14
    E0 = models.ground_energy(models.bond_hamiltonian(N, bonds))
    maxd = [16, 32, 64, 128, 256, 256, 256]
    sw = Sweeps(len(maxd), maxdim=maxd, cutoff=[1e-10])
    rec = []
    E = dmrg(H, random_mps(N, 2, 10, seed=2), sw, krylovdim=3, record=rec)
    reco = []
    Eo = odmrg.dmrg(Ws, odmrg.random_mps(N, 2, 10, seed=2), maxdims=maxd,
                    cutoffs=[1e-10] * len(maxd), krylovdim=3, record=reco)
    for r, ro in zip(rec, reco):
        print("cylinder 4x4 sweep: device %.12f oracle %.12f maxerr %.2e"
              % (r["energy_davidson"], ro["energy"], r["maxerr"]))
    print("cylinder 4x4: E %.12f oracle %.12f ED %.12f" % (E, Eo, E0))
    assert E0 - 1e-9 <= E < E0 + 1e-8
    assert abs(E - Eo) < 1e-8
    for mode in ("bf16x3", "fp32"):
        Em = dmrg(H, random_mps(N, 2, 10, seed=2), sw, krylovdim=3, mode=mode)
        print("cylinder 4x4 %s: |dE|/N %.2e" % (mode, abs(Em - E) / N))
        assert abs(Em - E) / N < 1e-9


def test_cylinder_4x16_w14_short_run(dev):
    """The 4 x 16 cylinder itself (64 sites, w = 14), random MPS seed 2,
    three sweeps at maxdim 24, 48, 48, cutoff 1e-10. The device MPS carries
    the energy the device reports (the oracle's <psi|H|psi> of the copied-
    back MPS, to 1e-10); energies fall sweep by sweep; the oracle's identical
    run and the split-BF16 / FP32 runs end within 1e-6 per site of it (at
    chi = 48 the runs are not converged, and truncation inside near-
    degenerate SU(2) multiplets picks different kept vectors from rounding-
    level differences, so trajectories agree only to this level)."""
    from itg.mps import Sweeps, dmrg, random_mps
    bonds, N, H, Ws = cylinder(16, 4)
    assert N == 64 and max(W.shape[1] for W in Ws) == 14
    maxd = [24, 48, 48]
    sw = Sweeps(3, maxdim=maxd, cutoff=[1e-10])
    res = {}
    for mode in ("native", "bf16x3", "fp32"):
        rec = []
        psi = random_mps(N, 2, 10, seed=2)
        res[mode] = dmrg(H, psi, sw, krylovdim=3, mode=mode, record=rec)
        print("cylinder 4x16 %s:" % mode,
              [("%.10f" % r["energy"], "maxerr %.1e" % r["maxerr"]) for r in rec])
        if mode == "native":
            Eh = odmrg.mps_energy(Ws, psi.to_host())
            assert abs(Eh - res[mode]) < 1e-10 * abs(Eh)
            es = [r["energy"] for r in rec]
            assert all(b < a for a, b in zip(es, es[1:]))
    reco = []
    Eo = odmrg.dmrg(Ws, odmrg.random_mps(N, 2, 10, seed=2), maxdims=maxd, cutoffs=[1e-10] * 3,
                    krylovdim=3, record=reco)
    print("cylinder 4x16 oracle %.10f" % Eo)
    assert abs(res["native"] - Eo) / N < 1e-6
    for mode in ("bf16x3", "fp32"):
        assert abs(res[mode] - res["native"]) / N < 1e-6


def test_heff_w14_chi3000_subblock(dev):
    """The cylinder's effective Hamiltonian at its largest shape: chi = 3000
    on both sides, w = 14, d = 2 (L and R 1 GB each in FP64). The device
    product in each arithmetic mode against the oracle's fixed-order product
    on three output rows a' (every s1', s2', b'): FP64 to 1e-12, FP32 and
    split-BF16 to 1e-4 relative."""
    chi, w, d = 3000, 14, 2
    rng = np.random.default_rng(14)
    L = rng.standard_normal((chi, w, chi))
    R = rng.standard_normal((chi, w, chi))
    W1 = rng.standard_normal((w, w, d, d))
    W2 = rng.standard_normal((w, w, d, d))
    psi = rng.standard_normal((chi, d, d, chi))
    rows = np.array([0, 1234, chi - 1])
    want = odmrg.heff_apply(L[rows], W1, W2, R, psi)          # (3, d, d, chi)
    Ld, Rd, W1d, W2d, pd = up(dev, L), up(dev, R), up(dev, W1), up(dev, W2), up(dev, psi)
    del L, R
    for mode, tol in (("native", 1e-12), ("fp32", 1e-4), ("bf16x3", 1e-4)):
        y = dev.heff2_apply(Ld, W1d, W2d, Rd, pd, chi, w, w, w, d, chi, mode=mode)
        got = ops.from_flat(dev.to_host(y), (chi, d, d, chi))[rows]
        err = np.linalg.norm(got - want) / np.linalg.norm(want)
        print("w=14 chi=3000 Heff %s: rel err %.2e on rows %s" % (mode, err, rows.tolist()))
        assert err < tol
        del y
