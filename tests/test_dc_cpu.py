"""CPU tier: the divide-and-conquer numerical core of libitg's eigensolver
(csrc/dc_core.h), driven on the CPU with the GPU's merge schedule, against
LAPACK (scipy eigh_tridiagonal)."""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest
from scipy.linalg import eigh_tridiagonal

from conftest import PKG

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def dc():
    out = os.path.join(tempfile.mkdtemp(), "libdc.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC",
                    "-I", os.path.join(PKG, "csrc"), os.path.join(HERE, "dc_harness.cpp"),
                    "-o", out], check=True)
    L = C.CDLL(out)
    L.dc_eig.argtypes = [C.c_int] + [C.c_void_p] * 4
    return L


def run(dc, d, e):
    n = d.size
    w = np.zeros(n)
    Q = np.zeros((n, n), order="F")
    dc.dc_eig(n, d.ctypes.data, e.ctypes.data, w.ctypes.data, Q.ctypes.data)
    return w, Q


def check(dc, d, e, tol=1e-13):
    w, Q = run(dc, d, e)
    wr = eigh_tridiagonal(d, e, eigvals_only=True)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    scale = max(np.abs(wr).max(), 1e-300)
    assert np.abs(w - wr).max() <= tol * scale * max(1, np.sqrt(d.size) / 8)
    n = d.size
    assert np.abs(Q.T @ Q - np.eye(n)).max() <= 50 * n * 1.1e-16
    assert np.abs(T @ Q - Q * w).max() <= 50 * n * 1.1e-16 * scale


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 17, 64, 100, 257])
def test_random_tridiagonal(dc, n):
    rng = np.random.default_rng(n)
    check(dc, rng.standard_normal(n), rng.standard_normal(max(n - 1, 0)))


def test_degenerate_and_decoupled(dc):
    n = 96
    d = np.repeat([1.0, 2.0, 3.0], 32)
    e = np.zeros(n - 1)
    e[::7] = 1e-3
    check(dc, d, e)
    # exact multiplets: identical blocks
    blk_d = np.array([2.0, -1.0, 0.5, 3.0])
    blk_e = np.array([1.0, 0.3, -0.7])
    d = np.tile(blk_d, 16)
    e = np.concatenate([np.concatenate([blk_e, [0.0]]) for _ in range(16)])[:-1]
    check(dc, d, e)


def test_graded_density_matrix_spectrum(dc):
    """Tridiagonal form of a PSD matrix whose eigenvalues decay over 20
    decades, as DMRG density matrices do."""
    rng = np.random.default_rng(7)
    n = 200
    lam = 10.0 ** (-np.linspace(0, 20, n))
    U, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = (U * lam) @ U.T
    from scipy.linalg import hessenberg
    H = hessenberg(A)
    d, e = np.diag(H).copy(), np.diag(H, -1).copy()
    w, Q = run(dc, d, e)
    wr = eigh_tridiagonal(d, e, eigvals_only=True)
    assert np.abs(w - wr).max() <= 1e-14 * wr.max() * 10
    assert np.abs(Q.T @ Q - np.eye(n)).max() < 1e-12


def test_rank_deficient_density_matrix(dc):
    """X Xᵀ with X of half rank: half the spectrum is a cluster at zero whose
    neighbours rotate into each other along long runs of the deflation
    pass (the shape of a two-site DMRG density matrix whose bond has grown
    past its rank)."""
    rng = np.random.default_rng(11)
    n, r = 512, 256
    X = rng.standard_normal((n, r)) / np.sqrt(r)
    from scipy.linalg import hessenberg
    H = hessenberg(X @ X.T)
    d, e = np.diag(H).copy(), np.diag(H, -1).copy()
    w, Q = run(dc, d, e)
    wr = eigh_tridiagonal(d, e, eigvals_only=True)
    assert np.abs(np.sort(w) - wr).max() <= 1e-13 * wr.max()
    assert np.abs(Q.T @ Q - np.eye(n)).max() < 1e-12
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    assert np.abs(T @ Q - Q * w).max() <= 1e-12 * wr.max()


def test_wilkinson(dc):
    m = 20
    d = np.abs(np.arange(-m, m + 1)).astype(float)
    e = np.ones(2 * m)
    check(dc, d, e)
