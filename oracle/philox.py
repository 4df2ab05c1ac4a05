"""Philox4x32-10 counter-based RNG (Salmon, Moraes, Dror, Shaw, SC'11) and the
Gaussian stream built on it.

Replaces CURAND `randn!` (src/cuitensor.jl:37,43-44; src/mps/cumps.jl:47) with
a stream that host and device compute identically up to libm rounding:
element g (g = offset + i) uses counter (g // 2, 0, 0, 0) and key
(lo32(seed), hi32(seed)); the four 32-bit outputs x0..x3 form two 53-bit
uniforms u1 = ((x1<<32|x0)>>11 + 0.5)*2^-53, u2 likewise from (x3, x2), and
Box-Muller gives z0 = r cos(2 pi u2), z1 = r sin(2 pi u2), r = sqrt(-2 ln u1);
element g is z0 if g is even, z1 if odd.
"""
import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint64(0x9E3779B9)
W1 = np.uint64(0xBB67AE85)
MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr, key):
    """ctr: (4, n) uint64 holding 32-bit words; key: (2,) or (2, n)."""
    c0, c1, c2, c3 = [np.asarray(c, dtype=np.uint64) & MASK for c in ctr]
    k0 = np.asarray(key[0], dtype=np.uint64) & MASK
    k1 = np.asarray(key[1], dtype=np.uint64) & MASK
    for r in range(10):
        if r > 0:
            k0 = (k0 + W0) & MASK
            k1 = (k1 + W1) & MASK
        p0 = M0 * c0
        p1 = M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK
        c0, c1, c2, c3 = (hi1 ^ c1 ^ k0) & MASK, lo1, (hi0 ^ c3 ^ k1) & MASK, lo0
    return c0, c1, c2, c3


def randn(n, seed, offset=0):
    """n Gaussian doubles of the stream keyed by `seed`, starting at `offset`."""
    g = np.arange(offset, offset + n, dtype=np.uint64)
    ctr = g >> np.uint64(1)
    zero = np.zeros_like(ctr)
    seed = np.uint64(seed)
    key = (seed & MASK, seed >> np.uint64(32))
    x0, x1, x2, x3 = philox4x32_10((ctr & MASK, ctr >> np.uint64(32), zero, zero), key)
    a = ((x1 << np.uint64(32)) | x0) >> np.uint64(11)
    b = ((x3 << np.uint64(32)) | x2) >> np.uint64(11)
    u1 = (a.astype(np.float64) + 0.5) * 2.0 ** -53
    u2 = (b.astype(np.float64) + 0.5) * 2.0 ** -53
    r = np.sqrt(-2.0 * np.log(u1))
    th = 2.0 * np.pi * u2
    odd = (g & np.uint64(1)).astype(bool)
    return np.where(odd, r * np.sin(th), r * np.cos(th))


def randn_complex(n, seed, offset=0):
    """n complex Gaussians of unit variance (randn(ComplexF64)): element i is
    (g[2(offset+i)] + i g[2(offset+i)+1]) / sqrt(2) over the real stream g."""
    g = randn(2 * n, seed, 2 * offset) * 0.70710678118654752440
    return g[0::2] + 1j * g[1::2]
