"""Gaussian alms of synfast (oracle, numpy) -- test infrastructure only.

Restates the draw of ``gmaster/utils.py:806-820`` (correlated complex Gaussian alms from a
per-l square root of the component covariance; real and imaginary parts scaled by
1/sqrt(2) for m > 0, imaginary part zero at m = 0) with the generator the device uses in
place of JAX's: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC11, 2011), key = the 64-bit
seed, counter = (element low, element high, draw, 0); each output gives two 53-bit
uniforms u1 in (0, 1], u2 in [0, 1) and two normals by Box-Muller.

Pinned in ``tests/test_oracle_synfast.py`` by the published Philox4x32-10 known-answer
vectors (Random123 ``kat_vectors``) and by the moments of the normals and of the alms.
"""

import numpy as np

M0, M1 = 0xD2511F53, 0xCD9E8D57
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK32 = 0xFFFFFFFF


def philox4x32_10(ctr, key):
    """``ctr``: (4, n) uint64 holding 32-bit words; ``key``: (2,) ints. Returns (4, n)."""
    c = [np.asarray(w, dtype=np.uint64) & np.uint64(MASK32) for w in ctr]
    k0, k1 = int(key[0]) & MASK32, int(key[1]) & MASK32
    for r in range(10):
        if r:
            k0 = (k0 + W0) & MASK32
            k1 = (k1 + W1) & MASK32
        p0 = np.uint64(M0) * c[0]
        p1 = np.uint64(M1) * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & np.uint64(MASK32)
        hi1, lo1 = p1 >> np.uint64(32), p1 & np.uint64(MASK32)
        c = [hi1 ^ c[1] ^ np.uint64(k0), lo1, hi0 ^ c[3] ^ np.uint64(k1), lo0]
    return c


def gauss_pairs(seed, elems, draw):
    """Normal pairs ``(z0, z1)`` of element indices ``elems`` for draw index ``draw``."""
    e = np.asarray(elems, dtype=np.uint64)
    ctr = [e & np.uint64(MASK32), e >> np.uint64(32), np.full_like(e, draw), np.zeros_like(e)]
    c = philox4x32_10(ctr, (seed & MASK32, (seed >> 32) & MASK32))
    a = ((c[0] << np.uint64(21)) ^ (c[1] >> np.uint64(11))) & np.uint64((1 << 53) - 1)
    b = ((c[2] << np.uint64(21)) ^ (c[3] >> np.uint64(11))) & np.uint64((1 << 53) - 1)
    u1 = (a.astype(np.float64) + 1.0) * 2.0**-53
    u2 = b.astype(np.float64) * 2.0**-53
    rad = np.sqrt(-2.0 * np.log(u1))
    return rad * np.cos(2 * np.pi * u2), rad * np.sin(2 * np.pi * u2)


def gaussian_alms(root, lmax, seed):
    """``(ncomp, nelem)`` complex alms; ``root`` (lmax+1, ncomp, ncomp)."""
    L = lmax + 1
    ncomp = root.shape[1]
    m = np.repeat(np.arange(L), L - np.arange(L))
    ell = np.concatenate([np.arange(mm, L) for mm in range(L)])
    n = len(ell)
    re = np.empty((n, ncomp))
    im = np.empty((n, ncomp))
    for i in range(ncomp):
        re[:, i], im[:, i] = gauss_pairs(seed, np.arange(n), i)
    R = root[ell]
    a = np.einsum("nij,nj->ni", R, re)
    b = np.einsum("nij,nj->ni", R, im)
    sc = np.where(m == 0, 1.0, np.sqrt(0.5))[:, None]
    return (a * sc + 1j * np.where(m[:, None] == 0, 0.0, b * sc)).T
