"""Pins of the synfast oracle (``oracle.synfast``), CPU only: Philox4x32-10 known-answer
vectors and the moments of the normals and of the alms."""

import numpy as np

from oracle import synfast as osf

# Random123 kat_vectors, philox4x32 with 10 rounds: (counter, key) -> output
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


def test_philox_known_answers():
    for ctr, key, want in KAT:
        got = osf.philox4x32_10([np.array([w], dtype=np.uint64) for w in ctr], key)
        assert [int(g[0]) for g in got] == list(want)


def test_normals_have_gaussian_moments():
    n = 400_000
    z0, z1 = osf.gauss_pairs(12345, np.arange(n), 0)
    z = np.concatenate([z0, z1])
    assert abs(z.mean()) < 5e-3
    assert abs(z.var() - 1) < 6e-3
    assert abs(np.mean(z**4) - 3) < 3e-2
    assert abs(np.mean(z0 * z1)) < 5e-3
    # different draws and seeds are independent streams
    w0, _ = osf.gauss_pairs(12345, np.arange(n), 1)
    v0, _ = osf.gauss_pairs(12346, np.arange(n), 0)
    assert abs(np.mean(z0 * w0)) < 5e-3 and abs(np.mean(z0 * v0)) < 5e-3
    # fraction beyond 2 sigma
    assert abs(np.mean(np.abs(z) > 2) - 0.0455) < 2e-3


def test_alms_follow_the_covariance():
    lmax = 200
    L = lmax + 1
    ell = np.arange(L)
    tt = 1.0 / (ell + 5.0) ** 2
    ee = 0.3 * tt
    te = 0.5 * np.sqrt(tt * ee)
    cov = np.zeros((L, 2, 2))
    cov[:, 0, 0], cov[:, 1, 1], cov[:, 0, 1], cov[:, 1, 0] = tt, ee, te, te
    w, v = np.linalg.eigh(cov)
    root = v * np.sqrt(w)[:, None, :]
    alm = osf.gaussian_alms(root, lmax, 7)
    m = np.repeat(np.arange(L), L - np.arange(L))
    lidx = np.concatenate([np.arange(mm, L) for mm in range(L)])
    assert np.all(alm[:, m == 0].imag == 0)
    wt = np.where(m == 0, 1.0, 2.0)
    for a, b, want in ((0, 0, tt), (1, 1, ee), (0, 1, te)):
        cl = np.bincount(lidx, weights=wt * (alm[a] * np.conj(alm[b])).real, minlength=L) / (2 * ell + 1)
        ratio = cl[100:] / want[100:]
        # cosmic variance ~ sqrt(2 / (2l + 1)) per l; averaged over 101 multipoles
        assert abs(ratio.mean() - 1) < 0.02
