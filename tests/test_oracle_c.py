"""The C/OpenMP oracle (CPU baseline) agrees with the numpy oracle (no GPU)."""

import os
import subprocess

import numpy as np
import pytest

from oracle import coupling, sht
from oracle.geometry import Rings

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def cb():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle", "c")], check=True)
    from oracle.c import cpu_baseline

    return cpu_baseline


def test_c_legendre_analysis_matches_numpy_oracle(cb):
    nside, lmax, s = 8, 23, 2
    L = lmax + 1
    rng = np.random.default_rng(0)
    R = Rings(nside)
    qu = rng.normal(size=(2, R.npix))
    orders = np.arange(-(L - 1), L)
    G = sht.phased_spectrum(qu[0] + 1j * qu[1], R, orders)
    ap = sht.analysis_plus(G, R, sht.AlmInfo(lmax), s)  # (L, 2L-1)
    out = cb.lat_ana_s(nside, lmax, s, G, np.arange(L))
    ell = np.arange(L)
    for m in range(L):
        direct = (-1) ** s * (out[m, :, 0] + 1j * out[m, :, 1])
        assert np.allclose(direct[m:], ap[m:, L - 1 + m], atol=1e-12)
        if m:
            mirror = (-1) ** s * (-1.0) ** (ell + s) * (out[m, :, 2] + 1j * out[m, :, 3])
            assert np.allclose(mirror[m:], ap[m:, L - 1 - m], atol=1e-12)


def test_c_coupling_matches_quadrature(cb):
    lmax = 40
    w = np.random.default_rng(4).uniform(size=lmax + 1)
    q = coupling.coupling_blocks(w, lmax)
    raw = cb.coupling_rows(lmax, w / (4 * np.pi), np.arange(lmax + 1))
    col = 2 * np.arange(lmax + 1) + 1
    for k in range(4):
        M = np.zeros((lmax + 1, lmax + 1))
        for l1 in range(lmax + 1):
            M[l1, l1:] = raw[l1, l1:, k] * col[l1:]
            M[l1:, l1] = raw[l1, l1:, k] * col[l1]
        # the quadrature oracle carries ~1e-12 round-off relative to max |M| (the l3 recursion
        # matches exact Racah sums; see tests/test_gpu_master.py)
        assert np.max(np.abs(M - q[k])) < 1e-11 * np.max(np.abs(q[k]))


def test_n3j_count():
    from oracle.c.cpu_baseline import n3j

    assert n3j(3) == sum(min(a + b, 3) - (b - a) + 1 for a in range(4) for b in range(a, 4))
    assert abs(n3j(12287) - 4.64e11) / 4.64e11 < 0.01


def test_column_entry_matches_full_spectrum_entry():
    """``lat_ana_s_cols`` (per-order columns, for spectra too large for the host) gives
    exactly what ``lat_ana_s`` gives from the full spectrum."""
    import subprocess
    import os
    subprocess.run(["make", "-s", "-C", os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                                      "oracle", "c")], check=True)
    from oracle.c import cpu_baseline as cb
    nside, lmax, spin = 8, 20, 2
    nring, L = 4 * nside - 1, lmax + 1
    rng = np.random.default_rng(3)
    spec = rng.normal(size=(nring, 2 * L - 1)) + 1j * rng.normal(size=(nring, 2 * L - 1))
    orders = np.array([0, 1, 5, 13, 20])
    full = cb.lat_ana_s(nside, lmax, spin, spec, orders)
    cols = cb.lat_ana_s_cols(nside, lmax, spin, spec[:, L - 1 + orders].T, spec[:, L - 1 - orders].T, orders)
    assert np.array_equal(full, cols)
