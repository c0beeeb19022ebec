"""Pins for the oracle's geometry and transforms by brute force (no GPU)."""

import math

import numpy as np
import pytest

from oracle import sht
from oracle.geometry import AlmInfo, Rings, pix2ang_ring
from oracle.legendre import wigner_d_explicit


@pytest.mark.parametrize("nside", [1, 2, 4, 8, 16])
def test_ring_geometry_matches_per_pixel_formulae(nside):
    R = Rings(nside)
    th, ph = pix2ang_ring(nside, np.arange(R.npix))
    assert R.nphi.sum() == R.npix
    assert np.all(R.offset[1:] == R.offset[:-1] + R.nphi[:-1])
    assert np.allclose(np.repeat(R.theta, R.nphi), th, atol=1e-14)
    phi = np.concatenate([R.phi0[r] + 2 * np.pi * np.arange(R.nphi[r]) / R.nphi[r] for r in range(R.nring)])
    assert np.allclose(phi, ph, atol=1e-13)
    assert np.allclose(R.theta[::-1], np.pi - R.theta, atol=1e-14)
    assert R.z[2 * nside - 1] == 0.0


def test_ring_dft_matches_direct_definition():
    rng = np.random.default_rng(0)
    R = Rings(4)
    x = rng.normal(size=R.npix)
    orders = np.arange(-11, 12)
    G = sht.phased_spectrum(x.astype(complex), R, orders)
    for r in range(R.nring):
        n, o = int(R.nphi[r]), int(R.offset[r])
        direct = sht.ring_dft_direct(x[o:o + n], n, orders)
        ref = R.weight * direct * np.exp(-1j * orders * R.phi0[r])
        assert np.allclose(G[r], ref, atol=1e-13)


@pytest.mark.parametrize("nside,lmax", [(2, 5), (4, 11)])
def test_spin0_matches_dense_sph_harm(nside, lmax):
    rng = np.random.default_rng(nside)
    npix = 12 * nside * nside
    Y = sht.dense_Y_s0(nside, lmax)
    x = rng.normal(size=npix)
    ref = (4 * np.pi / npix) * Y.conj().T @ x
    assert np.max(np.abs(sht.map2alm_s0(x, nside, lmax) - ref)) < 1e-13
    ai = AlmInfo(lmax)
    a = rng.normal(size=ai.nelem) + 1j * rng.normal(size=ai.nelem)
    a[ai.m == 0] = a[ai.m == 0].real
    mref = (Y @ a + np.conj(Y[:, ai.m > 0] @ a[ai.m > 0])).real
    assert np.max(np.abs(sht.alm2map_s0(a, nside, lmax) - mref)) < 1e-12


def test_spin2_matches_explicit_wigner_transform():
    """The catalogue-style direct sums of gmaster/utils.py:683-755 with Wigner's explicit d."""
    nside, lmax, s = 2, 5, 2
    npix = 12 * nside * nside
    ai = AlmInfo(lmax)
    L = lmax + 1
    th, ph = pix2ang_ring(nside, np.arange(npix))
    D = np.zeros((npix, L, 2 * L - 1))
    for ell in range(s, L):
        for m in range(-ell, ell + 1):
            for p in range(npix):
                D[p, ell, m + L - 1] = wigner_d_explicit(ell, m, -s, th[p])
    rng = np.random.default_rng(1)
    qu = rng.normal(size=(2, npix))
    sig = qu[0] + 1j * qu[1]
    full = np.zeros((L, 2 * L - 1), complex)
    for ell in range(s, L):
        for m in range(-ell, ell + 1):
            full[ell, m + L - 1] = math.sqrt((2 * ell + 1) / (4 * np.pi)) * np.sum(
                (4 * np.pi / npix) * sig * D[:, ell, m + L - 1] * np.exp(-1j * m * ph))
    E = np.zeros(ai.nelem, complex)
    B = np.zeros(ai.nelem, complex)
    for k in range(ai.nelem):
        ell, m = ai.ell[k], ai.m[k]
        plus, minus = full[ell, L - 1 + m], (-1) ** m * np.conj(full[ell, L - 1 - m])
        E[k], B[k] = -(plus + minus) / 2, 0.5j * (plus - minus)
    got = sht.map2alm_spin(qu, nside, lmax, s)
    assert np.max(np.abs(got[0] - E)) < 1e-13 and np.max(np.abs(got[1] - B)) < 1e-13
    ap = sht.eb_to_plus(got, ai)
    sigm = np.zeros(npix, complex)
    for ell in range(s, L):
        for m in range(-ell, ell + 1):
            sigm += math.sqrt((2 * ell + 1) / (4 * np.pi)) * D[:, ell, m + L - 1] * np.exp(1j * m * ph) * ap[ell, m + L - 1]
    mp = sht.alm2map_spin(got, nside, lmax, s)
    assert np.max(np.abs(mp[0] - sigm.real)) < 1e-12 and np.max(np.abs(mp[1] - sigm.imag)) < 1e-12


def test_jacobi_iterations_converge_for_band_limited_maps():
    nside, lmax = 8, 15
    ai = AlmInfo(lmax)
    rng = np.random.default_rng(3)
    a = rng.normal(size=ai.nelem) + 1j * rng.normal(size=ai.nelem)
    a[ai.m == 0] = a[ai.m == 0].real
    m = sht.alm2map_s0(a, nside, lmax)
    errs = [np.max(np.abs(sht.map2alm_s0(m, nside, lmax, n) - a)) for n in (0, 3, 10)]
    assert errs[0] > errs[1] > errs[2] and errs[2] < 1e-8
