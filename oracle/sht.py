"""Spin-0 and spin-s HEALPix transforms with Jacobi refinement (oracle, FP64).

Conventions (restating the reference):

* analysis   ``a_lm = (4 pi / npix) sum_p x_p conj(Y_lm(p))`` (``gmaster/utils.py:683-699``,
  Appendix "Spin-0 analysis"): per ring the azimuthal sum is the ring DFT
  ``sum_k x_k e^{-2 pi i k m / nphi}`` (``tests/test_sht.py:463-491``) times
  ``e^{-i m phi0}``.
* spin s     input ``Q + iU``; ``a+_lm = (-1)^s sqrt((2l+1)/4pi) sum w (Q+iU) d^l_{m,-s} e^{-i m phi}``,
  ``E = -(a+_m + (-1)^m conj a+_{-m}) / 2``, ``B = (i/2)(a+_m - (-1)^m conj a+_{-m})``
  (``gmaster/utils.py:701-707``, ``gmaster/_sht/healpix.py:806-814``); synthesis
  ``a+_m = -(E + iB)``, ``a+_{-m} = -(-1)^m (conj E + i conj B)``
  (``gmaster/_sht/healpix.py:599-611``) and maps ``(Re, Im)`` of
  ``sum a+_lm (-1)^s lambda_lm e^{i m phi}`` (``gmaster/utils.py:741-755``).
* Jacobi     ``alm <- alm - A(S alm - map)`` n_iter times (``gmaster/_sht/healpix.py:1459-1471``).

The oracle works entirely in map space (explicit inverse ring DFTs), so the HIP path's
ring-Fourier fold residual is checked against an independent formulation.

Pinned by: a dense brute-force ``Y`` matrix built from ``scipy.special.sph_harm_y``
(spin 0) and from Wigner's explicit sum (spin 2) at Nside 2-4, the direct per-ring DFT,
and the round trip ``map2alm(alm2map(a)) -> a`` (``tests/test_oracle_sht.py``).
"""

import numpy as np

from .geometry import AlmInfo, Rings
from .legendre import lambda_rows


def ring_dft_direct(ring_values, nphi, orders):
    """Direct ``sum_k x_k e^{-2 pi i k m / nphi}`` for the given orders (definition)."""
    k = np.arange(nphi)
    return np.exp(-2j * np.pi * np.outer(orders, k) / nphi) @ ring_values


def phased_spectrum(maps_complex, rings, orders):
    """``G[r, j] = w e^{-i m_j phi0_r} DFT_r[m_j mod nphi_r]`` for a complex or real map.

    ``orders`` may be negative.  Uses numpy's FFT (pinned against ``ring_dft_direct``).
    """
    out = np.zeros((rings.nring, len(orders)), dtype=np.complex128)
    orders = np.asarray(orders)
    for r in range(rings.nring):
        n = int(rings.nphi[r])
        o = int(rings.offset[r])
        spec = np.fft.fft(maps_complex[o:o + n])
        phase = np.exp(-1j * np.pi * (orders * rings.phi0_num[r] % (2 * rings.phi0_den[r])) / rings.phi0_den[r])
        out[r] = rings.weight * spec[orders % n] * phase
    return out


def ring_synthesis(F, rings, orders, npix):
    """Map ``x_p = sum_j F[r, j] e^{i m_j phi_p}`` on every ring (complex output)."""
    out = np.zeros(npix, dtype=np.complex128)
    orders = np.asarray(orders)
    for r in range(rings.nring):
        n = int(rings.nphi[r])
        o = int(rings.offset[r])
        phase = np.exp(1j * np.pi * (orders * rings.phi0_num[r] % (2 * rings.phi0_den[r])) / rings.phi0_den[r])
        X = np.zeros(n, dtype=np.complex128)
        np.add.at(X, orders % n, F[r] * phase)
        out[o:o + n] = n * np.fft.ifft(X)
    return out


def _lam_all(rings, L, m, s):
    return lambda_rows(L, m, s, rings.z, rings.shalf, rings.chalf)


def analysis_s0(G, rings, ainfo):
    """``a_lm = sum_r lambda_lm(theta_r) G[r, m]`` packed healpy-style."""
    L = ainfo.lmax + 1
    alm = np.zeros(ainfo.nelem, dtype=np.complex128)
    for m in range(L):
        lam = _lam_all(rings, L, m, 0)  # (L, nring)
        alm[ainfo.mstart[m] + np.arange(m, L)] = (lam @ G[:, m])[m:]
    return alm


def synthesis_s0(alm, rings, ainfo):
    """``F[r, m] = sum_l a_lm lambda_lm(theta_r)`` for m in [0, L)."""
    L = ainfo.lmax + 1
    F = np.zeros((rings.nring, L), dtype=np.complex128)
    for m in range(L):
        lam = _lam_all(rings, L, m, 0)
        a = np.zeros(L, dtype=np.complex128)
        a[m:] = alm[ainfo.mstart[m] + np.arange(m, L)]
        F[:, m] = a @ lam
    return F


def alm2map_s0(alm, nside, lmax):
    rings = Rings(nside)
    ainfo = AlmInfo(lmax)
    L = lmax + 1
    F = synthesis_s0(alm, rings, ainfo)
    # Hermitian extension: F[-m] = conj F[m]
    orders = np.concatenate([np.arange(L), -np.arange(1, L)])
    Ff = np.concatenate([F, np.conj(F[:, 1:])], axis=1)
    return ring_synthesis(Ff, rings, orders, rings.npix).real


def map2alm_s0(m_in, nside, lmax, n_iter=0):
    rings = Rings(nside)
    ainfo = AlmInfo(lmax)
    L = lmax + 1
    orders = np.arange(L)
    alm = analysis_s0(phased_spectrum(np.asarray(m_in, dtype=np.complex128), rings, orders), rings, ainfo)
    for _ in range(int(n_iter)):
        resid = alm2map_s0(alm, nside, lmax) - m_in
        alm = alm - analysis_s0(phased_spectrum(resid.astype(np.complex128), rings, orders), rings, ainfo)
    return alm


# ----------------------------------------------------------------------------------- spin s
def _mirror(rings):
    return rings.nring - 1 - np.arange(rings.nring)


def analysis_plus(G, rings, ainfo, s):
    """``a+[l, m]`` for m in (-L, L) as an ``(L, 2L-1)`` array (column ``m + L - 1``).

    Negative orders use ``d^l_{-m,-s}(theta) = (-1)^{l+s} d^l_{m,-s}(pi - theta)``, pinned
    separately against Wigner's explicit sum.
    """
    L = ainfo.lmax + 1
    mir = _mirror(rings)
    out = np.zeros((L, 2 * L - 1), dtype=np.complex128)
    sgn_s = (-1) ** s
    ell = np.arange(L)
    for m in range(L):
        lam = _lam_all(rings, L, m, s)  # (L, nring)
        out[:, L - 1 + m] = sgn_s * (lam @ G[:, L - 1 + m])
        if m:
            out[:, L - 1 - m] = sgn_s * ((-1.0) ** (ell + s)) * (lam[:, mir] @ G[:, L - 1 - m])
    return out


def synthesis_plus(ap, rings, ainfo, s):
    """``F+[r, m] = (-1)^s sum_l a+[l, m] lambda^{(s)}_{l,m}(theta_r)``, m in (-L, L)."""
    L = ainfo.lmax + 1
    mir = _mirror(rings)
    F = np.zeros((rings.nring, 2 * L - 1), dtype=np.complex128)
    sgn_s = (-1) ** s
    ell = np.arange(L)
    for m in range(L):
        lam = _lam_all(rings, L, m, s)
        F[:, L - 1 + m] = sgn_s * (ap[:, L - 1 + m] @ lam)
        if m:
            F[:, L - 1 - m] = sgn_s * ((ap[:, L - 1 - m] * (-1.0) ** (ell + s)) @ lam[:, mir])
    return F


def plus_to_eb(ap, ainfo, s):
    L = ainfo.lmax + 1
    E = np.zeros(ainfo.nelem, dtype=np.complex128)
    B = np.zeros(ainfo.nelem, dtype=np.complex128)
    for m in range(L):
        idx = ainfo.mstart[m] + np.arange(m, L)
        plus = ap[m:, L - 1 + m]
        minus = (-1) ** m * np.conj(ap[m:, L - 1 - m])
        E[idx] = -(plus + minus) / 2
        B[idx] = 0.5j * (plus - minus)
    ell = ainfo.ell
    E[ell < s] = 0
    B[ell < s] = 0
    return np.stack([E, B])


def eb_to_plus(alm, ainfo):
    L = ainfo.lmax + 1
    ap = np.zeros((L, 2 * L - 1), dtype=np.complex128)
    E, B = alm
    for m in range(L):
        idx = ainfo.mstart[m] + np.arange(m, L)
        ap[m:, L - 1 + m] = -(E[idx] + 1j * B[idx])
        if m:
            ap[m:, L - 1 - m] = -((-1) ** m) * (np.conj(E[idx]) + 1j * np.conj(B[idx]))
    return ap


def alm2map_spin(alm, nside, lmax, s):
    rings = Rings(nside)
    ainfo = AlmInfo(lmax)
    L = lmax + 1
    F = synthesis_plus(eb_to_plus(alm, ainfo), rings, ainfo, s)
    orders = np.arange(-(L - 1), L)
    x = ring_synthesis(F, rings, orders, rings.npix)
    return np.stack([x.real, x.imag])


def map2alm_spin(maps, nside, lmax, s, n_iter=0):
    rings = Rings(nside)
    ainfo = AlmInfo(lmax)
    L = lmax + 1
    orders = np.arange(-(L - 1), L)

    def once(qu):
        G = phased_spectrum(qu[0] + 1j * qu[1], rings, orders)
        return plus_to_eb(analysis_plus(G, rings, ainfo, s), ainfo, s)

    maps = np.asarray(maps, dtype=np.float64)
    alm = once(maps)
    for _ in range(int(n_iter)):
        alm = alm - once(alm2map_spin(alm, nside, lmax, s) - maps)
    return alm


def map2alm(maps, spin, nside, lmax, n_iter=0):
    """Public-shaped entry: ``(nmaps, npix)`` -> ``(nmaps, nelem)``."""
    maps = np.atleast_2d(np.asarray(maps, dtype=np.float64))
    if spin == 0:
        return map2alm_s0(maps[0], nside, lmax, n_iter)[None]
    return map2alm_spin(maps, nside, lmax, spin, n_iter)


def alm2map(alm, spin, nside, lmax):
    alm = np.atleast_2d(np.asarray(alm, dtype=np.complex128))
    if spin == 0:
        return alm2map_s0(alm[0], nside, lmax)[None]
    return alm2map_spin(alm, nside, lmax, spin)


# -------------------------------------------------------------------------- brute force
def dense_Y_s0(nside, lmax):
    """Dense ``Y[p, (l,m>=0)]`` from ``scipy.special.sph_harm_y`` (tiny Nside only)."""
    from scipy.special import sph_harm_y

    from .geometry import pix2ang_ring

    ainfo = AlmInfo(lmax)
    th, ph = pix2ang_ring(nside, np.arange(12 * nside * nside))
    Y = np.zeros((th.size, ainfo.nelem), dtype=np.complex128)
    for k in range(ainfo.nelem):
        Y[:, k] = sph_harm_y(ainfo.ell[k], ainfo.m[k], th, ph)
    return Y
