"""Geometry, transforms and input helpers (the counterpart of ``pymaster.utils``).

Mirrors ``gmaster/utils.py``: ``NmtMapInfo`` / ``NmtAlmInfo`` (``:336-524``),
``map2alm`` / ``alm2map`` of any spin (``:556-672``), C1/C2 ``mask_apodization``
(``:61-147``) and ``synfast_spherical`` (``:823-927``).  HEALPix only: CAR maps and
catalogue fields are outside this engine.
"""

import math

import numpy as np
import torch

from . import sht
from ._array import device_tensor, wrap
from ._config import (  # noqa: F401  (re-exported)
    NmtParams,
    get_default_params,
    nmt_params,
    set_n_iter_default,
    set_sht_calculator,
    set_tol_pinv_default,
)


class NmtMapInfo:
    """HEALPix map geometry (``wcs=None``, ``axes=(npix,)``)."""

    def __init__(self, wcs, axes):
        if wcs is not None:
            raise NotImplementedError("CAR (WCS) maps are not supported by this engine; use HEALPix")
        npix = int(axes[0])
        nside = 2
        while 12 * nside * nside != npix:
            nside *= 2
            if nside > 65536:
                raise ValueError("Something is wrong with your input arrays")
        self.is_healpix = True
        self.nside = nside
        self.npix = npix

    def __eq__(self, other):
        return isinstance(other, NmtMapInfo) and self.nside == other.nside

    def get_lmax(self):
        return 3 * self.nside - 1

    def reform_map(self, maps):
        if maps.shape[-1] != self.npix:
            raise ValueError("Incompatible map!")
        return maps


class NmtAlmInfo:
    """healpy m-major packing, ``a_lm`` at ``mstart[m] + l`` (``gmaster/utils.py:493-524``)."""

    def __init__(self, lmax):
        self.lmax = int(lmax)
        self.mmax = self.lmax
        m = np.arange(self.mmax + 1)
        self.mstart = (m * (2 * self.lmax + 1 - m) // 2).astype(np.uint64)
        self.nelem = (self.lmax + 1) * (self.lmax + 2) // 2

    def __eq__(self, other):
        return isinstance(other, NmtAlmInfo) and self.lmax == other.lmax

    @property
    def _ell(self):
        L = self.lmax + 1
        return np.concatenate([np.arange(mm, L) for mm in range(L)])

    @property
    def _m(self):
        L = self.lmax + 1
        return np.repeat(np.arange(L), L - np.arange(L))


def map2alm(map, spin, map_info, alm_info, *, n_iter):
    """Spherical-harmonic analysis ``(nmaps, npix) -> (nmaps, nelem)`` with ``n_iter`` Jacobi steps."""
    nmaps = 1 if spin == 0 else 2
    maps = device_tensor(map)
    if maps.ndim != 2 or tuple(maps.shape) != (nmaps, map_info.npix):
        raise ValueError("Input map has wrong shape")
    if spin < 0 or spin > alm_info.lmax:
        raise ValueError("spin must satisfy 0 <= spin <= lmax")
    return wrap(sht.map2alm(maps, int(spin), alm_info.lmax, n_iter=int(n_iter)))


def map2alm_pair(map_a, map_b, map_info, alm_info, *, n_iter):
    """Two spin-0 maps analysed in one shared latitudinal pass (``gmaster/utils.py:608-642``)."""
    a = device_tensor(map_a)
    b = device_tensor(map_b)
    if a.shape != (1, map_info.npix) or b.shape != (1, map_info.npix):
        raise ValueError("shared-band pair expects two (1, npix) spin-0 maps")
    out = sht.map2alm_s0(torch.cat([a, b]), alm_info.lmax, n_iter=int(n_iter))
    return wrap(out[0:1]), wrap(out[1:2])


def alm2map(alm, spin, map_info, alm_info):
    """Spherical-harmonic synthesis ``(nmaps, nelem) -> (nmaps, npix)``."""
    nmaps = 1 if spin == 0 else 2
    a = device_tensor(alm, dtype=torch.complex128)
    if a.ndim != 2 or tuple(a.shape) != (nmaps, alm_info.nelem):
        raise ValueError("Input alm has wrong shape")
    if spin < 0 or spin > alm_info.lmax:
        raise ValueError("spin must satisfy 0 <= spin <= lmax")
    return wrap(sht.alm2map(a, int(spin), map_info.nside, alm_info.lmax))


# ---------------------------------------------------------------------------- inputs
def _pix2vec(nside):
    """Unit vectors of every RING pixel (HEALPix closed forms), host FP64."""
    out = np.empty((12 * nside * nside, 3))
    for i in range(1, 4 * nside):
        north = 4 * nside - i if i > 2 * nside else i
        if north < nside:
            z = 1.0 - north * north / (3.0 * nside * nside)
            nphi, phi0 = 4 * north, math.pi / (4 * north)
            off = 2 * i * (i - 1) if i < nside else 12 * nside * nside - 2 * north * (north + 1)
        else:
            z = 2.0 * (2 * nside - north) / (3.0 * nside)
            nphi = 4 * nside
            phi0 = math.pi / (4 * nside) if (north - nside) % 2 == 0 else 0.0
            off = 2 * nside * (nside - 1) + (i - nside) * 4 * nside
        if i > 2 * nside:
            z = -z
        phi = phi0 + 2 * math.pi * np.arange(nphi) / nphi
        st = math.sqrt(max(0.0, (1 - z) * (1 + z)))
        out[off:off + nphi, 0] = st * np.cos(phi)
        out[off:off + nphi, 1] = st * np.sin(phi)
        out[off:off + nphi, 2] = z
    return out


def mask_apodization(mask_in, aposize, apotype="C1"):
    """C1 / C2 apodization of a HEALPix mask (``gmaster/utils.py:61-147``).

    ``x = sqrt((1 - cos d) / (1 - cos aposize))`` with ``d`` the angle to the nearest
    masked pixel; C1 multiplies by ``x - sin(2 pi x)/(2 pi)``, C2 by ``(1 - cos pi x)/2``.
    "Smooth" needs healpy's ``query_disc`` pixel sets and is not provided.
    """
    if apotype not in ("C1", "C2", "Smooth"):
        raise ValueError(f"Apodization type {apotype} unknown. Choose from ['C1', 'C2', 'Smooth']")
    if apotype == "Smooth":
        raise NotImplementedError("Smooth apodization is not provided by this engine")
    if aposize < 0:
        raise ValueError("Apodization scale must be a positive number")
    mask = np.asarray(torch.as_tensor(mask_in).cpu() if isinstance(mask_in, torch.Tensor) else mask_in,
                      dtype=np.float64)
    if mask.ndim != 1:
        raise ValueError("Mask must be a one-dimensional HEALPix map")
    if aposize == 0:
        return mask.copy()
    from scipy.spatial import cKDTree

    nside = int(round(math.sqrt(mask.size / 12)))
    radius = math.radians(aposize)
    if int(4 * mask.size * (1 - math.cos(1.2 * radius))) < 2:
        raise ValueError("Your apodization scale is too small for this pixel size")
    masked = np.flatnonzero(mask <= 0)
    if not len(masked):
        return mask.copy()
    vec = _pix2vec(nside)
    dist, _ = cKDTree(vec[masked]).query(vec)
    dm = 0.5 * dist**2
    thr = 1 - math.cos(radius)
    aff = (mask > 0) & (dm < thr)
    x = np.sqrt(np.maximum(dm[aff], 0) / thr)
    fac = x - np.sin(2 * np.pi * x) / (2 * np.pi) if apotype == "C1" else 0.5 * (1 - np.cos(np.pi * x))
    out = mask.copy()
    out[aff] *= fac
    return out


def synfast_spherical(nside, cls, spin_arr, beam=None, seed=-1, wcs=None, lmax=None):
    """Correlated Gaussian full-sky fields (``gmaster/utils.py:823-927``).

    ``cls`` holds the upper triangle of the field-component covariance, row-major over
    components (T = 1, spin-s = E, B).  Alms are drawn with NumPy (seeded), maps made by
    this engine's ``alm2map``.  Statistically equivalent to pymaster's, not bit-equal.
    """
    if wcs is not None:
        raise NotImplementedError("CAR (WCS) maps are not supported")
    spin_arr = [int(s) for s in spin_arr]
    ncomp = sum(1 if s == 0 else 2 for s in spin_arr)
    cls = np.atleast_2d(np.asarray(cls, dtype=np.float64))
    if len(cls) != ncomp * (ncomp + 1) // 2:
        raise ValueError("Wrong number of power spectra for these spins")
    lmax = 3 * nside - 1 if lmax is None else int(lmax)
    L = lmax + 1
    cov = np.zeros((L, ncomp, ncomp))
    k = 0
    for i in range(ncomp):
        for j in range(i, ncomp):
            cov[:, i, j] = cov[:, j, i] = cls[k, :L]
            k += 1
    if beam is not None:
        b = np.atleast_2d(beam)
        comp_beam = np.concatenate([[b[f]] * (1 if s == 0 else 2) for f, s in enumerate(spin_arr)])
        cov = cov * comp_beam[:, None, :L].T[:, :, None] * comp_beam[None, :, :L].T[:, None, :]
    # per-l square root (eigendecomposition, PSD clipping)
    w, v = np.linalg.eigh(cov)
    root = v * np.sqrt(np.clip(w, 0, None))[:, None, :]
    ai = NmtAlmInfo(lmax)
    ell, m = ai._ell, ai._m
    rng = np.random.default_rng(None if seed < 0 else seed)
    re = rng.standard_normal((ai.nelem, ncomp))
    im = rng.standard_normal((ai.nelem, ncomp))
    re = np.einsum("aij,aj->ai", root[ell], re)
    im = np.einsum("aij,aj->ai", root[ell], im)
    scale = np.where(m == 0, 1.0, 1 / np.sqrt(2))[:, None]
    alm = (re * scale + 1j * np.where(m[:, None] == 0, 0, im * scale)).T
    minfo = NmtMapInfo(None, (12 * nside * nside,))
    maps = []
    c = 0
    for s in spin_arr:
        n = 1 if s == 0 else 2
        maps.append(alm2map(alm[c:c + n], s, minfo, ai))
        c += n
    return wrap(torch.cat([mm.as_subclass(torch.Tensor) for mm in maps]))
