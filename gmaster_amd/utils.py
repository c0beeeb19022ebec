"""Geometry, transforms and input helpers (the counterpart of ``pymaster.utils``).

Mirrors ``gmaster/utils.py``: ``NmtMapInfo`` / ``NmtAlmInfo`` (``:336-524``),
``map2alm`` / ``alm2map`` of any spin (``:556-672``), C1/C2 ``mask_apodization``
(``:61-147``) and ``synfast_spherical`` (``:823-927``).  HEALPix only: CAR maps and
catalogue fields are outside this engine.
"""

import math

import numpy as np
import torch

from . import _lib, sht
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
def mask_apodization(mask_in, aposize, apotype="C1"):
    """C1 / C2 apodization of a HEALPix RING mask on the device (``gmaster/utils.py:61-147``).

    ``x = sqrt((1 - cos d) / (1 - cos aposize))`` with ``d`` the angle to the nearest
    masked pixel (mask <= 0); C1 multiplies by ``x - sin(2 pi x)/(2 pi)``, C2 by
    ``(1 - cos pi x)/2``.  The nearest-masked search is ``gm_mask_apodize``'s ring search.
    "Smooth" zeroes the pixels within 2.5 aposize (plus the largest pixel radius, healpy's
    inclusive disc rule) of a masked pixel with the same search, then smooths through the
    engine's FP64 transforms.
    """
    if apotype not in ("C1", "C2", "Smooth"):
        raise ValueError(f"Apodization type {apotype} unknown. Choose from ['C1', 'C2', 'Smooth']")
    if aposize < 0:
        raise ValueError("Apodization scale must be a positive number")
    mask = device_tensor(mask_in)
    if mask.ndim != 1:
        raise ValueError("Mask must be a one-dimensional HEALPix map")
    if aposize == 0:
        return wrap(mask.clone())
    npix = mask.numel()
    nside = NmtMapInfo(None, (npix,)).nside
    radius = math.radians(aposize)
    scale = 2.5 if apotype == "Smooth" else 1.2
    if int(4 * npix * (1 - math.cos(scale * radius))) < 2:
        raise ValueError("Your apodization scale is too small for this pixel size")
    mask = mask.contiguous()
    out = torch.empty_like(mask)
    work = torch.empty(int(_lib.query("gm_apodize_work_size", nside)), dtype=torch.uint8, device=mask.device)
    if apotype != "Smooth":
        _lib.call("gm_mask_apodize", _lib.ptr(mask), nside, radius, 1 if apotype == "C1" else 2, _lib.ptr(out),
                  _lib.ptr(work), _lib.stream_ptr(mask.device))
        return wrap(out)
    # "Smooth": zero every pixel within 2.5 aposize of a masked pixel, widened by the largest
    # pixel radius as healpy's inclusive query_disc does, then smooth with a Gaussian beam of
    # width aposize and multiply by the mask (gmaster/utils.py:119-134)
    if not bool((mask <= 0).any()):
        return wrap(mask.clone())
    rbig = min(2.5 * radius + _max_pixrad(nside), math.pi)
    _lib.call("gm_mask_apodize", _lib.ptr(mask), nside, rbig, 3, _lib.ptr(out), _lib.ptr(work),
              _lib.stream_ptr(mask.device))
    lmax = 3 * nside - 1
    alm = sht.map2alm_s0(out[None, :], lmax, n_iter=3, precision="fp64")
    ell = torch.arange(lmax + 1, dtype=torch.float64, device=mask.device)
    alm = _scale_components(alm, torch.exp(-0.5 * ell * (ell + 1) * radius * radius)[None], lmax)
    return wrap(sht.alm2map_s0(alm, nside, lmax)[0] * mask)


def _max_pixrad(nside):
    """Largest angle between a RING pixel's centre and a point of the pixel: from the centre
    (z = 2/3, phi = pi / (4 nside)) of the first belt ring's first pixel to its corner at
    (z = 1 - (1 - 1/nside)^2 / 3, phi = 0)."""
    za, pa = 2.0 / 3.0, math.pi / (4 * nside)
    zb = 1.0 - (1.0 - 1.0 / nside) ** 2 / 3.0
    va = (math.sqrt(1 - za * za) * math.cos(pa), math.sqrt(1 - za * za) * math.sin(pa), za)
    vb = (math.sqrt(1 - zb * zb), 0.0, zb)
    cross = (va[1] * vb[2] - va[2] * vb[1], va[2] * vb[0] - va[0] * vb[2], va[0] * vb[1] - va[1] * vb[0])
    return math.atan2(math.sqrt(sum(c * c for c in cross)), sum(a * b for a, b in zip(va, vb)))


def _scale_components(alm, fl, lmax):
    """``alm[c] * fl[c][l]`` per component (``gm_alm_filter`` with a diagonal filter)."""
    n = alm.shape[0]
    cl = torch.zeros((n, n, lmax + 1), dtype=torch.float64, device=alm.device)
    for c in range(n):
        cl[c, c] = fl[c]
    out = torch.empty_like(alm)
    _lib.call("gm_alm_filter", _lib.ptr(alm), n, alm.shape[1], _lib.ptr(cl), n, lmax, _lib.ptr(out), out.shape[1],
              _lib.stream_ptr(alm.device))
    return out


def synfast_spherical(nside, cls, spin_arr, beam=None, seed=-1, wcs=None, lmax=None):
    """Correlated Gaussian full-sky fields (``gmaster/utils.py:823-927``).

    ``cls`` holds the upper triangle of the component covariance, row-major over components
    (one per spin-0 field, E and B per spin-s field).  The per-l covariance root comes from
    an eigendecomposition on the host (``(lmax+1) x ncomp^2`` numbers); the alms are drawn on
    the device (``gm_gaussian_alms``, Philox4x32-10 keyed by ``seed``), the beams applied per
    component and the maps made by this engine's ``alm2map``.  Statistically equivalent to
    the reference's (whose generator is JAX's), reproducible for a given seed.
    """
    if wcs is not None:
        raise NotImplementedError("CAR (WCS) maps are not supported")
    if int(nside) <= 0:
        raise ValueError("nside must be positive")
    spins = np.asarray(spin_arr, dtype=np.int32)
    if spins.ndim != 1 or np.any(spins < 0):
        raise ValueError("Spins must be a one-dimensional array of positive values")
    comps = np.asarray([1 if s == 0 else 2 for s in spins])
    ncomp = int(comps.sum())
    spectra = np.asarray(cls, dtype=np.float64)
    expected = ncomp * (ncomp + 1) // 2
    if spectra.ndim != 2 or len(spectra) != expected:
        raise ValueError(f"Must provide all Cls necessary to simulate all fields ({expected})")
    if ncomp > 8:
        raise NotImplementedError("at most 8 map components per draw")
    minfo = NmtMapInfo(None, (12 * int(nside) ** 2,))
    lmax = minfo.get_lmax() if lmax is None else int(lmax)
    lmax = min(lmax, spectra.shape[1] - 1)
    L = lmax + 1
    cov = np.zeros((L, ncomp, ncomp))
    k = 0
    for i in range(ncomp):
        for j in range(i, ncomp):
            cov[:, i, j] = cov[:, j, i] = spectra[k, :L]
            k += 1
    w, v = np.linalg.eigh(cov)
    scale = np.max(np.abs(w), axis=1, keepdims=True)
    if np.any(w < -1e-12 * np.maximum(scale, 1)):
        raise ValueError("Input power spectra do not form positive-semidefinite matrices")
    root = np.ascontiguousarray(v * np.sqrt(np.maximum(w, 0))[:, None, :])
    if seed < 0:
        seed = int(np.random.randint(50_000_000))
    dev = torch.device("cuda", torch.cuda.current_device())
    root_d = torch.as_tensor(root, device=dev)
    ai = NmtAlmInfo(lmax)
    alm = torch.empty((ncomp, ai.nelem), dtype=torch.complex128, device=dev)
    _lib.call("gm_gaussian_alms", _lib.ptr(root_d), lmax, ncomp, int(seed) & 0xFFFFFFFFFFFFFFFF, _lib.ptr(alm),
              _lib.stream_ptr(dev))
    if beam is not None:
        b = np.asarray(beam, dtype=np.float64)
        if b.ndim != 2 or len(b) != len(spins):
            raise ValueError("Must provide one beam per field")
        if b.shape[1] < L:
            raise ValueError(f"The beam should be provided to ell = {lmax}")
        per_comp = torch.as_tensor(np.repeat(b[:, :L], comps, axis=0), device=dev)
        alm = _scale_components(alm, per_comp, lmax)
    maps = []
    c = 0
    for s, n in zip(spins, comps):
        maps.append(alm2map(alm[c:c + n], int(s), minfo, ai).as_subclass(torch.Tensor))
        c += n
    return wrap(torch.cat(maps))
