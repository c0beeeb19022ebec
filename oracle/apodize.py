"""C1 / C2 mask apodization (oracle).

Restates ``gmaster/utils.py:61-147``: for every unmasked pixel with chord-squared
distance ``d = 0.5 |v - v_hole|^2 = 1 - cos(theta)`` to the nearest masked pixel below
``1 - cos(aposize)``, ``x = sqrt(d / (1 - cos aposize))`` and the mask is multiplied by
``x - sin(2 pi x)/(2 pi)`` (C1) or ``(1 - cos(pi x))/2`` (C2).  Pixel vectors come from
the oracle's own ``pix2vec_ring``; nearest neighbours from ``scipy.spatial.cKDTree``.
Pinned by the definition (brute-force nearest neighbour at Nside 8 and 16 in the tests).
``smooth_apodization`` restates "Smooth"; its zeroed set uses ``geometry.max_pixrad`` (pinned
as the supremum of centre distances over random points with ``geometry.ang2pix_ring``);
healpy's own ``query_disc`` cannot run here, so its agreement with healpy is parity unpinned.
"""

import numpy as np
from scipy.spatial import cKDTree

from .geometry import pix2vec_ring


def mask_apodization(mask, aposize_deg, apotype="C1"):
    mask = np.asarray(mask, dtype=np.float64)
    npix = mask.size
    nside = int(round(np.sqrt(npix / 12)))
    radius = np.radians(aposize_deg)
    masked = np.flatnonzero(mask <= 0)
    if aposize_deg == 0 or masked.size == 0:
        return mask.copy()
    vec = pix2vec_ring(nside, np.arange(npix))
    dist, _ = cKDTree(vec[masked]).query(vec)
    dm = 0.5 * dist**2
    thr = 1 - np.cos(radius)
    aff = (mask > 0) & (dm < thr)
    x = np.sqrt(np.maximum(dm[aff], 0) / thr)
    if apotype == "C1":
        fac = x - np.sin(2 * np.pi * x) / (2 * np.pi)
    else:
        fac = 0.5 * (1 - np.cos(np.pi * x))
    out = mask.copy()
    out[aff] *= fac
    return out


def smooth_apodization(mask, aposize_deg, n_iter=3):
    """"Smooth" apodization (``gmaster/utils.py:119-134``): every pixel whose centre lies
    closer than 2.5 aposize + the largest pixel radius to a masked pixel's centre (healpy's
    inclusive ``query_disc`` with fact 1 around each masked pixel) is zeroed, the result is
    analysed (n_iter 3, lmax 3 nside - 1), smoothed by exp(-l(l+1) aposize^2 / 2),
    synthesised and multiplied by the mask."""
    from . import sht
    from .geometry import AlmInfo, max_pixrad

    mask = np.asarray(mask, dtype=np.float64)
    npix = mask.size
    nside = int(round(np.sqrt(npix / 12)))
    radius = np.radians(aposize_deg)
    masked = np.flatnonzero(mask <= 0)
    if aposize_deg == 0 or masked.size == 0:
        return mask.copy()
    vec = pix2vec_ring(nside, np.arange(npix))
    dist, _ = cKDTree(vec[masked]).query(vec)
    rbig = min(2.5 * radius + max_pixrad(nside), np.pi)
    binary = mask.copy()
    binary[0.5 * dist**2 < 1 - np.cos(rbig)] = 0.0
    lmax = 3 * nside - 1
    alm = sht.map2alm(binary[None], 0, nside, lmax, n_iter)[0]
    ell = AlmInfo(lmax).ell
    alm = alm * np.exp(-0.5 * ell * (ell + 1) * radius**2)
    return sht.alm2map(alm, 0, nside, lmax)[0] * mask
