"""Synthetic MASTER inputs: apodized hole masks and correlated Gaussian T, Q, U maps.

The workload of the Nside-4096 T+P benchmark (and its small-Nside parity cases):

* mask: 1 except ``nholes`` discs of radius ``radius_deg`` with centres uniform on the
  sphere (NumPy PCG64, ``seed``), apodized C1 (``x - sin(2 pi x)/(2 pi)``) or C2
  (``(1 - cos pi x)/2``) with ``x = sqrt((1 - cos d)/(1 - cos apo))``, where ``d`` is the
  angle from the pixel centre to the nearest disc edge.  (The reference apodizer measures
  ``d`` to the nearest masked pixel centre, ``gmaster/utils.py:61-147``; on a hole mask the
  two differ by less than a pixel, and this form needs no nearest-neighbour search over
  2e8 pixels.)  Optionally a galactic-latitude cut ``|b| > bcut_deg`` in the map's own frame.
* maps: correlated Gaussian T, E, B with ``C_TT = 400 / (l + 10)^2`` (unit at l = 10),
  ``C_EE = 0.1 C_TT``, ``C_TE = 0.3 sqrt(C_TT C_EE)``, ``C_BB = 0.01 C_EE``; alms drawn with
  NumPy (seeded), maps made by this engine's FP64 ``alm2map`` so every form sees identical
  inputs.
"""

import math

import numpy as np
import torch

from . import sht


def pixel_vectors(nside, device):
    """``(npix, 3)`` unit vectors of the RING pixels, built from the ring closed forms."""
    nring = 4 * nside - 1
    i = torch.arange(1, 4 * nside, device=device, dtype=torch.int64)
    north = torch.where(i > 2 * nside, 4 * nside - i, i)
    cap = north < nside
    nphi = torch.where(cap, 4 * north, torch.full_like(north, 4 * nside))
    q = north.double()
    zn = torch.where(cap, 1.0 - q * q / (3.0 * nside * nside), 2.0 * (2 * nside - q) / (3.0 * nside))
    z = torch.where(i > 2 * nside, -zn, zn)
    phi0 = torch.where(cap, math.pi / (4.0 * q),
                       torch.where(((north - nside) % 2) == 0, torch.full_like(q, math.pi / (4.0 * nside)),
                                   torch.zeros_like(q)))
    offs = torch.cumsum(nphi, 0) - nphi
    npix = 12 * nside * nside
    pix = torch.arange(npix, device=device, dtype=torch.int64)
    r = torch.searchsorted(offs, pix, right=True) - 1
    k = (pix - offs[r]).double()
    phi = phi0[r] + 2.0 * math.pi * k / nphi[r].double()
    zz = z[r]
    st = torch.sqrt(torch.clamp((1.0 - zz) * (1.0 + zz), min=0.0))
    del pix, r, k
    return torch.stack([st * torch.cos(phi), st * torch.sin(phi), zz], dim=1), nring


def hole_mask(nside, nholes=20, radius_deg=5.0, apo_deg=2.0, kind="C1", seed=1234, bcut_deg=None, device="cuda"):
    """Apodized mask with ``nholes`` random disc holes (and an optional |b| cut).
    ``radius_deg`` is one radius or one per hole."""
    rng = np.random.default_rng(seed)
    zc = rng.uniform(-1.0, 1.0, nholes)
    pc = rng.uniform(0.0, 2 * np.pi, nholes)
    sc = np.sqrt(1 - zc * zc)
    centres = torch.as_tensor(np.stack([sc * np.cos(pc), sc * np.sin(pc), zc], 1), device=device)
    vec, _ = pixel_vectors(nside, device)
    # one radius for all holes, or one per hole
    R = torch.as_tensor(np.radians(np.broadcast_to(np.asarray(radius_deg, dtype=np.float64), (nholes,))),
                        device=device)
    # angle to the nearest disc edge (negative inside a hole), processed in slabs
    npix = vec.shape[0]
    edge = torch.empty(npix, dtype=torch.float64, device=device)
    slab = max(1, (1 << 27) // max(nholes, 1))
    for s in range(0, npix, slab):
        v = vec[s:s + slab]
        cosang = torch.clamp(v @ centres.T, -1.0, 1.0)
        edge[s:s + slab] = (torch.acos(cosang) - R[None, :]).min(dim=1).values
    if bcut_deg is not None:
        lat = torch.asin(torch.clamp(vec[:, 2], -1.0, 1.0))
        edge = torch.minimum(edge, torch.abs(lat) - math.radians(bcut_deg))
    del vec
    mask = torch.ones(npix, dtype=torch.float64, device=device)
    mask[edge <= 0] = 0.0
    apo = math.radians(apo_deg)
    if apo > 0:
        thr = 1.0 - math.cos(apo)
        aff = (edge > 0) & (edge < apo)
        x = torch.sqrt((1.0 - torch.cos(edge[aff])) / thr)
        if kind == "C1":
            f = x - torch.sin(2 * math.pi * x) / (2 * math.pi)
        else:
            f = 0.5 * (1.0 - torch.cos(math.pi * x))
        mask[aff] = f
    return mask


def theory_cls(lmax):
    """``(TT, EE, BB, TE)`` of the benchmark, each ``lmax + 1`` long."""
    ell = np.arange(lmax + 1, dtype=np.float64)
    tt = 400.0 / (ell + 10.0) ** 2
    ee = 0.1 * tt
    te = 0.3 * np.sqrt(tt * ee)
    bb = 0.01 * ee
    return tt, ee, bb, te


def gaussian_alms(lmax, seed=4321):
    """``(3, nelem)`` complex alms (T, E, B) with the benchmark spectra (healpy packing)."""
    tt, ee, bb, te = theory_cls(lmax)
    L = lmax + 1
    m = np.repeat(np.arange(L), L - np.arange(L))
    ell = np.concatenate([np.arange(mm, L) for mm in range(L)])
    rng = np.random.default_rng(seed)
    n = len(ell)
    g = rng.standard_normal((4, n))  # re/im of two correlated + B handled below
    scale = np.where(m == 0, 1.0, 1.0 / np.sqrt(2.0))
    im_on = (m > 0).astype(np.float64)
    # Cholesky of [[TT, TE], [TE, EE]] per l
    a11 = np.sqrt(tt)
    a21 = np.where(tt > 0, te / np.where(tt > 0, a11, 1), 0.0)
    a22 = np.sqrt(np.maximum(ee - a21 * a21, 0.0))
    zt = (g[0] + 1j * g[1] * im_on) * scale
    ze = (g[2] + 1j * g[3] * im_on) * scale
    del g
    aT = a11[ell] * zt
    aE = a21[ell] * zt + a22[ell] * ze
    del zt, ze
    gb = rng.standard_normal((2, n))
    aB = np.sqrt(bb)[ell] * (gb[0] + 1j * gb[1] * im_on) * scale
    out = np.stack([aT, aE, aB])
    out[:, ell < 2] = np.where(np.arange(3)[:, None] == 0, out[:, ell < 2], 0)  # no E/B below l = 2
    return out


def tqu_maps(nside, lmax=None, seed=4321, device="cuda"):
    """``(3, npix)`` T, Q, U maps of the benchmark spectra, synthesised in FP64 on the GPU."""
    lmax = 3 * nside - 1 if lmax is None else int(lmax)
    alm = torch.as_tensor(gaussian_alms(lmax, seed), device=device)
    t = sht.alm2map_s0(alm[0:1], nside, lmax)
    qu = sht.alm2map_spin(alm[1:3], nside, lmax, 2)
    return torch.cat([t, qu])
