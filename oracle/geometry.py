"""HEALPix RING geometry and healpy alm packing (oracle, FP64).

Restates ``gmaster/utils.py:220-248`` (``_SHTInfo``) and ``gmaster/utils.py:493-524``
(``NmtAlmInfo``).  Pinned by the HEALPix closed forms (Gorski et al. 2005) and by the
invariants checked in ``tests/test_oracle_geometry.py``: sum(nphi) == npix, ring
offsets contiguous, theta_{N-1-i} == pi - theta_i, and pixel centres that match an
independent per-pixel ``pix2ang`` restatement.
"""

import numpy as np


class Rings:
    """Per-ring geometry of a HEALPix map in RING order.

    Attributes (arrays of length ``nring = 4 nside - 1``, ring ``r = i - 1``):
        north   north-hemisphere ring index ``1..2 nside`` (the equator is ``2 nside``)
        z       cos(theta), computed from the closed form (exact 0 on the equator)
        sin     sin(theta), from ``(1 - z)(1 + z)`` with ``1 - z`` exact in the caps
        shalf   sin(theta / 2), chalf cos(theta / 2)
        theta   colatitude
        nphi    pixels in the ring (``4 i`` in the caps, ``4 nside`` in the belt)
        offset  index of the ring's first pixel
        phi0_num, phi0_den  phi0 = pi * phi0_num / phi0_den (exact rational)
        phi0    first-pixel azimuth
    """

    def __init__(self, nside):
        nside = int(nside)
        self.nside = nside
        self.npix = 12 * nside * nside
        self.nring = 4 * nside - 1
        i = np.arange(1, 4 * nside, dtype=np.int64)
        north = np.where(i > 2 * nside, 4 * nside - i, i)
        cap = north < nside
        south = i > 2 * nside
        self.north = north
        # 1 - z and 1 + z of the northern ring from exact rationals:
        #   caps  z = 1 - north^2 / (3 nside^2)
        #   belt  z = 2 (2 nside - north) / (3 nside), 1 - z = (2 north - nside) / (3 nside)
        nf = north.astype(np.float64)
        one_minus = np.where(cap, nf * nf / (3.0 * nside * nside), (2.0 * nf - nside) / (3.0 * nside))
        z_north = np.where(cap, 1.0 - nf * nf / (3.0 * nside * nside),
                           2.0 * (2 * nside - nf) / (3.0 * nside))
        one_plus = 1.0 + z_north
        sin = np.sqrt(one_minus * one_plus)
        shalf = np.sqrt(0.5 * one_minus)
        chalf = np.sqrt(0.5 * one_plus)
        self.z = np.where(south, -z_north, z_north)
        self.sin = sin
        self.shalf = np.where(south, chalf, shalf)
        self.chalf = np.where(south, shalf, chalf)
        self.theta = np.where(south, np.pi - np.arctan2(sin, z_north), np.arctan2(sin, z_north))
        self.nphi = np.where(cap, 4 * north, 4 * nside).astype(np.int64)
        self.offset = np.concatenate([[0], np.cumsum(self.nphi)[:-1]]).astype(np.int64)
        # phi0 = pi / (4 north) in the caps; pi / (4 nside) on belt rings with
        # (north - nside) even, else 0 (gmaster/utils.py:235,238).
        belt_shift = ((north - nside) % 2) == 0
        self.phi0_num = np.where(cap, 1, np.where(belt_shift, 1, 0)).astype(np.int64)
        self.phi0_den = np.where(cap, 4 * north, 4 * nside).astype(np.int64)
        self.phi0 = np.pi * self.phi0_num / self.phi0_den
        self.weight = 4.0 * np.pi / self.npix


def pix2ang_ring(nside, pix):
    """Independent per-pixel (theta, phi) of HEALPix RING pixels (Gorski et al. 2005).

    Used only to pin ``Rings``; it follows the published pixel formulae directly.
    """
    pix = np.asarray(pix, dtype=np.int64)
    npix = 12 * nside * nside
    ncap = 2 * nside * (nside - 1)
    theta = np.empty(pix.shape)
    phi = np.empty(pix.shape)
    northcap = pix < ncap
    southcap = pix >= npix - ncap
    belt = ~(northcap | southcap)
    # north cap
    p = pix[northcap]
    ph = (p + 1) / 2.0
    i = np.floor(np.sqrt(ph - np.sqrt(np.floor(ph)))).astype(np.int64) + 1
    j = p + 1 - 2 * i * (i - 1)
    theta[northcap] = np.arccos(1 - i * i / (3.0 * nside * nside))
    phi[northcap] = (j - 0.5) * np.pi / (2.0 * i)
    # belt
    p = pix[belt] - ncap
    i = p // (4 * nside) + nside
    j = p % (4 * nside) + 1
    # first pixel at phi = 0 when (i - nside) is odd, pi / (4 nside) otherwise
    s = (i - nside + 1) % 2
    theta[belt] = np.arccos((2 * nside - i) * 2.0 / (3.0 * nside))
    phi[belt] = (j - 1 + s / 2.0) * np.pi / (2.0 * nside)
    # south cap
    p = npix - pix[southcap]
    ph = p / 2.0
    i = np.floor(np.sqrt(ph - np.sqrt(np.floor(ph)))).astype(np.int64) + 1
    j = 4 * i + 1 - (p - 2 * i * (i - 1))
    theta[southcap] = np.arccos(-(1 - i * i / (3.0 * nside * nside)))
    phi[southcap] = (j - 0.5) * np.pi / (2.0 * i)
    return theta, phi


def ang2pix_ring(nside, theta, phi):
    """RING pixel containing (theta, phi) (Gorski et al. 2005, the cap / belt formulae)."""
    z = np.cos(np.asarray(theta, dtype=np.float64))
    za = np.abs(z)
    tt = np.mod(np.asarray(phi, dtype=np.float64), 2 * np.pi) / (0.5 * np.pi)  # in [0, 4)
    npix = 12 * nside * nside
    ncap = 2 * nside * (nside - 1)
    pix = np.empty(z.shape, dtype=np.int64)
    belt = za <= 2.0 / 3.0
    t1 = nside * (0.5 + tt[belt])
    t2 = nside * z[belt] * 0.75
    jp = np.floor(t1 - t2).astype(np.int64)  # index of the ascending edge line
    jm = np.floor(t1 + t2).astype(np.int64)  # index of the descending edge line
    ir = nside + 1 + jp - jm  # ring index counted from z = 2/3, in 1 .. 2 nside + 1
    kshift = 1 - (ir & 1)
    ip = np.mod((jp + jm - nside + kshift + 1) // 2, 4 * nside)
    pix[belt] = ncap + (ir - 1) * 4 * nside + ip
    cap = ~belt
    tc = tt[cap]
    tp = tc - np.floor(tc)
    tmp = nside * np.sqrt(3.0 * (1.0 - za[cap]))
    jp = np.floor(tp * tmp).astype(np.int64)
    jm = np.floor((1.0 - tp) * tmp).astype(np.int64)
    ir = jp + jm + 1  # ring index counted from the nearer pole
    ip = np.mod(np.floor(tc * ir).astype(np.int64), 4 * ir)
    pix[cap] = np.where(z[cap] > 0, 2 * ir * (ir - 1) + ip, npix - 2 * ir * (ir + 1) + ip)
    return pix


def max_pixrad(nside):
    """Largest angle between a RING pixel centre and a point of that pixel: from the centre
    (z = 2/3, phi = pi / (4 nside)) of the first belt ring's first pixel to its corner at
    (z = 1 - (1 - 1/nside)^2 / 3, phi = 0)."""
    za, pa = 2.0 / 3.0, np.pi / (4 * nside)
    zb, pb = 1.0 - (1.0 - 1.0 / nside) ** 2 / 3.0, 0.0
    va = np.array([np.sqrt(1 - za * za) * np.cos(pa), np.sqrt(1 - za * za) * np.sin(pa), za])
    vb = np.array([np.sqrt(1 - zb * zb) * np.cos(pb), np.sqrt(1 - zb * zb) * np.sin(pb), zb])
    return float(np.arctan2(np.linalg.norm(np.cross(va, vb)), va @ vb))


def pix2vec_ring(nside, pix):
    theta, phi = pix2ang_ring(nside, pix)
    st = np.sin(theta)
    return np.stack([st * np.cos(phi), st * np.sin(phi), np.cos(theta)], axis=-1)


class AlmInfo:
    """healpy m-major packing: ``idx(l, m) = m (2 lmax + 1 - m) / 2 + l`` (gmaster/utils.py:493-524)."""

    def __init__(self, lmax):
        self.lmax = int(lmax)
        m = np.arange(self.lmax + 1, dtype=np.int64)
        self.mstart = m * (2 * self.lmax + 1 - m) // 2
        self.nelem = (self.lmax + 1) * (self.lmax + 2) // 2
        self.ell = np.concatenate([np.arange(mm, self.lmax + 1) for mm in m])
        self.m = np.repeat(m, self.lmax + 1 - m)

    def idx(self, ell, m):
        return self.mstart[m] + ell
