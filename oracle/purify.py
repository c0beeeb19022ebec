"""E/B purification of a spin-2 field (oracle, FP64, numpy) -- test infrastructure only.

Restates ``gmaster/field.py:433-499`` (Smith 2006): the masked alms plus counter-terms
from the spin-1 and spin-2 derivatives of the mask,

    w1 = alm2map(-sqrt(l(l+1)) a_mask, spin 1),  w2 = alm2map(sqrt(l(l+1)(l+2)(l-1)) a_mask, spin 2),
    a^X += 2/sqrt((l+2)(l-1)) map2alm_1(conj(w1) P)^X + 1/sqrt((l+2)(l+1)l(l-1)) map2alm_0(conj(w2) P)^X

for the purified components X (P = Q + iU the unmasked map).

Pinned (tests/test_oracle_purify.py) by what purification is for: on a pure-E sky seen
through a smooth (apodized) mask the purified B modes carry a small fraction of the E->B
leakage of the ordinary masked B modes; on the full sky (mask 1, no derivatives) the
counter-terms vanish; and the map products are conj(w) P.
"""

import numpy as np

from . import sht
from .geometry import AlmInfo


def _ell_scale(alm, f, lmax):
    return alm * f[AlmInfo(lmax).ell]


def spin_mul(w, p):
    """conj(w) p for (2, npix) two-component maps."""
    return np.stack([w[0] * p[0] + w[1] * p[1], w[0] * p[1] - w[1] * p[0]])


def purify(mask, alm_mask, maps_unmasked, nside, lmax, task, n_iter):
    ls = np.arange(lmax + 1, dtype=np.float64)
    alms = sht.map2alm(maps_unmasked * mask[None, :], 2, nside, lmax, n_iter)
    walm = np.zeros((2, len(alm_mask)), dtype=np.complex128)
    walm[0] = _ell_scale(alm_mask, -np.sqrt((ls + 1) * ls), lmax)
    wmap = sht.alm2map(walm, 1, nside, lmax)
    palm = sht.map2alm(spin_mul(wmap, maps_unmasked), 1, nside, lmax, n_iter)
    f1 = np.where(ls >= 2, 2.0 / np.sqrt(np.maximum((ls + 2) * (ls - 1), 1.0)), 0.0)
    for ipol, on in enumerate(task):
        if on:
            alms[ipol] = alms[ipol] + _ell_scale(palm[ipol], f1, lmax)
    f2 = np.where(ls >= 2, -np.sqrt(np.maximum((ls + 2) * (ls - 1), 0.0)), 0.0)
    walm[0] = _ell_scale(walm[0], f2, lmax)
    wmap = sht.alm2map(walm, 2, nside, lmax)
    prod = spin_mul(wmap, maps_unmasked)
    palm = np.stack([sht.map2alm(c[None], 0, nside, lmax, n_iter)[0] for c in prod])
    f3 = np.where(ls >= 2, 1.0 / np.sqrt(np.maximum((ls + 2) * (ls + 1) * ls * (ls - 1), 1.0)), 0.0)
    for ipol, on in enumerate(task):
        if on:
            alms[ipol] = alms[ipol] + _ell_scale(palm[ipol], f3, lmax)
    return alms
