"""E/B purification of spin-2 fields (Smith 2006; ``gmaster/field.py:433-499``).

The ordinary masked alms receive counter-terms built from the spin-1 and spin-2
derivatives of the mask: with ``w_1 = alm2map(-sqrt(l (l+1)) a^mask, spin 1)`` and
``w_2 = alm2map(sqrt(l (l+1) (l+2) (l-1)) a^mask, spin 2)`` (E parts, B = 0),

    a^X += 2 / sqrt((l+2)(l-1))            map2alm_spin1(conj(w_1) P)^X
         + 1 / sqrt((l+2)(l+1) l (l-1))    map2alm_spin0(conj(w_2) P)^X      (l >= 2)

for the purified components X (E and/or B), where P = Q + iU is the unmasked map.  The
products run in ``gm_spin_mul``, the l-dependent factors in ``gm_alm_filter``, the
transforms in the engine's spin-0/1/2 kernels.
"""

import torch

from . import _lib, sht


def _ell_factor(alm, fl, lmax):
    """``alm[k] * fl[l]`` for every component (``gm_alm_filter`` with a diagonal filter)."""
    a = alm if alm.ndim == 2 else alm[None]
    out = torch.empty_like(a)
    cl = torch.zeros((a.shape[0], a.shape[0], lmax + 1), dtype=torch.float64, device=a.device)
    for k in range(a.shape[0]):
        cl[k, k] = fl
    _lib.call("gm_alm_filter", _lib.ptr(a), a.shape[0], a.shape[1], _lib.ptr(cl), a.shape[0], lmax, _lib.ptr(out),
              out.shape[1], _lib.stream_ptr(a.device))
    return out if alm.ndim == 2 else out[0]


def _spin_mul(w, p):
    out = torch.empty_like(p)
    _lib.call("gm_spin_mul", _lib.ptr(w), _lib.ptr(p), p.shape[1], _lib.ptr(out), _lib.stream_ptr(p.device))
    return out


def purify(mask, alm_mask, maps_unmasked, lmax, task, n_iter, precision="fp64", return_maps=True):
    """Purified E/B alms ``(2, nelem)`` of the unmasked (2, npix) map (and their maps).
    ``task = (purify_e, purify_b)``."""
    dev = maps_unmasked.device
    nside = int(round((maps_unmasked.shape[1] / 12) ** 0.5))
    ls = torch.arange(lmax + 1, dtype=torch.float64, device=dev)
    alms = sht.map2alm(maps_unmasked * mask[None, :], 2, lmax, n_iter=n_iter, precision=precision)
    walm = torch.zeros((2, alm_mask.shape[-1]), dtype=torch.complex128, device=dev)
    walm[0] = _ell_factor(alm_mask.reshape(1, -1), -torch.sqrt((ls + 1) * ls), lmax)[0]
    # spin-1 derivative of the mask
    wmap = sht.alm2map(walm, 1, nside, lmax, precision=precision)
    palm = sht.map2alm(_spin_mul(wmap, maps_unmasked), 1, lmax, n_iter=n_iter, precision=precision)
    f1 = torch.where(ls >= 2, 2.0 / torch.sqrt(torch.clamp((ls + 2) * (ls - 1), min=1.0)), torch.zeros_like(ls))
    corr = _ell_factor(palm, f1, lmax)
    for ipol, on in enumerate(task):
        if on:
            alms[ipol] += corr[ipol]
    # spin-2 derivative of the mask
    f2 = torch.where(ls >= 2, -torch.sqrt(torch.clamp((ls + 2) * (ls - 1), min=0.0)), torch.zeros_like(ls))
    walm[0] = _ell_factor(walm[0:1], f2, lmax)[0]
    wmap = sht.alm2map(walm, 2, nside, lmax, precision=precision)
    prod = _spin_mul(wmap, maps_unmasked)
    palm = sht.map2alm_s0(prod, lmax, n_iter=n_iter, precision=precision)
    f3 = torch.where(ls >= 2, 1.0 / torch.sqrt(torch.clamp((ls + 2) * (ls + 1) * ls * (ls - 1), min=1.0)),
                     torch.zeros_like(ls))
    corr = _ell_factor(palm, f3, lmax)
    for ipol, on in enumerate(task):
        if on:
            alms[ipol] += corr[ipol]
    if return_maps:
        return alms, sht.alm2map(alms, 2, nside, lmax, precision=precision)
    return alms
