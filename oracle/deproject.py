"""Contaminant deprojection (oracle, FP64, numpy) — test infrastructure only.

Restates the template deprojection of a field (``gmaster/field.py:193-247``): masked
templates T_i, covariance M_ij = w sum T_i T_j (w = 4 pi / npix, ``gmaster/utils.py:245,
256-263``), its pseudo-inverse with relative eigenvalue cut (``moore_penrose_pinvh``,
``gmaster/utils.py:527-553``), alphas = M^+ (w sum T_i maps), maps -= sum alpha_i T_i; the
alm filter ``_filter_alms`` (``gmaster/workspaces.py:2258-2265``); and the coupled
deprojection biases ``deprojection_bias`` (``:2268-2386``) and
``uncorr_noise_deprojection_bias`` (``:2389-2454``), non-purified fields.

Pinned by closed forms (tests/test_oracle_deproject.py): on the full sky a template that
is one real spherical harmonic removes exactly one mode, so the signal bias is
-C_l0 / (2 l0 + 1) at l0 and zero elsewhere, and the uncorrelated-noise bias is
-var / (2 l0 + 1) for a uniform variance per steradian; two non-orthogonal templates
spanning two modes remove both; the cleaned map is orthogonal to every template; the
pseudo-inverse agrees with ``numpy.linalg.pinv`` (hermitian, same cut).
"""

import numpy as np

from . import sht
from .geometry import AlmInfo
from .master import coupled_cell


def pinvh(mat, tol):
    """Pseudo-inverse of a symmetric matrix: eigenvalues below tol * max are dropped
    (``gmaster/utils.py:546-553``); tol None or <= 0 gives the plain inverse."""
    mat = np.asarray(mat, dtype=np.float64)
    if tol is None or tol <= 0:
        return np.linalg.inv(mat)
    ev, vec = np.linalg.eigh(mat)
    inv = np.where(ev >= tol * np.max(ev), 1.0 / np.where(ev == 0, 1.0, ev), 0.0)
    return (vec * inv) @ vec.T


def deproject(mask, maps, templates, tol=1e-10):
    """Masked maps and templates -> (cleaned maps, masked templates, alphas, iM)
    (``gmaster/field.py:220-247``, not pre-masked)."""
    mask = np.asarray(mask, dtype=np.float64)
    maps = np.atleast_2d(np.asarray(maps, dtype=np.float64)) * mask[None, :]
    temps = np.asarray(templates, dtype=np.float64) * mask[None, None, :]
    w = 4 * np.pi / mask.size
    cov = np.einsum("iap,jap->ij", temps, temps) * w
    iM = pinvh(cov, tol)
    prod = np.einsum("iap,ap->i", temps, maps) * w
    alphas = iM @ prod
    return maps - np.einsum("i,iap->ap", alphas, temps), temps, alphas, iM


def filter_alms(alms, spectra, lmax):
    """``out[o] = sum_k spectra[k, o][l] alms[k]`` (``gmaster/workspaces.py:2258-2265``)."""
    ell = AlmInfo(lmax).ell
    return np.array([sum(alms[k] * spectra[k, o][ell] for k in range(len(alms)))
                     for o in range(spectra.shape[1])])


class Field:
    """Field with templates, optionally E/B-purified (``task``): the pieces the bias
    formulas read (``gmaster/field.py:193-315``, not pre-masked)."""

    def __init__(self, mask, maps, templates, spin, nside, lmax, n_iter, tol=1e-10, task=(False, False)):
        from .purify import purify
        self.mask = np.asarray(mask, dtype=np.float64)
        self.spin, self.nside, self.lmax, self.n_iter = spin, nside, lmax, n_iter
        self.nmaps = 1 if spin == 0 else 2
        self.task = tuple(task)
        self.pure = any(self.task)
        unmasked = np.atleast_2d(np.asarray(maps, dtype=np.float64)).copy()
        if templates is None:
            self.maps = unmasked * self.mask[None, :]
            self.temp, self.iM, self.alphas = None, None, None
            self.n_temp = 0
        else:
            self.maps, self.temp, self.alphas, self.iM = deproject(self.mask, maps, templates, tol)
            self.n_temp = len(self.temp)
            unmasked = unmasked - np.einsum("i,iap->ap", self.alphas, np.asarray(templates, dtype=np.float64))
        if self.pure:
            self.alm_mask = sht.map2alm(self.mask[None], 0, nside, lmax, n_iter)[0]
            self.alm = purify(self.mask, self.alm_mask, unmasked, nside, lmax, self.task, n_iter)
            self.maps = sht.alm2map(self.alm, 2, nside, lmax)
            if self.n_temp:
                self.alm_temp = np.array([purify(self.mask, self.alm_mask, t, nside, lmax, self.task, n_iter)
                                          for t in np.asarray(templates, dtype=np.float64)])
        else:
            self.alm = sht.map2alm(self.maps, spin, nside, lmax, n_iter)
            if self.n_temp:
                self.alm_temp = np.array([sht.map2alm(t, spin, nside, lmax, n_iter) for t in self.temp])

    def transformed(self, maps):
        """Alms of unmasked maps as the field forms its own (masked, or purified)."""
        from .purify import purify
        if self.pure:
            return purify(self.mask, self.alm_mask, maps, self.nside, self.lmax, self.task, self.n_iter)
        return sht.map2alm(maps * self.mask[None, :], self.spin, self.nside, self.lmax, self.n_iter)


def _spectra(first, second, lmax):
    return coupled_cell(first, second, lmax).reshape(len(first), len(second), lmax + 1)


def deprojection_bias(f1, f2, cl_guess):
    """Coupled deprojection bias, rows a * n2 + b (``gmaster/workspaces.py:2312-2386``)."""
    lmax, nside, n_iter = f1.lmax, f1.nside, f1.n_iter
    spectra = np.asarray(cl_guess, dtype=np.float64).reshape(f1.nmaps, f2.nmaps, lmax + 1)
    bias = np.zeros((f1.nmaps, f2.nmaps, lmax + 1))
    w = 4 * np.pi / f1.mask.size
    if f1.n_temp:
        ff = []
        for tj in f1.temp:
            alms = sht.map2alm(tj * f1.mask[None, :], f1.spin, nside, lmax, n_iter)
            alms = filter_alms(alms, spectra, lmax)
            maps = sht.alm2map(alms, f2.spin, nside, lmax)
            filtered = f2.transformed(maps)
            ff.append(np.array([_spectra(ti, filtered, lmax) for ti in f1.alm_temp]))
        ff = np.stack(ff, axis=1)
        bias -= np.einsum("ij,ijklm->klm", f1.iM, ff)
    if f2.n_temp:
        gg, products = [], []
        for tj in f2.temp:
            alms = sht.map2alm(tj * f2.mask[None, :], f2.spin, nside, lmax, n_iter)
            alms = filter_alms(alms, spectra.transpose(1, 0, 2), lmax)
            maps = sht.alm2map(alms, f1.spin, nside, lmax)
            if f1.n_temp:
                products.append(np.array([w * np.sum(ti * maps * f1.mask[None, :]) for ti in f1.temp]))
            filtered = f1.transformed(maps)
            gg.append(np.array([_spectra(filtered, ti, lmax) for ti in f2.alm_temp]))
        gg = np.stack(gg, axis=1)
        bias -= np.einsum("ij,ijklm->klm", f2.iM, gg)
        if f1.n_temp:
            products = np.stack(products, axis=1)
            fg = np.array([[_spectra(a, b, lmax) for b in f2.alm_temp] for a in f1.alm_temp])
            bias += np.einsum("ij,rs,jr,isklm->klm", f1.iM, f2.iM, products, fg)
    return bias.reshape(f1.nmaps * f2.nmaps, lmax + 1)


def uncorr_noise_deprojection_bias(f1, map_var):
    """Uncorrelated-noise deprojection bias (``gmaster/workspaces.py:2414-2454``)."""
    lmax = f1.lmax
    shape = (f1.nmaps * f1.nmaps, lmax + 1)
    if not f1.n_temp:
        return np.zeros(shape)
    w = 4 * np.pi / f1.mask.size
    wv = f1.mask ** 2 * np.asarray(map_var, dtype=np.float64)
    first = []
    for tj in f1.temp:
        weighted = sht.map2alm(tj * wv[None, :], f1.spin, f1.nside, lmax, f1.n_iter)
        first.append(np.array([_spectra(weighted, ti, lmax) for ti in f1.alm_temp]))
    first = np.stack(first, axis=1)
    bias = -2 * np.einsum("ij,ijklm->klm", f1.iM, first)
    cells = np.array([[_spectra(a, b, lmax) for b in f1.alm_temp] for a in f1.alm_temp])
    products = np.array([[w * np.sum(a * b * wv[None, :]) for b in f1.temp] for a in f1.temp])
    bias += np.einsum("ij,rs,jr,isklm->klm", f1.iM, f1.iM, products, cells)
    return bias.reshape(shape)
