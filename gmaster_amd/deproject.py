"""Contaminant deprojection: device primitives and the deprojection biases.

* ``map_dots`` — weighted inner products of maps (``gm_map_dots``): the template
  covariance and template-map products of ``gmaster/field.py:231-241`` and ``dot_map``
  (``gmaster/utils.py:256-263``, HEALPix weight 4 pi / npix).
* ``deproject`` — alphas = M^+ (T . maps) and maps -= sum alpha_i T_i (``gm_map_deproject``;
  ``gmaster/field.py:231-247``); the pseudo-inverse of the small covariance is taken on the
  host (``moore_penrose_pinvh``, ``gmaster/utils.py:527-553``).
* ``filter_alms`` — ``gm_alm_filter`` (``gmaster/workspaces.py:2258-2265``).
* ``deprojection_bias``, ``uncorr_noise_deprojection_bias``, ``compute_full_master`` —
  ``gmaster/workspaces.py:2268-2520``, composed from the engine's transforms (purified for
  pure fields), ``gm_coupled_cell`` and the primitives above.
"""

import numpy as np
import torch

from . import _lib, sht
from ._array import device_tensor, wrap
from ._config import APPROXIMATE_FORMS, nmt_params


def moore_penrose_pinvh(mat, tol_pinv):
    """Pseudo-inverse of a symmetric matrix: eigenvalues below ``tol_pinv`` times the largest
    get a zero inverse; ``tol_pinv`` None or <= 0 gives the plain inverse
    (``gmaster/utils.py:527-553``).  Host FP64 (the matrix is ntemp x ntemp)."""
    m = np.asarray(mat, dtype=np.float64)
    if tol_pinv is None or tol_pinv <= 0:
        return np.linalg.inv(m)
    ev, vec = np.linalg.eigh(m)
    keep = ev >= tol_pinv * np.max(ev)
    inv = np.zeros_like(ev)
    inv[keep] = 1.0 / ev[keep]
    return (vec * inv) @ vec.T


def map_dots(a, b, weight_map=None, w=None):
    """``out[i, j] = w sum_p a_i[p] b_j[p] v[p]`` over flattened maps: a ``(na, ...)``,
    b ``(nb, ...)`` with the same trailing size; v a per-pixel weight (last axis)."""
    a = device_tensor(a)
    b = device_tensor(b, device=a.device)
    na, nb = a.shape[0], b.shape[0]
    n = a[0].numel()
    if b[0].numel() != n:
        raise ValueError("map_dots: vectors of different sizes")
    npix = a.shape[-1]
    if w is None:
        w = 4.0 * np.pi / npix
    v = None
    if weight_map is not None:
        v = device_tensor(weight_map, device=a.device).reshape(-1)
        if v.numel() != npix:
            raise ValueError("map_dots: weight map does not match the pixelization")
    out = torch.empty((na, nb), dtype=torch.float64, device=a.device)
    work = torch.empty(int(_lib.query("gm_map_dots_work_size", na, nb, n)), dtype=torch.float64, device=a.device)
    _lib.call("gm_map_dots", _lib.ptr(a), na, n, _lib.ptr(b), nb, n, n, _lib.ptr(v), npix, float(w), _lib.ptr(out),
              _lib.ptr(work), _lib.stream_ptr(a.device))
    return out


def deproject(maps, templates, tol_pinv):
    """In place on ``maps`` (nmaps, npix): subtract the best-fit masked ``templates``
    (ntemp, nmaps, npix).  Returns ``(alphas, iM)`` as host arrays."""
    cov = map_dots(templates, templates).cpu().numpy()
    iM = moore_penrose_pinvh(cov, tol_pinv)
    prod = map_dots(templates, maps[None]).cpu().numpy()[:, 0]
    alphas = iM @ prod
    alpha_d = torch.as_tensor(alphas, dtype=torch.float64, device=maps.device)
    _lib.call("gm_map_deproject", _lib.ptr(maps), maps.numel(), _lib.ptr(templates), templates.shape[0],
              maps.numel(), _lib.ptr(alpha_d), _lib.stream_ptr(maps.device))
    return alphas, iM


def filter_alms(alms, spectra, lmax):
    """``out[o] = sum_k spectra[k, o][l] alms[k]`` for healpy-packed ``alms`` (nin, nelem)."""
    a = device_tensor(alms, dtype=torch.complex128)
    cl = device_tensor(spectra, device=a.device)
    nin, nout = cl.shape[0], cl.shape[1]
    if a.shape[0] != nin or cl.shape[2] != lmax + 1:
        raise ValueError("filter_alms: shapes do not match")
    out = torch.empty((nout, a.shape[1]), dtype=torch.complex128, device=a.device)
    _lib.call("gm_alm_filter", _lib.ptr(a), nin, a.shape[1], _lib.ptr(cl), nout, lmax, _lib.ptr(out), out.shape[1],
              _lib.stream_ptr(a.device))
    return out


def _precision():
    return "mixed" if nmt_params.form in APPROXIMATE_FORMS else "fp64"


def _map2alm(maps, spin, lmax, n_iter):
    return sht.map2alm(maps, spin, lmax, n_iter=n_iter, precision=_precision())


def _alm2map(alm, spin, nside, lmax):
    return sht.alm2map(alm, spin, nside, lmax, precision=_precision())


def _spectra(first, second, lmax):
    from .workspaces import _coupled_cell
    return _coupled_cell(first, second, lmax).reshape(len(first), len(second), lmax + 1)


def _iM(field):
    return torch.as_tensor(field.iM, dtype=torch.float64)


def _transformed(field, maps, n_iter):
    """The field's alms of unmasked ``maps``: masked and analysed, or purified
    (``transformed`` of ``gmaster/workspaces.py:2297-2310``)."""
    lmax = field.ainfo.lmax
    if field.pure_e or field.pure_b:
        from .purify import purify
        amask = field.get_mask_alms().as_subclass(torch.Tensor)
        return purify(field.mask, amask, maps, lmax, (field.pure_e, field.pure_b), n_iter, precision=_precision(),
                      return_maps=False)
    return _map2alm(maps * field.mask[None, :], field.spin, lmax, n_iter)


def deprojection_bias(f1, f2, cl_guess, n_iter=None):
    """Coupled bias of the pseudo-spectrum from template deprojection, ``(n1 n2, lmax+1)``
    (``gmaster/workspaces.py:2268-2386``)."""
    if not f1.is_compatible(f2):
        raise ValueError("Fields have incompatible pixelizations")
    lmax = f1.ainfo.lmax
    expected = (f1.nmaps * f2.nmaps, lmax + 1)
    guess = device_tensor(cl_guess, device=f1.mask.device)
    if tuple(guess.shape) != expected:
        raise ValueError(f"Guess Cl should have shape {expected}")
    if f1.lite or f2.lite:
        raise ValueError("Can't compute deprojection bias for lightweight fields")
    n_iter = f1.n_iter if n_iter is None else int(n_iter)
    nside = f1.minfo.nside
    spectra = guess.reshape(f1.nmaps, f2.nmaps, lmax + 1)
    bias = torch.zeros((f1.nmaps, f2.nmaps, lmax + 1), dtype=torch.float64)
    if f1.n_temp:
        ff = []
        for tj in f1.temp:
            alms = filter_alms(_map2alm(tj * f1.mask[None, :], f1.spin, lmax, n_iter), spectra, lmax)
            maps = _alm2map(alms, f2.spin, nside, lmax)
            filtered = _transformed(f2, maps, n_iter)
            ff.append(torch.stack([_spectra(ti, filtered, lmax) for ti in f1.alm_temp]).cpu())
        ff = torch.stack(ff, dim=1)
        bias -= torch.einsum("ij,ijklm->klm", _iM(f1), ff)
    if f2.n_temp:
        gg, products = [], []
        for tj in f2.temp:
            alms = filter_alms(_map2alm(tj * f2.mask[None, :], f2.spin, lmax, n_iter),
                               spectra.permute(1, 0, 2).contiguous(), lmax)
            maps = _alm2map(alms, f1.spin, nside, lmax)
            masked = maps * f1.mask[None, :]
            if f1.n_temp:
                products.append(map_dots(f1.temp, masked[None])[:, 0].cpu())
            filtered = _transformed(f1, maps, n_iter)
            gg.append(torch.stack([_spectra(filtered, ti, lmax) for ti in f2.alm_temp]).cpu())
        gg = torch.stack(gg, dim=1)
        bias -= torch.einsum("ij,ijklm->klm", _iM(f2), gg)
        if f1.n_temp:
            products = torch.stack(products, dim=1)
            fg = torch.stack([torch.stack([_spectra(a, b, lmax) for b in f2.alm_temp]) for a in f1.alm_temp]).cpu()
            bias += torch.einsum("ij,rs,jr,isklm->klm", _iM(f1), _iM(f2), products, fg)
    return wrap(bias.reshape(expected).to(f1.mask.device))


def uncorr_noise_deprojection_bias(f1, map_var, n_iter=None):
    """Deprojection bias of the auto-spectrum for uncorrelated noise of variance ``map_var``
    per steradian (``gmaster/workspaces.py:2389-2454``)."""
    if f1.lite:
        raise ValueError("Can't compute deprojection bias for lightweight fields")
    var = device_tensor(map_var, device=f1.mask.device).reshape(-1)
    if var.numel() != f1.minfo.npix:
        raise ValueError("Variance map doesn't match map resolution")
    lmax = f1.ainfo.lmax
    shape = (f1.nmaps * f1.nmaps, lmax + 1)
    if not f1.n_temp:
        return wrap(torch.zeros(shape, dtype=torch.float64, device=f1.mask.device))
    n_iter = f1.n_iter if n_iter is None else int(n_iter)
    wv = f1.mask * f1.mask * var
    first = []
    for tj in f1.temp:
        weighted = _map2alm(tj * wv[None, :], f1.spin, lmax, n_iter)
        first.append(torch.stack([_spectra(weighted, ti, lmax) for ti in f1.alm_temp]).cpu())
    first = torch.stack(first, dim=1)
    iM = _iM(f1)
    bias = -2 * torch.einsum("ij,ijklm->klm", iM, first)
    cells = torch.stack([torch.stack([_spectra(a, b, lmax) for b in f1.alm_temp]) for a in f1.alm_temp]).cpu()
    products = map_dots(f1.temp, f1.temp, weight_map=wv).cpu()
    bias += torch.einsum("ij,rs,jr,isklm->klm", iM, iM, products, cells)
    return wrap(bias.reshape(shape).to(f1.mask.device))


def compute_full_master(f1, f2, b=None, cl_noise=None, cl_guess=None, workspace=None, l_toeplitz=-1, l_exact=-1,
                        dl_band=-1, normalization="MASTER"):
    """Coupled spectrum minus deprojection and noise biases, decoupled into bandpowers
    (``gmaster/workspaces.py:2457-2520``)."""
    from .workspaces import NmtWorkspace, compute_coupled_cell
    if b is None and workspace is None:
        raise SyntaxError("Must supply either workspace or bins")
    if not f1.is_compatible(f2, strict=False):
        raise ValueError("Fields have incompatible pixelizations")
    shape = (f1.nmaps * f2.nmaps, f1.ainfo.lmax + 1)
    dev = f1.mask.device
    noise = torch.zeros(shape, dtype=torch.float64, device=dev) if cl_noise is None else device_tensor(cl_noise, device=dev)
    guess = torch.zeros(shape, dtype=torch.float64, device=dev) if cl_guess is None else device_tensor(cl_guess, device=dev)
    if tuple(noise.shape) != shape:
        raise ValueError(f"Noise Cl should have shape {shape}")
    if tuple(guess.shape) != shape:
        raise ValueError(f"Guess Cl should have shape {shape}")
    bias = deprojection_bias(f1, f2, guess).as_subclass(torch.Tensor)
    coupled = compute_coupled_cell(f1, f2).as_subclass(torch.Tensor)
    if workspace is None:
        workspace = NmtWorkspace.from_fields(f1, f2, b, l_toeplitz=l_toeplitz, l_exact=l_exact, dl_band=dl_band,
                                             normalization=normalization)
    return workspace.decouple_cell(coupled - bias - noise)
