"""Batched transforms: spin-0 maps and spin-2 fields of one geometry through ONE spin-0
recurrence per (m, ring), with the contraction on the matrix unit (``gm_bat_*``).

A T+P realisation (T, its mask, Q/U) is 12 real channels: T and the mask contribute the
ring-spectrum sums ``S0_l = sum_r lambda_l G_r``; Q and U contribute ``S0_l`` and
``S1_l = sum_r lambda_l G_r / sin^2(theta_r)``, from which the finish kernel forms E and B
with the HEALPix F1/F2 identities (the spin-2 rows as combinations of the spin-0 rows
lambda_{l-2} .. lambda_{l+1}).  Rings farther than ~14.5 degrees from the poles contract
in FP32 on the matrix unit over the FP64 recurrence; the polar ring pairs, where 1/sin^2
grows and the spin-2 combinations cancel, contract in FP64.

The Jacobi refinement (``gmaster/_sht/healpix.py:1459-1471``) runs on corrections as in
``sht._refine``: after ``R_0 = fold(S alm_0) - G`` each pass analyses the residual and
the next residual is updated by the synthesis of the correction, so the rounding of every
pass after the first scales with the correction, not with the signal.
"""

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, sht

_PLANS = {}

# Ring pairs with sin^2(theta) below this contract in FP64 (the 1/sin^2 channel's weight
# stays <= 16 in FP32).
POLAR_SIN2 = 1.0 / 16.0
# The FP16 hi/lo analysis (every pass of a refinement: the next pass removes a pass's
# rounding, and the last pass's rounding scales with its residual) keeps FP64 only nearer
# the poles: its split is the largest multiple of its workgroup's ring pairs (ANA_BLOCK)
# between the sin^2(theta) = 1/24 and 1/96 splits, or the 1/48 split where no multiple
# falls in that range.  Whole workgroups matter: a last workgroup with a few live waves
# costs about as much as a full one for every order (Nside 4096, 8192 ring pairs: split
# 7680 = 15 workgroups; analysis pass 143.7 ms against 156.8 ms at the 1/16 split 6928,
# 175.8 ms at 7749; bandpowers unchanged, alm errors vs FP64 within 1e-6).
RES_SIN2_RANGE = (1.0 / 24.0, 1.0 / 96.0)
RES_SIN2_FALLBACK = 1.0 / 48.0
ANA_BLOCK = 512  # ring pairs per FP16/FP32 analysis workgroup (8 waves of 64)
# With refinement passes (n_iter >= 1) the first analysis contracts on FP16 hi/lo products
# too: the refinement subtracts A S (alm_0 - a) from the first alms, which removes their
# rounding with the rest of the quadrature error (measured at Nside 4096: alm errors vs
# FP64 1.2e-7 (T, E), 3.8e-7 (B) either way).  With n_iter = 0 the first pass is the
# result and contracts in FP32.
FIRST_PASS_FP16 = True

SUPPORTED = ((2, 1), (1, 1), (0, 1), (0, 2), (2, 0), (4, 0), (8, 0))

# zero entries after the alpha table and spare degrees after the synthesis coefficients
# (the kernels read whole 32-degree blocks past an order's last degree)
ALPHA_TAIL = 128
COEF_TAIL = 96


@dataclass
class BatPlan:
    nside: int
    lmax: int
    device: torch.device
    tab: sht.Plan          # spin-0 tables of lmax + 2 (alpha, P valid up to degree lmax + 2)
    alpha: torch.Tensor    # tab.alpha followed by ALPHA_TAIL zeros (whole-block reads stay in bounds)
    nact: torch.Tensor     # ring pairs visited by order m (prefix from the equator)
    jsplit: int            # first ring pair contracted in FP64
    mpol: int              # orders m < mpol reach beyond jsplit
    nb32: int
    nb64: int
    tsize: int
    tsize64: int           # table entries of the orders m < mpol
    res: tuple             # (jsplit, mpol, nb32, nb64, tsize64) of the FP16 hi/lo analysis


def _nactive_pairs(nside, lmax, polar_skip):
    """Ring pairs (equator first) that order m visits: the larger of ducc0's spin-0 and
    spin-2 polar-skip bounds (``sht._nactive``), evaluated on the folded rings."""
    z, sin = sht._ring_slots(nside, 0)
    n = len(z)
    L = lmax + 1
    if not polar_skip:
        return np.full(L, n, dtype=np.int32)
    ofs = max(100.0, 0.01 * lmax)
    b = lmax * sin + ofs
    mlim0 = b
    mlim2 = 2.0 * np.abs(z) + np.sqrt(np.maximum(b * b - 4.0 * sin * sin, 0.0))
    best = np.maximum.accumulate(np.maximum(mlim0, mlim2)[::-1])[::-1]
    return np.searchsorted(-best, -np.arange(L), side="right").astype(np.int32)


def _split_at(nside, sin2):
    _, sin = sht._ring_slots(nside, 0)
    return int(np.count_nonzero(sin * sin >= sin2))


def polar_split(nside):
    """Number of ring pairs (from the equator) with sin^2(theta) >= POLAR_SIN2."""
    return _split_at(nside, POLAR_SIN2)


def residual_split(nside):
    """First ring pair of the FP16 hi/lo analysis contracted in FP64 (RES_SIN2_RANGE)."""
    lo, hi = _split_at(nside, RES_SIN2_RANGE[0]), _split_at(nside, RES_SIN2_RANGE[1])
    j = hi // ANA_BLOCK * ANA_BLOCK
    return j if j > 0 and j >= lo else _split_at(nside, RES_SIN2_FALLBACK)


def _split_fields(nside, nact, Lt, jsplit):
    mpol = int(np.count_nonzero(nact > jsplit))
    return (jsplit, mpol, int(_lib.query("gm_bat_blocks", nside, jsplit, 0)),
            int(_lib.query("gm_bat_blocks", nside, jsplit, 1)), mpol * (Lt + 8) - mpol * (mpol - 1) // 2)


def plan(nside, lmax, device=None, polar_skip=True):
    device = torch.device(device if device is not None else "cuda")
    if device.index is None:
        device = torch.device("cuda", torch.cuda.current_device())
    key = (int(nside), int(lmax), device, bool(polar_skip))
    bp = _PLANS.get(key)
    if bp is not None:
        return bp
    nside, lmax = int(nside), int(lmax)
    tab = sht.plan(nside, lmax + 2, 0, device, polar_skip)
    if nside >= ANA_BLOCK and int(_lib.query("gm_bat_blocks", nside, ANA_BLOCK + 1, 0)) != 2:
        raise RuntimeError("ANA_BLOCK does not match the library's analysis workgroup")
    nact = _nactive_pairs(nside, lmax, polar_skip)
    Lt = lmax + 3  # table degrees
    jsplit, mpol, nb32, nb64, tsize64 = _split_fields(nside, nact, Lt, polar_split(nside))
    alpha = torch.zeros(tab.alpha.numel() + ALPHA_TAIL, dtype=torch.float64, device=device)
    alpha[: tab.alpha.numel()].copy_(tab.alpha)
    bp = BatPlan(
        nside=nside, lmax=lmax, device=device, tab=tab, alpha=alpha,
        nact=torch.as_tensor(nact, dtype=torch.int32, device=device),
        jsplit=jsplit, mpol=mpol, tsize64=tsize64, nb32=nb32, nb64=nb64,
        tsize=int(_lib.query("gm_table_size", lmax + 2)),
        res=_split_fields(nside, nact, Lt, residual_split(nside)),
    )
    _PLANS[key] = bp
    return bp


def clear_plans():
    _PLANS.clear()


def _check(k0, k2):
    if (k0, k2) not in SUPPORTED:
        raise ValueError(f"unsupported batch: {k0} spin-0 maps and {k2} spin-2 fields")


def analysis(bp, spec0, spec2, alm0, alm2, mode, residual=False):
    """Spectra -> alms (mode 0 writes, 1 subtracts, 2 adds).  ``spec0`` ``(K0, nring, L)``,
    ``spec2`` ``(K2, nring, 2L-1)`` (F+ of Q + iU), ``alm0`` ``(K0, nelem)``, ``alm2``
    ``(K2, 2, nelem)``; absent kinds are None.  ``residual=True`` (refinement passes, whose
    input is a residual map) contracts with FP16 hi/lo operands (products to ~2^-21)."""
    k0 = 0 if spec0 is None else spec0.shape[0]
    k2 = 0 if spec2 is None else spec2.shape[0]
    _check(k0, k2)
    st = _lib.stream_ptr(bp.device)
    S = int(_lib.query("gm_bat_stride", k0, k2, 0))
    jsplit, mpol, nb32, nb64, tsize64 = bp.res if residual else (bp.jsplit, bp.mpol, bp.nb32, bp.nb64, bp.tsize64)
    part32 = torch.empty(nb32 * bp.tsize * S, dtype=torch.float32, device=bp.device) if nb32 else None
    part64 = torch.empty(nb64 * tsize64 * S, dtype=torch.float64, device=bp.device) if nb64 and mpol else None
    t = bp.tab
    _lib.call("gm_bat_ana_res" if residual else "gm_bat_ana", _lib.ptr(spec0), k0, spec0.shape[2] if k0 else 0, spec0.stride(0) if k0 else 0,
              _lib.ptr(spec2), k2, spec2.shape[2] if k2 else 0, spec2.stride(0) if k2 else 0, bp.nside, bp.lmax,
              _lib.ptr(bp.alpha), _lib.ptr(t.smant), _lib.ptr(t.sexp), _lib.ptr(bp.nact), jsplit, mpol,
              _lib.ptr(part32), _lib.ptr(part64), st)
    _lib.call("gm_bat_finish", _lib.ptr(part32), _lib.ptr(part64), k0, k2, bp.nside, bp.lmax, _lib.ptr(bp.nact),
              jsplit, mpol, _lib.ptr(t.P), _lib.ptr(alm0), alm0.stride(0) if k0 else 0, _lib.ptr(alm2),
              alm2.stride(1) if k2 else 0, int(mode), st)


def synthesis(bp, alm0, alm2, spec0, spec2):
    """Alms -> ring spectra (the inverse of ``analysis``'s layouts)."""
    k0 = 0 if alm0 is None else alm0.shape[0]
    k2 = 0 if alm2 is None else alm2.shape[0]
    _check(k0, k2)
    st = _lib.stream_ptr(bp.device)
    NS = int(_lib.query("gm_bat_stride", k0, k2, 1))
    coef = torch.empty((bp.tsize + COEF_TAIL) * NS, dtype=torch.float64, device=bp.device)
    t = bp.tab
    _lib.call("gm_bat_prep", _lib.ptr(alm0), k0, alm0.stride(0) if k0 else 0, _lib.ptr(alm2), k2,
              alm2.stride(1) if k2 else 0, _lib.ptr(t.P), bp.lmax, _lib.ptr(coef), st)
    _lib.call("gm_bat_syn", _lib.ptr(coef), k0, k2, bp.nside, bp.lmax, _lib.ptr(bp.alpha), _lib.ptr(t.smant),
              _lib.ptr(t.sexp), _lib.ptr(bp.nact), bp.jsplit, _lib.ptr(spec0), spec0.shape[2] if k0 else 0,
              spec0.stride(0) if k0 else 0, _lib.ptr(spec2), spec2.shape[2] if k2 else 0,
              spec2.stride(0) if k2 else 0, st)


def _fft_fwd(bp, maps0, maps2):
    nside, lmax = bp.nside, bp.lmax
    L, W, nring, npix = lmax + 1, 2 * lmax + 1, 4 * nside - 1, 12 * nside * nside
    st = _lib.stream_ptr(bp.device)
    spec0 = spec2 = None
    if maps0 is not None:
        spec0 = torch.empty((maps0.shape[0], nring, L), dtype=torch.complex128, device=bp.device)
        _lib.call("gm_ring_fft_fwd_s0", _lib.ptr(maps0), maps0.shape[0], npix, nside, lmax, _lib.ptr(spec0), L,
                  spec0.stride(0), st)
    if maps2 is not None:
        spec2 = torch.empty((maps2.shape[0], nring, W), dtype=torch.complex128, device=bp.device)
        for f in range(maps2.shape[0]):
            _lib.call("gm_ring_fft_fwd_s", _lib.ptr(maps2[f, 0]), _lib.ptr(maps2[f, 1]), nside, lmax,
                      _lib.ptr(spec2[f]), W, st)
    return spec0, spec2


def _fold(bp, syn0, syn2, g0, g2, out0, out2):
    """out = fold(syn) - g for every spectrum (the Jacobi residual in ring-Fourier space)."""
    nside, lmax = bp.nside, bp.lmax
    st = _lib.stream_ptr(bp.device)
    if syn0 is not None:
        _lib.call("gm_ring_fold_residual", _lib.ptr(syn0), _lib.ptr(g0), _lib.ptr(out0), syn0.shape[0], lmax + 1,
                  syn0.stride(0), nside, lmax, st)
    if syn2 is not None:
        for f in range(syn2.shape[0]):
            _lib.call("gm_ring_fold_residual_c", _lib.ptr(syn2[f]), _lib.ptr(g2[f]), _lib.ptr(out2[f]),
                      syn2.shape[2], nside, lmax, st)


def _like(t):
    return None if t is None else torch.empty_like(t)


def map2alm(maps0, maps2, lmax, n_iter=0, polar_skip=True, out=None):
    """``maps0`` ``(K0, npix)`` real spin-0 maps and ``maps2`` ``(K2, 2, npix)`` (Q, U) ->
    ``(alm0 (K0, nelem), alm2 (K2, 2, nelem))`` with ``n_iter`` Jacobi steps; either input
    may be None.  ``out = (alm0, alm2)`` gives the output tensors (complex128, contiguous,
    those shapes) instead of new ones."""
    dev = (maps0 if maps0 is not None else maps2).device
    maps0 = None if maps0 is None else maps0.to(device=dev, dtype=torch.float64).contiguous()
    maps2 = None if maps2 is None else maps2.to(device=dev, dtype=torch.float64).contiguous()
    npix = (maps0 if maps0 is not None else maps2).shape[-1]
    nside = sht._check_nside(npix)
    if maps0 is not None and (maps0.ndim != 2 or maps0.shape[1] != npix):
        raise ValueError("Input map has wrong shape")
    if maps2 is not None and (maps2.ndim != 3 or maps2.shape[1] != 2 or maps2.shape[2] != npix):
        raise ValueError("Input map has wrong shape")
    k0 = 0 if maps0 is None else maps0.shape[0]
    k2 = 0 if maps2 is None else maps2.shape[0]
    _check(k0, k2)
    lmax = int(lmax)
    bp = plan(nside, lmax, dev, polar_skip)
    L = lmax + 1
    nelem = L * (L + 1) // 2
    spec0, spec2 = _fft_fwd(bp, maps0, maps2)
    if out is not None:
        alm0, alm2 = out
        _check_out(alm0, (k0, nelem), torch.complex128, dev)
        _check_out(alm2, (k2, 2, nelem), torch.complex128, dev)
    else:
        alm0 = torch.empty((k0, nelem), dtype=torch.complex128, device=dev) if k0 else None
        alm2 = torch.empty((k2, 2, nelem), dtype=torch.complex128, device=dev) if k2 else None
    analysis(bp, spec0, spec2, alm0, alm2, 0, residual=FIRST_PASS_FP16 and n_iter > 0)
    if n_iter <= 0:
        return alm0, alm2
    syn0, syn2 = _like(spec0), _like(spec2)
    ra0, ra2 = _like(spec0), _like(spec2)
    synthesis(bp, alm0, alm2, syn0, syn2)
    _fold(bp, syn0, syn2, spec0, spec2, ra0, ra2)
    rb0, rb2 = spec0, spec2  # the input spectra are not read again
    d0, d2 = _like(alm0), _like(alm2)
    sign = 1
    for it in range(int(n_iter)):
        if it == n_iter - 1:  # the last correction goes straight into the alms
            analysis(bp, ra0, ra2, alm0, alm2, 1 if sign > 0 else 2, residual=True)
            break
        analysis(bp, ra0, ra2, d0, d2, 0, residual=True)
        for a, d in ((alm0, d0), (alm2, d2)):
            if a is not None:
                if sign > 0:
                    a.sub_(d)
                else:
                    a.add_(d)
        synthesis(bp, d0, d2, syn0, syn2)
        _fold(bp, syn0, syn2, ra0, ra2, rb0, rb2)
        ra0, rb0, ra2, rb2 = rb0, ra0, rb2, ra2
        sign = -sign
    return alm0, alm2


def _check_out(t, shape, dtype, dev):
    if shape[0] == 0:
        if t is not None:
            raise ValueError("output given for an absent kind")
        return
    if t is None or tuple(t.shape) != shape or t.dtype != dtype or t.device != dev or not t.is_contiguous():
        raise ValueError(f"output must be a contiguous {dtype} tensor of shape {shape} on {dev}")


def alm2map(alm0, alm2, nside, lmax, polar_skip=True, out=None):
    """Inverse of ``map2alm`` (no refinement): ``(maps0 (K0, npix), maps2 (K2, 2, npix))``;
    ``out = (maps0, maps2)`` gives the output tensors (float64, contiguous)."""
    dev = (alm0 if alm0 is not None else alm2).device
    k0 = 0 if alm0 is None else alm0.shape[0]
    k2 = 0 if alm2 is None else alm2.shape[0]
    _check(k0, k2)
    lmax = int(lmax)
    bp = plan(nside, lmax, dev, polar_skip)
    L, W, nring, npix = lmax + 1, 2 * lmax + 1, 4 * nside - 1, 12 * nside * nside
    spec0 = torch.empty((k0, nring, L), dtype=torch.complex128, device=dev) if k0 else None
    spec2 = torch.empty((k2, nring, W), dtype=torch.complex128, device=dev) if k2 else None
    synthesis(bp, None if alm0 is None else alm0.contiguous(), None if alm2 is None else alm2.contiguous(),
              spec0, spec2)
    st = _lib.stream_ptr(dev)
    maps0 = maps2 = None
    if out is not None:
        _check_out(out[0], (k0, npix), torch.float64, dev)
        _check_out(out[1], (k2, 2, npix), torch.float64, dev)
    if k0:
        maps0 = out[0] if out is not None else torch.empty((k0, npix), dtype=torch.float64, device=dev)
        _lib.call("gm_ring_fft_inv_s0", _lib.ptr(spec0), k0, L, spec0.stride(0), nside, lmax, _lib.ptr(maps0), npix, st)
    if k2:
        maps2 = out[1] if out is not None else torch.empty((k2, 2, npix), dtype=torch.float64, device=dev)
        for f in range(k2):
            _lib.call("gm_ring_fft_inv_s", _lib.ptr(spec2[f]), W, nside, lmax, _lib.ptr(maps2[f, 0]),
                      _lib.ptr(maps2[f, 1]), st)
    return maps0, maps2
