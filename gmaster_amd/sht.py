"""HEALPix spherical-harmonic transforms of any spin over the gfx950 kernels.

``map2alm`` (with Jacobi refinement) and ``alm2map`` for spin 0 (one or two maps that
share one latitudinal pass, e.g. a field and its mask) and spin s (Q, U -> E, B).

Pipeline of one analysis (``gmaster/_sht/healpix.py:1741-1926`` is the reference router):

1. ring FFT of every ring into the phased spectrum ``G[r][m]`` (kernel ``gm_ring_fft_fwd_*``);
2. Legendre / Wigner-d analysis per order m over all rings (``gm_lat_ana_*``), then a
   finish kernel that sums the ring chunks, applies the per-(l, m) scale and (spin s)
   combines +-m into E/B;
3. each Jacobi iteration ``alm <- alm - A(S alm - map)`` (``gmaster/_sht/healpix.py:1459-1471``)
   runs synthesis (``gm_lat_prep_*`` + ``gm_lat_syn_*``), the ring-Fourier fold residual
   (``gm_ring_fold_residual*``, no FFT) and another analysis whose finish subtracts.

All device buffers are torch tensors; every kernel runs on the current torch stream.
"""

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib

_PLANS = {}


def _seed_constants(L, s):
    """Per order m: ``(-1)^{m+s} sqrt((2M+1)/4pi) sqrt(C(2M, M + min(m, s)))`` as
    (mantissa, exponent) pairs, M = max(m, s).  Log-binomials are summed in long double
    (``C(2m, m+s) = C(2m-2, m-1+s) (2m)(2m-1) / ((m+s)(m-s))`` for m > s)."""
    ld = np.longdouble
    m = np.arange(L)
    log2c = np.zeros(L, dtype=ld)
    if s < L:
        # m >= s branch by recursion from C(2s, 2s) = 1 at m = s
        mm = np.arange(s + 1, L).astype(ld)
        steps = np.log2(2 * mm) + np.log2(2 * mm - 1) - np.log2(mm + s) - np.log2(mm - s)
        log2c[s + 1:] = np.cumsum(steps)
    for mi in range(min(s, L)):  # m < s: C(2s, s + m), exact integers
        log2c[mi] = ld(math.log2(math.comb(2 * s, s + mi)))
    M = np.maximum(m, s).astype(ld)
    total = 0.5 * np.log2((2 * M + 1) / (4 * ld(np.pi))) + 0.5 * log2c
    e = np.floor(total)
    mant = np.exp2(total - e)
    sign = np.where((m + s) % 2 == 0, 1.0, -1.0)
    return (sign * mant.astype(np.float64)), e.astype(np.int32)


def _ring_slots(nside, spin):
    """(z, sin) of the rings in the order the Legendre kernels visit them."""
    i = np.arange(1, 4 * nside)
    north = np.where(i > 2 * nside, 4 * nside - i, i)
    cap = north < nside
    q = north.astype(np.float64)
    zn = np.where(cap, 1.0 - q * q / (3.0 * nside * nside), 2.0 * (2 * nside - q) / (3.0 * nside))
    one_minus = np.where(cap, q * q / (3.0 * nside * nside), (2.0 * q - nside) / (3.0 * nside))
    sin = np.sqrt(one_minus * (1.0 + zn))
    z = np.where(i > 2 * nside, -zn, zn)
    if spin == 0:
        r = 2 * nside - 1 - np.arange(2 * nside)  # folded: equator first, then towards the pole
    else:
        j = np.arange(4 * nside - 1)
        u = (j + 1) // 2
        r = np.where(j == 0, 2 * nside - 1, np.where(j % 2 == 1, 2 * nside - 1 - u, 2 * nside - 1 + u))
    return z[r], sin[r]


def _nactive(nside, lmax, spin, polar_skip):
    """Prefix length of the ring slots that order m visits (ducc0's mlim rule)."""
    z, sin = _ring_slots(nside, spin)
    n = len(z)
    L = lmax + 1
    if not polar_skip:
        return np.full(L, n, dtype=np.int32)
    ofs = max(100.0, 0.01 * lmax)
    b = lmax * sin + ofs
    mlim = spin * np.abs(z) + np.sqrt(np.maximum(b * b - spin * spin * sin * sin, 0.0))
    # running max from the pole end so the active set is a prefix
    best = np.maximum.accumulate(mlim[::-1])[::-1]
    m = np.arange(L)
    # number of slots j with best[j] >= m  (best is non-increasing)
    return np.searchsorted(-best, -m, side="right").astype(np.int32)


@dataclass
class Plan:
    nside: int
    lmax: int
    spin: int
    device: torch.device
    alpha: torch.Tensor
    gamma: torch.Tensor
    P: torch.Tensor
    smant: torch.Tensor
    sexp: torch.Tensor
    nact: torch.Tensor
    nchunks: int
    tsize: int
    nchunks_mixed: int = 0


def plan(nside, lmax, spin, device=None, polar_skip=True):
    """Cached device tables for one (nside, lmax, spin) geometry."""
    device = torch.device(device if device is not None else "cuda")
    if device.index is None:
        device = torch.device("cuda", torch.cuda.current_device())
    key = (int(nside), int(lmax), int(spin), device, bool(polar_skip))
    p = _PLANS.get(key)
    if p is not None:
        return p
    nside, lmax, spin = int(nside), int(lmax), int(spin)
    L = lmax + 1
    tsize = int(_lib.query("gm_table_size", lmax))
    f64 = dict(dtype=torch.float64, device=device)
    alpha = torch.empty(tsize, **f64)
    gamma = torch.empty(tsize, **f64) if spin else None
    P = torch.empty(tsize, **f64)
    _lib.call("gm_lat_tables", lmax, spin, _lib.ptr(alpha), _lib.ptr(gamma), _lib.ptr(P), _lib.stream_ptr(device))
    sm, se = _seed_constants(L, spin)
    p = Plan(
        nside=nside, lmax=lmax, spin=spin, device=device, alpha=alpha, gamma=gamma, P=P,
        smant=torch.as_tensor(sm, **f64),
        sexp=torch.as_tensor(se, dtype=torch.int32, device=device),
        nact=torch.as_tensor(_nactive(nside, lmax, spin, polar_skip), dtype=torch.int32, device=device),
        nchunks=int(_lib.query("gm_lat_nchunks", nside, spin)),
        tsize=tsize,
        nchunks_mixed=int(_lib.query("gm_lat_nchunks_mixed", nside, spin, 1)),
    )
    _PLANS[key] = p
    return p


def clear_plans():
    _PLANS.clear()


def _check_nside(npix):
    nside = int(round(math.sqrt(npix / 12)))
    if 12 * nside * nside != npix or nside & (nside - 1):
        raise ValueError("Input map has wrong shape")
    return nside


def _make_room(nside):
    """Above Nside 4096 one transform's temporaries reach ~100 GB (three ring spectra and
    the partial sums of a spin-2 field): hand cached but unused blocks back first, so the
    allocator is not left fragmented by the previous transform's differently sized ones."""
    if nside > 4096:
        torch.cuda.empty_cache()


def _as_f64(t, device):
    t = torch.as_tensor(t)
    return t.to(device=device, dtype=torch.float64).contiguous()


# ------------------------------------------------------------------------------ spin 0
def _passes(p, spin, mixed, K, nslots, mode):
    """``(chunk0, finish mode)`` of each analysis pass: one pass up to Nside 4096; above, the
    ring chunks are covered ``nslots`` at a time so the partial buffer stays bounded, the
    later passes adding to (or, for the Jacobi update, subtracting from) the alms."""
    total = int(_lib.query("gm_lat_ring_chunks", p.nside, spin, int(bool(mixed)), K))
    return [(c0, mode if c0 == 0 else (1 if mode == 1 else 2)) for c0 in range(0, total, nslots)]


def _analysis_s0(p, spec, alm, mode, mixed=False):
    K = spec.shape[0]
    st = _lib.stream_ptr(p.device)
    nch = int(_lib.query("gm_lat_nchunks_mixed", p.nside, 0, K)) if mixed else p.nchunks
    sfx = "_mixed" if mixed else ""
    partial = torch.empty(nch * p.tsize * 2 * K, dtype=torch.float64, device=p.device)
    ld = spec.shape[2]
    for c0, md in _passes(p, 0, mixed, K, nch, int(mode)):
        _lib.call("gm_lat_ana_s0" + sfx, _lib.ptr(spec), K, ld, spec.stride(0), p.nside, p.lmax, _lib.ptr(p.alpha),
                  _lib.ptr(p.smant), _lib.ptr(p.sexp), _lib.ptr(p.nact), _lib.ptr(partial), nch, c0, st)
        _lib.call("gm_lat_finish_s0" + sfx, _lib.ptr(partial), K, p.nside, _lib.ptr(p.nact), _lib.ptr(p.P), p.lmax,
                  _lib.ptr(alm), alm.stride(0), md, c0, st)


def _synthesis_s0(p, alm, spec, mixed=False):
    K = alm.shape[0]
    st = _lib.stream_ptr(p.device)
    coef = torch.empty(p.tsize * K, dtype=torch.complex128, device=p.device)
    _lib.call("gm_lat_prep_s0", _lib.ptr(alm), K, alm.stride(0), _lib.ptr(p.P), p.lmax, _lib.ptr(coef), st)
    # one map has only two FP32 FMAs per recurrence step to save, fewer than the step's
    # FP64 -> FP32 conversion costs: a single map synthesises in FP64 in every form
    # (measured at Nside 1024: 1.6 ms FP64 against 1.9 ms FP32 vector, 2.4 ms matrix unit)
    name, alpha = ("gm_lat_syn_s0_mixed" if mixed and K > 1 else "gm_lat_syn_s0"), p.alpha
    _lib.call(name, _lib.ptr(coef), K, p.nside, p.lmax, _lib.ptr(alpha), _lib.ptr(p.smant),
              _lib.ptr(p.sexp), _lib.ptr(p.nact), _lib.ptr(spec), spec.shape[2], spec.stride(0), st)


def _refine(alm, n_iter, analysis, synthesis, fold, spec, mixed):
    """Jacobi refinement ``alm <- alm - A(S alm - map)`` in ring-Fourier space
    (``gmaster/_sht/healpix.py:1459-1472``).

    FP64 (``mixed`` False): every pass recomputes the full residual from ``alm``.
    Mixed (FP32 contraction, FP64 recurrence): after R_0 = fold(S alm_0) - G the passes act
    on corrections only, D = A R, alm -= D, R <- R - fold(S D), so the rounding of each
    pass scales with the correction, not with the signal.  The residual buffer holds s R
    (s = +-1, flipped by each fold)."""
    if n_iter <= 0:
        return alm
    synth = torch.empty_like(spec)
    if not mixed:
        resid = torch.empty_like(spec)
        for _ in range(int(n_iter)):
            synthesis(alm, synth, False)
            fold(synth, spec, resid)
            analysis(resid, alm, 1, False)
        return alm
    ra = torch.empty_like(spec)
    synthesis(alm, synth, True)
    fold(synth, spec, ra)
    rb = spec  # the input spectrum is not read again: its buffer holds the next residual
    sign = 1
    d = torch.empty_like(alm)
    for it in range(int(n_iter)):
        analysis(ra, d, 0, True)
        if sign > 0:
            alm.sub_(d)
        else:
            alm.add_(d)
        if it < n_iter - 1:
            synthesis(d, synth, True)
            fold(synth, ra, rb)
            ra, rb = rb, ra
            sign = -sign
    return alm


def _groups(n, mixed, max_group=2):
    """Split n maps into recurrence-sharing groups of at most ``max_group`` (2 by default;
    4 or 8 in mixed precision: measured slower per map at Nside 1024, register and LDS
    pressure lowering occupancy more than the shared recurrence saves)."""
    sizes = tuple(k for k in ((8, 4, 2, 1) if mixed else (2, 1)) if k <= max_group)
    out, start = [], 0
    while start < n:
        k = next(s for s in sizes if s <= n - start)
        out.append((start, k))
        start += k
    return out


def map2alm_s0(maps, lmax, n_iter=0, polar_skip=True, precision="fp64", max_group=2):
    """``(K, npix)`` real maps -> ``(K, nelem)`` alms.  Maps share one recurrence in groups
    of up to ``max_group`` (2 in FP64; 2, 4 or 8 with ``precision="mixed"``, which runs every
    pass with FP32 contraction over the FP64 recurrence)."""
    maps = torch.as_tensor(maps)
    dev = maps.device if maps.is_cuda else torch.device("cuda")
    maps = _as_f64(maps, dev)
    if maps.ndim != 2 or maps.shape[0] < 1:
        raise ValueError("Input map has wrong shape")
    groups = _groups(maps.shape[0], _mixed(precision), max_group)
    if len(groups) > 1:
        return torch.cat([map2alm_s0(maps[a:a + k], lmax, n_iter, polar_skip, precision, max_group)
                          for a, k in groups])
    K, npix = maps.shape
    nside = _check_nside(npix)
    lmax = int(lmax)
    _make_room(nside)
    p = plan(nside, lmax, 0, dev, polar_skip)
    L = lmax + 1
    nring = 4 * nside - 1
    st = _lib.stream_ptr(dev)
    spec = torch.empty((K, nring, L), dtype=torch.complex128, device=dev)
    _lib.call("gm_ring_fft_fwd_s0", _lib.ptr(maps), K, npix, nside, lmax, _lib.ptr(spec), L, spec.stride(0), st)
    nelem = L * (L + 1) // 2
    alm = torch.empty((K, nelem), dtype=torch.complex128, device=dev)
    _analysis_s0(p, spec, alm, 0, _mixed(precision))

    def fold(f, g, out):
        _lib.call("gm_ring_fold_residual", _lib.ptr(f), _lib.ptr(g), _lib.ptr(out), K, L, spec.stride(0), nside, lmax,
                  st)

    return _refine(alm, n_iter, lambda s_, a, mode, mx: _analysis_s0(p, s_, a, mode, mx),
                   lambda a, s_, mx: _synthesis_s0(p, a, s_, mx), fold, spec, _mixed(precision))


def _mixed(precision):
    if precision not in ("fp64", "mixed"):
        raise ValueError("precision must be 'fp64' or 'mixed'")
    return precision == "mixed"


def alm2map_s0(alm, nside, lmax, polar_skip=True, precision="fp64", max_group=2):
    alm = torch.as_tensor(alm)
    dev = alm.device if alm.is_cuda else torch.device("cuda")
    alm = alm.to(device=dev, dtype=torch.complex128).contiguous()
    lmax = int(lmax)
    L = lmax + 1
    if alm.ndim != 2 or alm.shape[0] < 1 or alm.shape[1] != L * (L + 1) // 2:
        raise ValueError("Input alm has wrong shape")
    groups = _groups(alm.shape[0], _mixed(precision), max_group)
    if len(groups) > 1:
        return torch.cat([alm2map_s0(alm[a:a + k], nside, lmax, polar_skip, precision, max_group)
                          for a, k in groups])
    K = alm.shape[0]
    p = plan(nside, lmax, 0, dev, polar_skip)
    nring = 4 * nside - 1
    spec = torch.empty((K, nring, L), dtype=torch.complex128, device=dev)
    _synthesis_s0(p, alm, spec, _mixed(precision))
    npix = 12 * nside * nside
    maps = torch.empty((K, npix), dtype=torch.float64, device=dev)
    _lib.call("gm_ring_fft_inv_s0", _lib.ptr(spec), K, L, spec.stride(0), nside, lmax, _lib.ptr(maps), npix,
              _lib.stream_ptr(dev))
    return maps


# ------------------------------------------------------------------------------ spin s
def _analysis_s(p, spec, ae, ab, mode, mixed=False):
    st = _lib.stream_ptr(p.device)
    nch = p.nchunks_mixed if mixed else p.nchunks
    sfx = "_mixed" if mixed else ""
    partial = torch.empty(nch * p.tsize * 4, dtype=torch.float64, device=p.device)
    for c0, md in _passes(p, p.spin, mixed, 1, nch, int(mode)):
        _lib.call("gm_lat_ana_s" + sfx, _lib.ptr(spec), spec.shape[1], p.nside, p.lmax, p.spin, _lib.ptr(p.alpha),
                  _lib.ptr(p.gamma), _lib.ptr(p.smant), _lib.ptr(p.sexp), _lib.ptr(p.nact), _lib.ptr(partial), nch,
                  c0, st)
        _lib.call("gm_lat_finish_s" + sfx, _lib.ptr(partial), p.nside, _lib.ptr(p.nact), _lib.ptr(p.P), p.lmax,
                  p.spin, _lib.ptr(ae), _lib.ptr(ab), md, c0, st)


def _synthesis_s(p, ae, ab, spec, mixed=False):
    st = _lib.stream_ptr(p.device)
    coef = torch.empty(p.tsize * 2, dtype=torch.complex128, device=p.device)
    _lib.call("gm_lat_prep_s", _lib.ptr(ae), _lib.ptr(ab), _lib.ptr(p.P), p.lmax, p.spin, _lib.ptr(coef), st)
    name, (alpha, gamma) = ("gm_lat_syn_s_mixed" if mixed else "gm_lat_syn_s"), (p.alpha, p.gamma)
    _lib.call(name, _lib.ptr(coef), p.nside, p.lmax, p.spin, _lib.ptr(alpha), _lib.ptr(gamma),
              _lib.ptr(p.smant), _lib.ptr(p.sexp), _lib.ptr(p.nact), _lib.ptr(spec), spec.shape[1], st)


def map2alm_spin(maps, lmax, spin, n_iter=0, polar_skip=True, precision="fp64"):
    """``(2, npix)`` (Q, U) -> ``(2, nelem)`` (E, B).  ``precision="mixed"`` as for spin 0."""
    maps = torch.as_tensor(maps)
    dev = maps.device if maps.is_cuda else torch.device("cuda")
    maps = _as_f64(maps, dev)
    if maps.ndim != 2 or maps.shape[0] != 2:
        raise ValueError("Input map has wrong shape")
    npix = maps.shape[1]
    nside = _check_nside(npix)
    lmax, spin = int(lmax), int(spin)
    _make_room(nside)
    p = plan(nside, lmax, spin, dev, polar_skip)
    L = lmax + 1
    W = 2 * L - 1
    nring = 4 * nside - 1
    st = _lib.stream_ptr(dev)
    spec = torch.empty((nring, W), dtype=torch.complex128, device=dev)
    _lib.call("gm_ring_fft_fwd_s", _lib.ptr(maps[0]), _lib.ptr(maps[1]), nside, lmax, _lib.ptr(spec), W, st)
    nelem = L * (L + 1) // 2
    alm = torch.empty((2, nelem), dtype=torch.complex128, device=dev)
    _analysis_s(p, spec, alm[0], alm[1], 0, _mixed(precision))

    def fold(f, g, out):
        _lib.call("gm_ring_fold_residual_c", _lib.ptr(f), _lib.ptr(g), _lib.ptr(out), W, nside, lmax, st)

    return _refine(alm, n_iter, lambda s_, a, mode, mx: _analysis_s(p, s_, a[0], a[1], mode, mx),
                   lambda a, s_, mx: _synthesis_s(p, a[0], a[1], s_, mx), fold, spec, _mixed(precision))


def alm2map_spin(alm, nside, lmax, spin, polar_skip=True, precision="fp64"):
    alm = torch.as_tensor(alm)
    dev = alm.device if alm.is_cuda else torch.device("cuda")
    alm = alm.to(device=dev, dtype=torch.complex128).contiguous()
    lmax, spin = int(lmax), int(spin)
    L = lmax + 1
    if alm.ndim != 2 or alm.shape != (2, L * (L + 1) // 2):
        raise ValueError("Input alm has wrong shape")
    p = plan(nside, lmax, spin, dev, polar_skip)
    W = 2 * L - 1
    nring = 4 * nside - 1
    spec = torch.empty((nring, W), dtype=torch.complex128, device=dev)
    _synthesis_s(p, alm[0], alm[1], spec, _mixed(precision))
    npix = 12 * nside * nside
    maps = torch.empty((2, npix), dtype=torch.float64, device=dev)
    _lib.call("gm_ring_fft_inv_s", _lib.ptr(spec), W, nside, lmax, _lib.ptr(maps[0]), _lib.ptr(maps[1]),
              _lib.stream_ptr(dev))
    return maps


def map2alm(maps, spin, lmax, n_iter=0, polar_skip=True, precision="fp64"):
    if int(spin) == 0:
        return map2alm_s0(maps, lmax, n_iter, polar_skip, precision)
    return map2alm_spin(maps, lmax, spin, n_iter, polar_skip, precision)


def alm2map(alm, spin, nside, lmax, polar_skip=True, precision="fp64"):
    if int(spin) == 0:
        return alm2map_s0(alm, nside, lmax, polar_skip, precision)
    return alm2map_spin(alm, nside, lmax, spin, polar_skip, precision)
