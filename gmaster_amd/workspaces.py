"""MASTER workspaces (``pymaster.NmtWorkspace``) and coupled spectra.

Restates ``gmaster/workspaces.py``:

* ``compute_coupled_cell`` (``:86-125``) -- kernel ``gm_coupled_cell``;
* the mode-coupling blocks 00, 02 (TE), 22+ / 22- (``:352-517, 878-924``) -- one pass of
  the Wigner-3j kernel ``gm_coupling`` produces all four, so workspaces built from the
  same mask window share it (cached on the window's content);
* the Toeplitz approximation (``:520-590, :1125-1194, :1834-1840``) -- kernels
  ``gm_coupling_pairs`` (exact entries only) and ``gm_toeplitz`` (the fill);
* the slot / sign layout of the blocks (``:1851-1903``), binning with weights, f_ell and
  beams (``:1452-1470, :292-310``) -- kernel ``gm_bin_block``;
* ``decouple_cell`` (``:2106-2147``): binned spectra (``gm_bin_cell``) and an FP64 LU
  solve of the binned MCM: its LU factors (``gm_lu_factor``: blocked, without row
  exchanges, with a flag that falls back to ``torch.linalg.lu_factor`` when partial
  pivoting would exchange rows) are computed once per binned matrix and reused by every
  ``decouple_cell`` (``gm_lu_solve``) and by the bandpower windows (getrs).
"""

import hashlib

import numpy as np
import torch

from . import _lib
from ._array import device_tensor, wrap
from ._config import nmt_params

_COUPLING_CACHE = {}
_G_CACHE = {}


def _g_table(n, device):
    """``g(p) = Gamma(p + 1/2) / (sqrt(pi) Gamma(p + 1))`` for p < n, long-double recursion."""
    key = (n, str(device))
    t = _G_CACHE.get(key)
    if t is None:
        p = np.arange(1, n, dtype=np.longdouble)
        g = np.concatenate([[np.longdouble(1)], np.cumprod((2 * p - 1) / (2 * p))])
        t = torch.as_tensor(g.astype(np.float64), device=device)
        _G_CACHE[key] = t
    return t


def toeplitz_sanity(l_toeplitz, l_exact, dl_band, lmax, fields=()):
    """Parameter checks of the Toeplitz approximation (``gmaster/workspaces.py:1125-1138``)."""
    if l_toeplitz <= 0:
        return
    if any(getattr(f, "pure_e", False) or getattr(f, "pure_b", False) for f in fields):
        raise ValueError("Can't use Toeplitz approximation with purification.")
    if l_exact <= 0 or dl_band < 0:
        raise ValueError("`l_exact` and `dl_band` must be positive numbers")
    if l_exact > l_toeplitz:
        raise ValueError("`l_exact` must be `<= l_toeplitz")
    if l_toeplitz >= lmax or l_exact >= lmax or dl_band >= lmax:
        raise ValueError("`l_toeplitz`, `l_exact` and `dl_band` must be smaller than `l_max`")


def _runs(starts, counts, device):
    """Concatenated integer runs ``starts[i] + (0 .. counts[i] - 1)``."""
    counts = counts.clamp(min=0)
    total = int(counts.sum())
    if total == 0:
        return torch.empty(0, dtype=torch.int64, device=device), torch.empty(0, dtype=torch.int64, device=device)
    idx = torch.repeat_interleave(torch.arange(len(counts), device=device), counts)
    first = torch.cumsum(counts, 0) - counts
    return idx, torch.arange(total, device=device) - first[idx] + starts[idx]


def _toeplitz_pairs(L, l_toeplitz, l_exact, dl_band, device):
    """``(lo, hi)`` int32: the upper-triangle pairs whose exact values the Toeplitz fill reads
    (diagonal, column l_toeplitz, columns <= l_exact, dl_band diagonals;
    ``gmaster/workspaces.py:520-539``), each once, built on the device as disjoint runs:
    rows lo <= l_exact (every hi >= lo), band rows lo > l_exact (hi <= lo + dl_band), and
    what column l_toeplitz adds outside both."""
    lt, le, dl = int(l_toeplitz), int(l_exact), int(dl_band)
    ar = lambda a, b: torch.arange(a, max(a, b), dtype=torch.int64, device=device)
    lo_a = ar(0, min(le, L - 1) + 1)
    ia, hi_a = _runs(lo_a, L - lo_a, device)
    lo_b = ar(le + 1, L)
    ib, hi_b = _runs(lo_b, torch.clamp(L - lo_b, max=dl + 1), device)
    lo_c1 = ar(le + 1, lt - dl)
    hi_c2 = ar(lt + dl + 1, L) if lt > le else ar(0, 0)
    lo = torch.cat([lo_a[ia], lo_b[ib], lo_c1, torch.full_like(hi_c2, lt)])
    hi = torch.cat([hi_a, hi_b, torch.full_like(lo_c1, lt), hi_c2])
    return lo.to(torch.int32), hi.to(torch.int32)


_ITEMS_CACHE = {}


def _toeplitz_items(L, l_toeplitz, l_exact, dl_band, device):
    """``(items, lo, hi)`` covering exactly the pairs of ``_toeplitz_pairs``: ``items``
    ``(n, 3)`` int32 for ``gm_coupling_rows`` -- row segments (rows lo <= l_exact, the band
    rows, row l_toeplitz past its band) cut into 64-column waves -- and int32 pair lists for
    ``gm_coupling_pairs``: the entries of column l_toeplitz above the band and, when a band
    row's length leaves at most 8 entries past its last whole wave, those entries (one
    diagonal each, consecutive rows sharing a wave there instead of a wave holding 1-8
    lanes).  Built on the host once per geometry and kept on the device (the 00, 02 and 22
    workspaces of a mask share them, and a rebuild per call would be a host round trip
    while the device waits)."""
    lt, le, dl = int(l_toeplitz), int(l_exact), int(dl_band)
    key = (int(L), lt, le, dl, str(device))
    hit = _ITEMS_CACHE.get(key)
    if hit is not None:
        return hit
    ra = np.arange(0, min(le, L - 1) + 1)
    rb = np.arange(le + 1, L)
    full = 64 * ((dl + 1) // 64)
    split = full > 0 and (dl + 1) - full <= 8
    bw = full if split else dl + 1
    rows, starts, counts = [ra, rb], [ra, rb], [L - ra, np.minimum(L - rb, bw)]
    if lt > le and lt + dl + 1 < L:
        rows.append(np.array([lt]))
        starts.append(np.array([lt + dl + 1]))
        counts.append(np.array([L - (lt + dl + 1)]))
    row, start, count = np.concatenate(rows), np.concatenate(starts), np.concatenate(counts)
    keep = count > 0
    row, start, count = row[keep], start[keep], count[keep]
    nch = (count + 63) // 64
    seg = np.repeat(np.arange(len(nch)), nch)
    k = np.arange(len(seg)) - np.repeat(np.cumsum(nch) - nch, nch)  # chunk of its segment
    items = np.stack([row[seg], start[seg] + 64 * k, np.minimum(count[seg] - 64 * k, 64)], 1).astype(np.int32)
    c1 = np.arange(le + 1, max(le + 1, lt - dl))  # (lo, lt) above the band
    plo, phi = [c1], [np.full_like(c1, lt)]
    if split:
        for t in range(full, dl + 1):  # band entries past the whole waves, one diagonal each
            r = rb[rb + t < L]
            plo.append(r)
            phi.append(r + t)
    lo, hi = np.concatenate(plo), np.concatenate(phi)
    out = (torch.as_tensor(np.ascontiguousarray(items), device=device),
           torch.as_tensor(lo.astype(np.int32), device=device), torch.as_tensor(hi.astype(np.int32), device=device))
    if len(_ITEMS_CACHE) >= 8:
        _ITEMS_CACHE.clear()
    _ITEMS_CACHE[key] = out
    return out


def coupling_blocks(window, lmax, toeplitz=None):
    """``{"00", "02", "22p", "22m"}`` coupling blocks ``(lmax+1, lmax+1)`` for a mask window.

    ``window`` is the mask pseudo-C_ell (entries above 2 lmax cannot couple).  With
    ``toeplitz = (l_toeplitz, l_exact, dl_band)`` only the entries the Toeplitz fill reads
    are evaluated (``gm_coupling_pairs``) and every block is then filled (``gm_toeplitz``),
    as ``gmaster/workspaces.py:1834-1840`` does to the TE / EE blocks and ``:542-590`` to the
    TT block.  Results are cached on (lmax, Toeplitz parameters, window bytes) so the 00, 02
    and 22 workspaces of one mask share a pass.
    """
    w = device_tensor(window)[: 2 * lmax + 1].contiguous()
    tkey = tuple(int(v) for v in toeplitz) if toeplitz is not None and toeplitz[0] > 0 else None
    key = (int(lmax), tkey, str(w.device), hashlib.sha1(w.cpu().numpy().tobytes()).hexdigest())
    hit = _COUPLING_CACHE.get(key)
    if hit is not None:
        return hit
    L = lmax + 1
    dev = w.device
    names = ("00", "02", "22p", "22m")
    g = _g_table(L + 1, dev)
    wp = (w * (2 * torch.arange(len(w), dtype=torch.float64, device=dev) + 1) / (4 * np.pi)).contiguous()
    stream = _lib.stream_ptr(dev)
    if tkey is None:
        blocks = {k: torch.empty((L, L), dtype=torch.float64, device=dev) for k in names}
        _lib.call("gm_coupling", _lib.ptr(wp), len(w) - 1, _lib.ptr(g), lmax, *(_lib.ptr(blocks[k]) for k in names),
                  stream)
    else:
        lt, le, dl = tkey
        items, plo, phi = _toeplitz_items(L, lt, le, dl, dev)
        exact = {k: torch.zeros((L, L), dtype=torch.float64, device=dev) for k in names}
        _lib.call("gm_coupling_rows", _lib.ptr(wp), len(w) - 1, _lib.ptr(g), lmax, _lib.ptr(items),
                  items.shape[0], *(_lib.ptr(exact[k]) for k in names), stream)
        _lib.call("gm_coupling_pairs", _lib.ptr(wp), len(w) - 1, _lib.ptr(g), lmax, _lib.ptr(plo), _lib.ptr(phi),
                  plo.numel(), *(_lib.ptr(exact[k]) for k in names), stream)
        blocks = {}
        for k in names:
            blocks[k] = torch.empty((L, L), dtype=torch.float64, device=dev)
            _lib.call("gm_toeplitz", _lib.ptr(exact[k]), lmax, lt, le, dl, _lib.ptr(blocks[k]), stream)
        del exact
    if len(_COUPLING_CACHE) >= 2:
        _COUPLING_CACHE.clear()
    _COUPLING_CACHE[key] = blocks
    return blocks


def fast_coupling_blocks(window, lmax, toeplitz=None, tol=None, n_check=4096, seed=0):
    """The ``"fast"`` form: Toeplitz approximation with a wide exact region (only the exact
    entries it reads are evaluated), checked against exact values of ``n_check`` random non-exact
    entries; falls back to the exact coupling when their largest error in correlation units
    exceeds ``tol``.  Returns ``(blocks, info)``."""
    from .learned import spot_check

    lt, le, dl = toeplitz or nmt_params.fast_toeplitz
    tol = nmt_params.fast_tol if tol is None else tol
    if lmax < 2 * lt:
        return coupling_blocks(window, lmax), {"form": "fast", "fallback": True, "reason": "lmax < 2 l_toeplitz"}
    toe = coupling_blocks(window, lmax, (lt, le, dl))
    err = spot_check(window, lmax, (lt, le, dl), toe, n_check, seed)
    info = {"form": "fast", "toeplitz": (lt, le, dl), "sample_err": err, "n_check": n_check, "tol": tol,
            "fallback": bool(err > tol)}
    if err > tol:
        return coupling_blocks(window, lmax), info
    return toe, info


def clear_coupling_cache():
    _COUPLING_CACHE.clear()
    from .learned import clear_learned_cache

    clear_learned_cache()


_COUPLING_PRECISION = ["auto"]


def set_coupling_precision(name):
    """Operand precision of the coupling builders (``gmaster/workspaces.py:56-78``): "fp64",
    "fp32" or "auto".  Accepted for compatibility; every coupling kernel here evaluates in
    FP64 (the 3j recursion and the quadrature GEMMs), i.e. at or above the requested width."""
    if name not in ("fp64", "fp32", "auto"):
        raise KeyError("GMaster coupling precision must be 'fp64', 'fp32' or 'auto'")
    _COUPLING_PRECISION[0] = name


def coupling_precision():
    """The current coupling-precision setting ("fp64", "fp32" or "auto")."""
    return _COUPLING_PRECISION[0]


def _general_blocks(mask_cls, lmax, s1, s2, n1, n2, split=True):
    """``(even, odd)`` (or their sum when ``split`` is False) ``(lmax+1)^2`` general coupling
    blocks from the mask spectrum ``mask_cls`` (multipoles L < len(mask_cls); those above
    2 lmax cannot couple), by Gauss-Legendre quadrature (``gmaster/workspaces.py:682-732,
    878-905``): ``M = (2l' + 1)/(8 pi) sum_k w_k c(x_k) d^l_{n1 n2}(x_k) d^l'_{-s1 -s2}(x_k)``."""
    from . import purecoupling as pc
    for a, b in ((n1, n2), (-s1, -s2), (s1 - n1, s2 - n2)):
        if max(abs(a), abs(b)) > 16:
            raise NotImplementedError("spin indices above 16 are not supported")
    w = mask_cls[: 2 * lmax + 1].contiguous()
    lw = w.numel() - 1
    L = lmax + 1
    dev = w.device
    # the integrand is a polynomial of degree <= 2 lmax + lw in cos(beta)
    x, wq = pc.gauss_legendre((2 * lmax + lw) // 2 + 1, dev)
    ellw = torch.arange(lw + 1, dtype=torch.float64, device=dev)
    ell = torch.arange(L, dtype=torch.float64, device=dev)
    c, sg = pc.wigner_d_corr(x, s1 - n1, s2 - n2, lw, ((2 * ellw + 1) * w).contiguous(), signed=split)
    left = pc.wigner_d_table(x, n1, n2, lmax)
    right = pc.wigner_d_table(x, -s1, -s2, lmax, ((2 * ell + 1) / (8 * np.pi)).contiguous())
    total = torch.matmul(pc.table_combine([(1, left, c, None)], wq), right.T)
    if not split:
        return total
    signed = torch.matmul(pc.table_combine([(1, left, sg, None)], wq), right.T)
    return pc.parity_split(total, signed)


def general_spin_blocks(mask_cls, lmax, spin1, spin2, is_teb=False, toeplitz=None):
    """``{"02", "22p", "22m"}`` blocks of a spin pair outside {0, 2} x {0, 2}
    (``gmaster/workspaces.py:1801-1827``): the 0-s block sums both parities of
    (s1, s2, n1, n2) = (spin1, spin2, spin1, spin2); an s x s pair splits them; a joint TEB
    workspace takes its s x s blocks from (spin2, spin2, spin2, spin2).  With
    ``toeplitz = (l_toeplitz, l_exact, dl_band)`` each block is then filled by ``gm_toeplitz``."""
    w = device_tensor(mask_cls)
    L = lmax + 1
    out = {}
    if spin1 == 0 or spin2 == 0:
        out["02"] = _general_blocks(w, lmax, spin1, spin2, spin1, spin2, split=False)
    else:
        out["22p"], out["22m"] = _general_blocks(w, lmax, spin1, spin2, spin1, spin2)
    if is_teb:
        out["22p"], out["22m"] = _general_blocks(w, lmax, spin2, spin2, spin2, spin2)
    if toeplitz is not None and toeplitz[0] > 0:
        lt, le, dl = (int(v) for v in toeplitz)
        for k in list(out):
            filled = torch.empty((L, L), dtype=torch.float64, device=w.device)
            _lib.call("gm_toeplitz", _lib.ptr(out[k].contiguous()), lmax, lt, le, dl, _lib.ptr(filled),
                      _lib.stream_ptr(w.device))
            out[k] = filled
    return out


def get_general_coupling_matrix(pcl_mask, s1, s2, n1, n2, parity="all"):
    """General mode-coupling matrix, as in ``pymaster`` (``gmaster/workspaces.py:1069-1122``):

        M[l, l'] = sum_L (2l' + 1)(2L + 1)/(4 pi) C_L P(l + l' + L)
                   (l l' L; n1 -s1 s1-n1)(l l' L; n2 -s2 s2-n2)

    over ``0 <= l, l', L <= nl - 1`` (nl = len(pcl_mask)), by Gauss-Legendre quadrature on the
    device: the product of two 3j symbols is half the integral of d^l_{n1 n2} d^l'_{-s1 -s2}
    d^L_{s1-n1, s2-n2}, so M = (2l' + 1)/(8 pi) sum_k w_k c(x_k) d^l_{n1 n2}(x_k) d^l'_{-s1 -s2}(x_k)
    with the correlation c = sum_L (2L + 1) C_L d^L (``gm_wigner_d_corr``; its (-1)^L form
    gives the parity split), one FP64 GEMM per parity part.  Rows and columns below
    max(s1, s2, n1, n2) are zero.  ``parity``: "all", "even", "odd" or "both" (even, odd)."""
    if parity not in ("all", "even", "odd", "both"):
        raise ValueError('`parity` must be "all", "even", "odd", or "both".')
    w = device_tensor(pcl_mask)
    if w.ndim != 1:
        raise ValueError("pcl_mask must be one-dimensional")
    s1, s2, n1, n2 = (int(v) for v in (s1, s2, n1, n2))
    lmax = w.numel() - 1
    if parity == "all":
        out = _general_blocks(w, lmax, s1, s2, n1, n2, split=False)
    else:
        even, odd = _general_blocks(w, lmax, s1, s2, n1, n2)
        out = {"even": even, "odd": odd}.get(parity)
        if parity == "both":
            out = torch.stack([even, odd])
    lstart = max(s1, s2, n1, n2)
    if lstart > 0:
        out[..., :lstart, :] = 0
        out[..., :, :lstart] = 0
    return wrap(out)


def compute_coupled_cell(f1, f2):
    """``(n1 n2, lmax + 1)`` coupled pseudo-spectra of two fields (rows ``a * n2 + b``)."""
    if not f1.is_compatible(f2, strict=False):
        raise ValueError("You're trying to correlate incompatible fields")
    return wrap(_coupled_cell(f1.get_alms(), f2.get_alms(), min(f1.ainfo.lmax, f2.ainfo.lmax)))


def _coupled_cell(alm1, alm2, lmax):
    a1 = device_tensor(alm1, dtype=torch.complex128)
    a2 = device_tensor(alm2, dtype=torch.complex128, device=a1.device)
    if a1.ndim == 1:
        a1 = a1[None]
    if a2.ndim == 1:
        a2 = a2[None]
    n1, n2 = a1.shape[0], a2.shape[0]
    if a1.shape[1] != a2.shape[1]:
        raise ValueError("alms of different sizes")
    L = lmax + 1
    cl = torch.empty((n1 * n2, L), dtype=torch.float64, device=a1.device)
    work = torch.empty(int(_lib.query("gm_cl_work_size", lmax, n1 * n2)), dtype=torch.float64, device=a1.device)
    _lib.call("gm_coupled_cell", _lib.ptr(a1), n1, _lib.ptr(a2), n2, a1.shape[1], lmax, _lib.ptr(cl), _lib.ptr(work),
              _lib.stream_ptr(a1.device))
    return cl


_TE_LEVELS = ("02", "02pure")
_EVEN_LEVELS = ("22p", "22p1", "22p2")
_ODD_LEVELS = ("22m", "22m1", "22m2")


def _layout(ncls, spin1, spin2, purity=(False, False, False, False)):
    """``[(block name, (row channel, column channel), sign)]`` (``gmaster/workspaces.py:1851-1903``).
    ncls 7 is the joint TEB layout: TT, then TE/TB with sign (-1)^spin2, then the spin x spin
    slots at offset 3 with the M+ / M- signs of spin_sign = 1.  With purification
    ``(pure_e1, pure_b1, pure_e2, pure_b2)`` a slot takes the block of its level, the number
    of purified fields among the slot's two channels (``:1845-1850, 1865-1901``; ncls 7 uses
    the spin-2 field's flags for both channels)."""
    pe1, pb1, pe2, pb2 = (int(bool(v)) for v in purity)
    if ncls == 1:
        return [("00", (0, 0), 1.0)]
    if ncls == 2:
        sign = float((-1) ** (spin1 + spin2))
        return [(_TE_LEVELS[pe1 + pe2], (0, 0), sign), (_TE_LEVELS[pb1 + pb2], (1, 1), sign)]
    if ncls == 7:
        out = [("00", (0, 0), 1.0)]
        out += [(_TE_LEVELS[lv], (1 + c, 1 + c), float((-1) ** spin2)) for c, lv in enumerate((pe2, pb2))]
        pe1, pb1 = pe2, pb2
        offset, sign = 3, 1.0
    else:
        out, offset, sign = [], 0, float((-1) ** (spin1 + spin2))
    levels = (pe1 + pe2, pe1 + pb2, pb1 + pe2, pb1 + pb2)
    out += [(_EVEN_LEVELS[lv], (offset + c, offset + c), sign) for c, lv in enumerate(levels)]
    out += [(_ODD_LEVELS[lv], (offset + c, offset + 3 - c), -sign if c in (1, 2) else sign)
            for c, lv in enumerate(levels)]
    return out



SOLVE_RHS_MAX = 64  # right-hand sides solved by gm_lu_solve; more go through getrs


def lu_factor(M):
    """LU factors of a square FP64 device matrix for ``lu_solve``: ``gm_lu_factor`` (no row
    exchanges, the binned coupling matrices being near the identity) unless its flag says
    that partial pivoting would exchange rows, then ``torch.linalg.lu_factor``."""
    n = M.shape[0]
    lu = M.contiguous().clone()
    flag = torch.zeros(1, dtype=torch.int32, device=M.device)
    _lib.call("gm_lu_factor", _lib.ptr(lu), n, _lib.ptr(flag), _lib.stream_ptr(M.device))
    if int(flag.item()) != 0:
        return ("pivoted", torch.linalg.lu_factor(M))
    return ("plain", lu, lu.T.contiguous())


def lu_solve(fac, rhs):
    """``M^-1 rhs`` for ``fac = lu_factor(M)`` and ``rhs`` ``(n, k)``."""
    kind, f = fac[0], fac[1]
    if kind == "pivoted":
        return torch.linalg.lu_solve(*f, rhs)
    n = f.shape[0]
    if rhs.shape[1] > SOLVE_RHS_MAX:
        piv = torch.arange(1, n + 1, dtype=torch.int32, device=f.device)
        return torch.linalg.lu_solve(f, piv, rhs)
    out = rhs.to(torch.float64).contiguous().clone()
    _lib.call("gm_lu_solve", _lib.ptr(f), _lib.ptr(fac[2]), n, _lib.ptr(out), out.shape[1], _lib.stream_ptr(f.device))
    return out


class NmtWorkspace:
    """Mode-coupling matrix, its binned form and the bandpower operators of two fields."""

    def __init__(self, fl1=None, fl2=None, bins=None, is_teb=False, l_toeplitz=-1, l_exact=-1, dl_band=-1,
                 fname=None, normalization="MASTER"):
        self.blocks = None
        self.mcm_binned = None
        self._bpws = None
        if fname is not None:
            self.read_from(fname)
            return
        if fl1 is not None or fl2 is not None or bins is not None:
            self.compute_coupling_matrix(fl1, fl2, bins, is_teb=is_teb, l_toeplitz=l_toeplitz, l_exact=l_exact,
                                         dl_band=dl_band, normalization=normalization)

    @classmethod
    def from_fields(cls, fl1, fl2, bins, **kwargs):
        return cls(fl1, fl2, bins, **kwargs)

    @classmethod
    def from_file(cls, fname):
        """Workspace read from a NaMaster-compatible FITS file."""
        return cls(fname=fname)

    @property
    def mcm(self):
        """Dense coupling matrix, ``(ncls (lmax+1))^2`` in ``l * ncls + c`` ordering."""
        return self.get_coupling_matrix()

    # ------------------------------------------------------------------ FITS I/O
    def write_to(self, fname):
        """Write the workspace as a NaMaster-compatible FITS file (``gmaster/workspaces.py:2211-2255``):
        primary image WSP_PRIMARY (dense MCM + keywords), tables BEAMS, PCL_MASKS, BANDPOWERS."""
        from . import fits

        tl, te, db = self.toeplitz if getattr(self, "toeplitz", None) else (-1, -1, -1)
        header = {"EXTNAME": "WSP_PRIMARY", "LMAX": self.lmax, "LMAX_MASK": self.lmax_mask,
                  "IS_TEB": bool(getattr(self, "is_teb", False)),
                  "NCLS": self.ncls, "NORM_TYPE": self.norm_type, "WAWB": float(self.wawb), "SPIN1": self.spin1,
                  "SPIN2": self.spin2, "ANISO1": False, "ANISO2": False, "PURE_E1": bool(self.purity[0]),
                  "PURE_B1": bool(self.purity[1]), "PURE_E2": bool(self.purity[2]), "PURE_B2": bool(self.purity[3]),
                  "L_TOEPLITZ": tl, "L_EXACT": te, "DL_BAND": db}
        L = self.lmax + 1
        ell = np.arange(L, dtype=np.int32)
        b = self.bins
        bands, weights, f_ell = b.per_ell_table()
        pcl = self.pcl_mask.detach().cpu().numpy().astype(np.float64)
        fits.write(fname, self.get_coupling_matrix().cpu().numpy(), header, [
            ("BEAMS", [("L", ell), ("BEAM1", self.beam1.cpu().numpy().astype(np.float64)),
                       ("BEAM2", self.beam2.cpu().numpy().astype(np.float64))], {}),
            ("PCL_MASKS", [("L", np.arange(len(pcl), dtype=np.int32)), ("PCL_MASKS", pcl)], {}),
            ("BANDPOWERS", [("BAND", bands), ("ELLS", ell), ("WEIGHTS", weights), ("F_ELL", f_ell)],
             {"ELL_MAX": self.lmax, "N_BANDS": b.get_n_bands()}),
        ])

    def read_from(self, fname):
        """Read a NaMaster-compatible FITS workspace (``gmaster/workspaces.py:2162-2209``)."""
        from . import fits
        from .bins import NmtBin

        hdus = fits.read(fname)
        h, dense = fits.find(hdus, "WSP_PRIMARY")
        self.is_teb = bool(h.get("IS_TEB", False))
        if any(bool(h.get(k, False)) for k in ("ANISO1", "ANISO2")):
            raise NotImplementedError("anisotropic workspaces are not supported by this engine")
        self.purity = tuple(bool(h.get(k, False)) for k in ("PURE_E1", "PURE_B1", "PURE_E2", "PURE_B2"))
        self.lmax = int(h["LMAX"])
        self.lmax_mask = int(h["LMAX_MASK"])
        self.ncls = int(h["NCLS"])
        self.norm_type = int(h.get("NORM_TYPE", 0))
        self.normalization = "MASTER" if self.norm_type == 0 else "FKP"
        self.wawb = float(h.get("WAWB", 0.0))
        self.spin1 = int(h.get("SPIN1", 0 if self.ncls in (1, 7) else 2))
        self.spin2 = int(h.get("SPIN2", 0 if self.ncls == 1 else 2))
        self.nmaps1 = 1 if self.spin1 == 0 else 2
        self.nmaps2 = 1 if self.spin2 == 0 else 2
        lt = int(h.get("L_TOEPLITZ", -1))
        self.toeplitz = (lt, int(h.get("L_EXACT", -1)), int(h.get("DL_BAND", -1))) if lt > 0 else None
        self.coupling_info = {"form": "file"}
        L = self.lmax + 1
        dev = torch.device("cuda", torch.cuda.current_device())
        if dense is None or dense.shape != (L * self.ncls, L * self.ncls):
            raise ValueError("WSP_PRIMARY does not hold a coupling matrix of the stated size")
        self.layout = _layout(self.ncls, self.spin1, self.spin2, self.purity)
        # the distinct blocks, read back from their first slot in the dense matrix
        d4 = torch.as_tensor(dense, dtype=torch.float64, device=dev).reshape(L, self.ncls, L, self.ncls)
        self.blocks = {}
        for name, (c1, c2), sign in self.layout:
            if name not in self.blocks:
                self.blocks[name] = (sign * d4[:, c1, :, c2]).contiguous()
        for name in ("00", "02", "22p", "22m"):
            self.blocks.setdefault(name, torch.zeros((L, L), dtype=torch.float64, device=dev))
        # a matrix that the slot layout does not reproduce (one set by update_coupling_matrix)
        # is kept as one block per channel pair
        again = torch.zeros_like(d4)
        for name, (c1, c2), sign in self.layout:
            again[:, c1, :, c2] += sign * self.blocks[name]
        if not torch.equal(again, d4):
            self.blocks = {"00": d4[:, 0, :, 0].contiguous()}
            self.layout = []
            for c1 in range(self.ncls):
                for c2 in range(self.ncls):
                    self.blocks[f"c{c1}_{c2}"] = d4[:, c1, :, c2].contiguous()
                    self.layout.append((f"c{c1}_{c2}", (c1, c2), 1.0))
        del again
        _, beams = fits.find(hdus, "BEAMS")
        if "BEAMS" in beams:
            common = np.sqrt(beams["BEAMS"].astype(np.float64))
            b1 = b2 = common
        else:
            b1, b2 = beams["BEAM1"].astype(np.float64), beams["BEAM2"].astype(np.float64)
        self.beam1 = torch.as_tensor(b1, device=dev)
        self.beam2 = torch.as_tensor(b2, device=dev)
        _, pcl = fits.find(hdus, "PCL_MASKS")
        self.pcl_mask = torch.as_tensor(pcl["PCL_MASKS"].astype(np.float64), device=dev)
        name = "BINS" if any(str(hh.get("EXTNAME", "")).strip() == "BINS" for hh, _ in hdus) else "BANDPOWERS"
        bh, bd = fits.find(hdus, name)
        keep = bd["BAND"] >= 0
        self.bins = NmtBin(bpws=bd["BAND"][keep].astype(np.int64), ells=bd["ELLS"][keep].astype(np.int64),
                           lmax=int(bh["ELL_MAX"]), weights=bd["WEIGHTS"][keep].astype(np.float64),
                           f_ell=bd["F_ELL"][keep].astype(np.float64))
        self.nbands = self.bins.get_n_bands()
        self._postprocess()

    def compute_coupling_matrix(self, fl1, fl2, bins, is_teb=False, l_toeplitz=-1, l_exact=-1, dl_band=-1,
                                normalization="MASTER"):
        if not fl1.is_compatible(fl2, strict=False):
            raise ValueError("Fields have incompatible pixelizations")
        if fl1.ainfo.lmax != bins.lmax:
            raise ValueError(f"Maximum multipoles in bins ({bins.lmax}) and fields ({fl1.ainfo.lmax}) are not the same")
        if is_teb and (fl1.spin != 0 or fl2.spin == 0):
            raise ValueError("If is_teb=True, fl1 must be spin-0 and fl2 non-zero spin")
        toeplitz_sanity(l_toeplitz, l_exact, dl_band, fl1.ainfo.lmax, (fl1, fl2))
        if normalization not in ("MASTER", "FKP"):
            raise ValueError(f"Unknown normalization type {normalization}. Allowed options are 'MASTER' and 'FKP'")
        self.purity = tuple(bool(v) for v in (fl1.pure_e, fl1.pure_b, fl2.pure_e, fl2.pure_b))
        self.spin1, self.spin2 = fl1.spin, fl2.spin
        self.nmaps1 = 1 if self.spin1 == 0 else 2
        self.nmaps2 = 1 if self.spin2 == 0 else 2
        self.is_teb = bool(is_teb)
        self.ncls = 7 if self.is_teb else self.nmaps1 * self.nmaps2
        self.lmax = fl1.ainfo.lmax
        self.lmax_mask = fl1.ainfo_mask.lmax
        self.beam1 = fl1.beam
        self.beam2 = fl2.beam
        self.bins = bins
        self.nbands = bins.get_n_bands()
        self.normalization = normalization
        self.norm_type = int(normalization == "FKP")
        if any(self.purity):
            # the pure blocks amplify mask-spectrum rounding (weights up to ~L^4): FP64 mask alms
            am1 = fl1.mask_alms_at("fp64")
            am2 = am1 if fl2 is fl1 else fl2.mask_alms_at("fp64")
        else:
            am1 = fl1.get_mask_alms()
            am2 = am1 if fl2 is fl1 else fl2.get_mask_alms()
        self.pcl_mask = _coupled_cell(am1, am2, self.lmax_mask)[0]
        if fl2 is fl1:
            self.pcl_mask = self.pcl_mask - fl1.Nw
        self.wawb = 0.0
        if self.norm_type:
            npix = fl1.minfo.npix
            self.wawb = float(torch.sum(fl1.mask * fl2.mask)) * (4 * np.pi / npix) / (4 * np.pi)
        self.toeplitz = (l_toeplitz, l_exact, dl_band) if l_toeplitz > 0 else None
        self.coupling_info = {"form": "toeplitz" if self.toeplitz else nmt_params.form}
        if self.spin1 not in (0, 2) or self.spin2 not in (0, 2):
            # general spins: the TT block from the 3j kernel (joint TEB), the others by
            # quadrature over the mask spectrum up to lmax_mask
            if self.is_teb:
                self.blocks = dict(coupling_blocks(self.pcl_mask, self.lmax, self.toeplitz))
            else:
                self.blocks = {"00": torch.zeros(1, dtype=torch.float64, device=self.pcl_mask.device)}
            self.blocks.update(general_spin_blocks(self.pcl_mask, self.lmax, self.spin1, self.spin2, self.is_teb,
                                                   self.toeplitz))
            self.coupling_info = {"form": "general-spin quadrature" + (" + toeplitz" if self.toeplitz else "")}
        elif any(self.purity):
            # exact standard blocks (Toeplitz is rejected with purification) and the
            # quadrature blocks of the purified slots
            from .purecoupling import pure_coupling_blocks
            self.blocks = dict(coupling_blocks(self.pcl_mask, self.lmax))
            self.blocks.update(pure_coupling_blocks(self.pcl_mask, self.lmax))
            self.coupling_info = {"form": "exact+pure"}
        elif self.toeplitz is None and nmt_params.form == "fast":
            self.blocks, self.coupling_info = fast_coupling_blocks(self.pcl_mask, self.lmax)
        elif self.toeplitz is None and nmt_params.form == "learned":
            from .learned import learned_coupling_blocks
            self.blocks, self.coupling_info = learned_coupling_blocks(self.pcl_mask, self.lmax)
        else:
            self.blocks = coupling_blocks(self.pcl_mask, self.lmax, self.toeplitz)
        self.layout = _layout(self.ncls, self.spin1, self.spin2, self.purity)
        self._postprocess()

    def _postprocess(self):
        dev = self.blocks["00"].device
        d = self.bins.device_view(dev)
        beam = (self.beam1 * self.beam2).to(dev)
        tin = (beam[d["ells"].long()] * d["inv_f"]).contiguous()
        nb, ncls = self.nbands, self.ncls
        st = _lib.stream_ptr(dev)
        binned = {}
        for name in {b for b, _, _ in self.layout}:
            out = torch.empty((nb, nb), dtype=torch.float64, device=dev)
            _lib.call("gm_bin_block", _lib.ptr(self.blocks[name]), self.lmax, _lib.ptr(d["ells"]), _lib.ptr(d["start"]),
                      nb, _lib.ptr(d["w"]), _lib.ptr(tin), _lib.ptr(out), st)
            binned[name] = out
        mb = torch.zeros((nb, ncls, nb, ncls), dtype=torch.float64, device=dev)
        for name, (c1, c2), sign in self.layout:
            mb[:, c1, :, c2] += sign * binned[name]
        mb = mb.reshape(nb * ncls, nb * ncls)
        if self.norm_type:
            mb = self.wawb * torch.eye(nb * ncls, dtype=torch.float64, device=dev)
        self.mcm_binned = mb
        self._bpws = None
        self._lu = None

    # ------------------------------------------------------------------ products
    def decouple_cell(self, cl_in, cl_bias=None, cl_noise=None):
        """``(ncls, >= lmax+1)`` coupled spectra -> ``(ncls, nbands)`` bandpowers."""
        cl = device_tensor(cl_in, device=self.mcm_binned.device)
        if cl.ndim != 2 or cl.shape[0] != self.ncls or cl.shape[1] < self.lmax + 1:
            raise ValueError(f"Input power spectrum has wrong shape. Expected ({self.ncls}, {self.lmax + 1}), "
                             f"but got {tuple(cl.shape)}")
        total = cl[:, : self.lmax + 1]
        for name, bias in (("bias", cl_bias), ("noise", cl_noise)):
            if bias is not None:
                b = device_tensor(bias, device=cl.device)
                if b.ndim != 2 or b.shape[0] != self.ncls or b.shape[1] < self.lmax + 1:
                    raise ValueError(f"Input {name} power spectrum has wrong shape")
                total = total - b[:, : self.lmax + 1]
        binned = self.bins.bin_cell(total.contiguous()).as_subclass(torch.Tensor)
        rhs = binned.T.reshape(-1)
        sol = self._solve(rhs[:, None])[:, 0]
        return wrap(sol.reshape(self.nbands, self.ncls).T.contiguous())

    def _solve(self, rhs):
        """``mcm_binned^-1 rhs`` from LU factors computed once per binned matrix."""
        if getattr(self, "_lu", None) is None:
            self._lu = lu_factor(self.mcm_binned)
        return lu_solve(self._lu, rhs)

    def couple_cell(self, cl_in):
        """Theory ``(ncls, >= lmax+1)`` -> expected coupled spectra ``(ncls, lmax+1)``."""
        cl = device_tensor(cl_in, device=self.mcm_binned.device)
        if cl.ndim != 2 or cl.shape[0] != self.ncls or cl.shape[1] < self.lmax + 1:
            raise ValueError(f"Input power spectrum has wrong shape. Expected ({self.ncls}, {self.lmax + 1}), "
                             f"but got {tuple(cl.shape)}")
        th = cl[:, : self.lmax + 1] * (self.beam1 * self.beam2)[None, :]
        out = torch.zeros_like(th)
        for name, (c1, c2), sign in self.layout:
            out[c1] += sign * (self.blocks[name] @ th[c2])
        return wrap(out)

    def get_coupling_matrix(self):
        """Dense ``(ncls (lmax+1))^2`` MCM in ``l * ncls + c`` ordering."""
        L, ncls = self.lmax + 1, self.ncls
        out = torch.zeros((L, ncls, L, ncls), dtype=torch.float64, device=self.mcm_binned.device)
        for name, (c1, c2), sign in self.layout:
            out[:, c1, :, c2] += sign * self.blocks[name]
        return wrap(out.reshape(L * ncls, L * ncls))

    def get_bandpower_windows(self):
        """``(ncls, nbands, ncls, lmax+1)`` windows ``solve(mcm_binned, W M diag(beam))``."""
        if self._bpws is None:
            dev = self.mcm_binned.device
            L, nb, ncls = self.lmax + 1, self.nbands, self.ncls
            wbin = self.bins.binning_matrix(dev)
            beam = self.beam1 * self.beam2
            one = torch.zeros((nb, ncls, L, ncls), dtype=torch.float64, device=dev)
            for name, (c1, c2), sign in self.layout:
                one[:, c1, :, c2] += sign * (wbin @ self.blocks[name]) * beam[None, :]
            sol = self._solve(one.reshape(nb * ncls, L * ncls))
            self._bpws = sol.reshape(nb, ncls, L, ncls).permute(1, 0, 3, 2).contiguous()
        return wrap(self._bpws)

    def update_coupling_matrix(self, new_matrix):
        """Replace the unbinned MCM (``(ncls (lmax+1))^2``, ``l * ncls + c`` ordering) and
        recompute the binned operators (``gmaster/workspaces.py:2023-2040``)."""
        size = self.ncls * (self.lmax + 1)
        expected = (size, size)
        if tuple(np.shape(new_matrix)) != expected:
            raise ValueError(f"Input matrix has an inconsistent shape. Expected {expected}, "
                             f"but got {tuple(np.shape(new_matrix))}")
        L, ncls = self.lmax + 1, self.ncls
        dense = device_tensor(new_matrix, device=self.mcm_binned.device).reshape(L, ncls, L, ncls)
        # every channel pair becomes its own block with sign 1
        self.blocks = {"00": dense[:, 0, :, 0].contiguous()}
        self.layout = []
        for c1 in range(ncls):
            for c2 in range(ncls):
                name = f"c{c1}_{c2}"
                self.blocks[name] = dense[:, c1, :, c2].contiguous()
                self.layout.append((name, (c1, c2), 1.0))
        self._postprocess()

    def update_beams(self, beam1, beam2):
        L = self.lmax + 1
        b1 = device_tensor(beam1, device=self.mcm_binned.device)
        b2 = device_tensor(beam2, device=self.mcm_binned.device)
        if b1.shape != (L,) or b2.shape != (L,):
            raise ValueError(f"The new beams must go up to ell = {self.lmax}")
        self.beam1, self.beam2 = b1, b2
        self._postprocess()

    def update_bins(self, bins):
        if bins.lmax != self.lmax:
            raise ValueError("The new binning scheme has a different maximum multipole")
        self.bins = bins
        self.nbands = bins.get_n_bands()
        self._postprocess()


def _coefficient_blocks(spin1, spin2, is_teb):
    """Names of the distinct blocks a spin pair couples through: the spin-0 pair ("00"), the
    spin-0 x spin-s pair ("0s") and the spin-s pair's M+ / M- ("pp", "mm").  A joint TEB
    workspace has all four."""
    zero1, zero2 = spin1 == 0, spin2 == 0
    names = set()
    if is_teb or (zero1 and zero2):
        names.add("00")
    if is_teb or zero1 != zero2:
        names.add("0s")
    if is_teb or not (zero1 or zero2):
        names.update(("pp", "mm"))
    return names


def _check_coefficient_request(spin1, spin2, is_teb, pure_any):
    rules = (
        (is_teb and (spin1 != 0 or spin2 == 0), "is_teb is only valid for spin1=0 and spin2!=0"),
        (pure_any and 2 not in (spin1, spin2), "pure_any is only valid for spin 2"),
        (pure_any and not {spin1, spin2} <= {0, 2}, "Purification is only implemented for spin 2"),
    )
    for broken, message in rules:
        if broken:
            raise ValueError(message)


def _mask_coefficients(window, lmax, spin1, spin2, is_teb, pure_any, toe):
    """Blocks of one mask window (already cut or zero-padded to 2 lmax + 1 entries), each
    divided by 2 l' + 1; "0s", "pp", "mm" carry the purification-level axis."""
    blocks = dict(coupling_blocks(window, lmax, toe))
    if not {spin1, spin2} <= {0, 2}:
        # other spins: the 0-s block pairs the non-zero spin with 0, the s x s blocks use
        # (spin1, spin2), or (spin2, spin2) for a joint TEB pair (gmaster/workspaces.py:1279-1316)
        names = _coefficient_blocks(spin1, spin2, is_teb)
        if "0s" in names:
            blocks.update(general_spin_blocks(window, lmax, *((0, spin2) if spin1 == 0 else (spin1, 0)),
                                              toeplitz=toe))
        if "pp" in names:
            gen = general_spin_blocks(window, lmax, spin2 if is_teb else spin1, spin2, toeplitz=toe)
            blocks["22p"], blocks["22m"] = gen["22p"], gen["22m"]
    inv_cols = 1.0 / (2 * torch.arange(lmax + 1, dtype=torch.float64, device=window.device) + 1)
    levels = {"0s": [blocks["02"]], "pp": [blocks["22p"]], "mm": [blocks["22m"]]}
    if pure_any:
        # purification levels 0, 1, 2 (gmaster/workspaces.py:1317-1325)
        from .purecoupling import pure_coupling_blocks

        pure = pure_coupling_blocks(window, lmax)
        levels["0s"].append(pure["02pure"])
        levels["pp"] += [pure["22p1"], pure["22p2"]]
        levels["mm"] += [pure["22m1"], pure["22m2"]]
    res = {"00": blocks["00"] * inv_cols}
    for name, mats in levels.items():
        res[name] = torch.stack(mats) * inv_cols
    return res


def get_master_coefficients(pcl_mask, lmax, spin1, spin2, is_teb=False, pure_any=False, l_toeplitz=-1, l_exact=-1,
                            dl_band=-1):
    """Distinct MASTER coupling blocks of a spin pair, divided by ``2 l' + 1``
    (``gmaster/workspaces.py:1197-1350``): ``"00"``, ``"0s"``, ``"pp"`` (M+) and ``"mm"`` (M-),
    None where the spin pair has no such block; ``"0s"``, ``"pp"`` and ``"mm"`` carry a
    leading axis of the purification levels (1, or 2 / 3 with ``pure_any``), and a mask axis
    before the last two when ``pcl_mask`` is 2D.  Spin pairs outside {0, 2} take their 0-s and
    s x s blocks from the general quadrature (``general_spin_blocks``)."""
    spin1, spin2, lmax = int(spin1), int(spin2), int(lmax)
    _check_coefficient_request(spin1, spin2, is_teb, pure_any)
    toeplitz_sanity(l_toeplitz, l_exact, dl_band, lmax)
    windows = device_tensor(pcl_mask)
    if windows.ndim not in (1, 2):
        raise ValueError("pcl_mask must be one- or two-dimensional")
    stacked = windows.ndim == 2
    windows = windows.reshape(-1, windows.shape[-1])
    nw = windows.shape[1]
    # every window cut or zero-padded to the 2 lmax + 1 multipoles the 3j sums reach
    cut = torch.zeros((windows.shape[0], 2 * lmax + 1), dtype=torch.float64, device=windows.device)
    keep = min(nw, 2 * lmax + 1)
    cut[:, :keep] = windows[:, :keep]
    toe = (l_toeplitz, l_exact, dl_band) if l_toeplitz > 0 else None
    per_window = [_mask_coefficients(w, lmax, spin1, spin2, is_teb, pure_any, toe) for w in cut]
    present = _coefficient_blocks(spin1, spin2, is_teb)
    out = {}
    for name in ("00", "0s", "pp", "mm"):
        if name not in present:
            out[name] = None
        elif stacked:
            out[name] = wrap(torch.stack([c[name] for c in per_window], dim=-3))
        else:
            out[name] = wrap(per_window[0][name])
    out["pure_any"] = bool(pure_any)
    out["toeplitz"] = {"l_toeplitz": l_toeplitz, "l_exact": l_exact, "dl_band": dl_band}
    out["spins"] = (spin1, spin2)
    out["lmax"] = lmax
    out["lmax_mask"] = nw - 1
    return out
