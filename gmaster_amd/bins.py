"""Bandpowers (``pymaster.NmtBin``; the reference's is ``gmaster/bins.py:39-302``).

A binning is a sparse band x multipole operator.  It is stored once, as a compressed
row table (``_Bands``): the band rows in order, each row's multipoles, the normalised
weight of each (row sums are one) and its ``f_l`` factor.  The binned spectrum is
``C_b = sum_{e in row b} w_e f_e C_{l_e}`` and unbinning puts ``C_b / f_e`` back on every
multipole of row ``b``.  The device kernel ``gm_bin_cell`` walks the same rows; the host
accessors slice them.

The public surface (constructors, accessors, the messages of the errors) is pymaster's;
the linear band rule is NaMaster's: bands of ``nlb`` multipoles starting at l = 2, with
l < 2 and an incomplete last band left out.
"""

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._array import device_tensor, wrap


@dataclass(frozen=True)
class _Bands:
    """Compressed rows of the band operator: row ``b`` is entries ``row[b]:row[b+1]``."""

    row: np.ndarray     # (nbands + 1,) int32 row offsets
    ell: np.ndarray     # (nnz,) int32 multipole of each entry
    weight: np.ndarray  # (nnz,) float64, sums to one over every row
    f_ell: np.ndarray   # (nnz,) float64, positive

    @property
    def nbands(self):
        return len(self.row) - 1

    def band_of_entry(self):
        return np.repeat(np.arange(self.nbands, dtype=np.int32), np.diff(self.row))

    def row_sums(self, values):
        """Sum of ``values`` (one per entry) over each row; 0 for an empty row."""
        out = np.zeros(self.nbands)
        full = np.flatnonzero(np.diff(self.row) > 0)
        if len(full):
            # each segment runs to the next non-empty row's start, i.e. to its own row's end
            out[full] = np.add.reduceat(np.asarray(values, dtype=np.float64), self.row[full])
        return out


def _linear_band_index(lmax, nlb):
    """Band index of every multipole 0..lmax for bands of ``nlb`` from l = 2 (-1: unbinned).
    A last band with fewer than ``nlb`` multipoles is dropped (NaMaster's rule, as in
    ``gmaster/bins.py:30-37``)."""
    ell = np.arange(lmax + 1)
    nfull = max((lmax - 1) // nlb, 0)  # complete bands within 2..lmax
    band = np.where((ell >= 2) & (ell < 2 + nfull * nlb), (ell - 2) // nlb, -1)
    return ell.astype(np.int32), band.astype(np.int32)


def _d_ell_factor(ells):
    ells = np.asarray(ells, dtype=np.float64)
    return ells * (ells + 1.0) / (2.0 * np.pi)


def _build_bands(ells, bpws, weights, f_ell, lmax):
    """Keep the entries with a band and l <= lmax, group them by band (stable within a
    band) and normalise each band's weights."""
    use = np.flatnonzero((bpws >= 0) & (ells <= lmax))
    order = use[np.argsort(bpws[use], kind="stable")]
    band = bpws[order]
    nbands = int(band[-1]) + 1 if len(band) else 1
    counts = np.zeros(nbands, dtype=np.int64)
    np.add.at(counts, band, 1)
    row = np.zeros(nbands + 1, dtype=np.int32)
    row[1:] = np.cumsum(counts)
    raw = _Bands(row, ells[order].astype(np.int32), weights[order].astype(np.float64),
                 f_ell[order].astype(np.float64))
    total = raw.row_sums(raw.weight)
    bad = np.flatnonzero(~(total > 0))
    if len(bad):
        raise RuntimeError(f"Weights in band {int(bad[0])} are wrong")
    norm = raw.weight / np.repeat(total, counts)
    fac = np.where(raw.f_ell > 0, raw.f_ell, 1.0)
    return _Bands(row, raw.ell, norm, fac)


class NmtBin:
    """Bandpower definition (keyword-only arguments, as in pymaster)."""

    def __init__(self, *, bpws, ells, lmax=None, weights=None, f_ell=None):
        ells = np.asarray(ells, dtype=np.int32)
        bpws = np.asarray(bpws, dtype=np.int32)
        if ells.ndim != 1 or bpws.shape != ells.shape or not len(ells):
            raise ValueError("ells and bpws must be non-empty 1-D arrays of equal size")
        n = len(ells)
        weights = np.ones(n) if weights is None else np.asarray(weights, dtype=np.float64)
        f_ell = np.ones(n) if f_ell is None else np.asarray(f_ell, dtype=np.float64)
        if weights.shape != ells.shape or f_ell.shape != ells.shape:
            raise ValueError("weights and f_ell must have the same shape as ells")
        self.lmax = int(ells.max()) if lmax is None else int(lmax)
        self._bands = _build_bands(ells, bpws, weights, f_ell, self.lmax)
        self.n_bands = self._bands.nbands
        self._views = {}

    # ------------------------------------------------------------------ constructors
    @classmethod
    def from_nside_linear(cls, nside, nlb, is_Dell=False, f_ell=None):
        return cls.from_lmax_linear(3 * nside - 1, nlb, is_Dell, f_ell)

    @classmethod
    def from_lmax_linear(cls, lmax, nlb, is_Dell=False, f_ell=None):
        ell, band = _linear_band_index(lmax, nlb)
        if f_ell is None and is_Dell:
            f_ell = _d_ell_factor(ell)
        return cls(bpws=band, ells=ell, lmax=lmax, f_ell=f_ell)

    @classmethod
    def from_edges(cls, ell_ini, ell_end, is_Dell=False, f_ell=None):
        """Top-hat bands ``[ell_ini[b], ell_end[b])``; ``f_ell`` holds one value per binned
        multipole, in band order."""
        lo = np.asarray(ell_ini, dtype=np.int64)
        hi = np.asarray(ell_end, dtype=np.int64)
        width = hi - lo
        band = np.repeat(np.arange(len(lo), dtype=np.int32), width)
        first = np.repeat(lo - np.concatenate([[0], np.cumsum(width)[:-1]]), width)
        ell = (first + np.arange(int(width.sum()))).astype(np.int32)
        if f_ell is None and is_Dell:
            f_ell = _d_ell_factor(ell)
        return cls(bpws=band, ells=ell, lmax=int(ell.max()), weights=np.repeat(1.0 / width, width), f_ell=f_ell)

    # ------------------------------------------------------------------ accessors
    def _entries(self, b):
        b = int(b)
        return slice(int(self._bands.row[b]), int(self._bands.row[b + 1]))

    def get_n_bands(self):
        return self.n_bands

    def get_nell_list(self):
        return np.diff(self._bands.row).astype(np.int64)

    def get_ell_list(self, b):
        return self._bands.ell[self._entries(b)].copy()

    def get_ell_min(self, b):
        return self.get_ell_list(b)[0]

    def get_ell_max(self, b):
        return self.get_ell_list(b)[-1]

    def get_weight_list(self, b):
        return self._bands.weight[self._entries(b)].copy()

    def get_fell_list(self, b):
        return self._bands.f_ell[self._entries(b)].copy()

    def get_effective_ells(self):
        return self._bands.row_sums(self._bands.ell * self._bands.weight)

    def per_ell_table(self):
        """Dense per-multipole view (l = 0..lmax): band index (-1: unbinned), weight and
        f_l (0 where unbinned), the layout of NaMaster's workspace files."""
        L = self.lmax + 1
        band = np.full(L, -1, dtype=np.int32)
        weight = np.zeros(L)
        fac = np.zeros(L)
        bt = self._bands
        band[bt.ell] = bt.band_of_entry()
        weight[bt.ell] = bt.weight
        fac[bt.ell] = bt.f_ell
        return band, weight, fac

    # ------------------------------------------------------------------ device views
    def device_view(self, device):
        """The row table on ``device``: ``ells``, ``start`` (row offsets, int32), ``band``
        (row of each entry, int64), ``w`` (weight x f_l) and ``inv_f`` (1 / f_l)."""
        key = str(device)
        view = self._views.get(key)
        if view is None:
            bt = self._bands
            view = {
                "ells": torch.as_tensor(bt.ell, dtype=torch.int32, device=device),
                "start": torch.as_tensor(bt.row, dtype=torch.int32, device=device),
                "band": torch.as_tensor(bt.band_of_entry(), dtype=torch.long, device=device),
                "w": torch.as_tensor(bt.weight * bt.f_ell, dtype=torch.float64, device=device),
                "inv_f": torch.as_tensor(1.0 / bt.f_ell, dtype=torch.float64, device=device),
            }
            self._views[key] = view
        return view

    def binning_matrix(self, device):
        """Dense ``(n_bands, lmax+1)`` operator ``B[b, l] = w_l f_l`` of ``bin_cell``."""
        v = self.device_view(device)
        out = torch.zeros((self.n_bands, self.lmax + 1), dtype=torch.float64, device=device)
        out.index_put_((v["band"], v["ells"].long()), v["w"], accumulate=True)
        return out

    def bin_cell(self, cls_in):
        """``(n_cls, lmax+1) -> (n_cls, n_bands)`` (1-D in, 1-D out)."""
        cls = device_tensor(cls_in)
        squeeze = cls.ndim == 1
        if squeeze:
            cls = cls[None]
        if cls.ndim != 2 or cls.shape[1] != self.lmax + 1:
            raise ValueError("Input Cl has wrong size")
        cls = cls.contiguous()
        v = self.device_view(cls.device)
        out = torch.empty((cls.shape[0], self.n_bands), dtype=torch.float64, device=cls.device)
        _lib.call("gm_bin_cell", _lib.ptr(cls), cls.shape[0], cls.shape[1], _lib.ptr(v["ells"]), _lib.ptr(v["start"]),
                  self.n_bands, _lib.ptr(v["w"]), _lib.ptr(out), _lib.stream_ptr(cls.device))
        return wrap(out[0] if squeeze else out)

    def unbin_cell(self, cls_in):
        """``(n_cls, n_bands) -> (n_cls, lmax+1)``: each band's value / f_l on its multipoles."""
        cls = device_tensor(cls_in)
        squeeze = cls.ndim == 1
        if squeeze:
            cls = cls[None]
        if cls.ndim != 2 or cls.shape[1] != self.n_bands:
            raise ValueError("Input Cl has wrong size")
        v = self.device_view(cls.device)
        out = torch.zeros((cls.shape[0], self.lmax + 1), dtype=torch.float64, device=cls.device)
        out[:, v["ells"].long()] = cls[:, v["band"]] * v["inv_f"]
        return wrap(out[0] if squeeze else out)
